// FFT-based 2-D convolution with a reverse-mode backward pass.
//
// Forward ("same" output, kernel origin at its centre c = ((KH-1)/2, (KW-1)/2)):
//   y_b = crop_c( c2r( K^ . r2c(pad_0 x_b) ) ) / (P Q),   K^ = r2c(pad_0 k)
// on a P x Q grid, P >= H + KH - 1 and Q >= W + KW - 1 (13-smooth, Q even).
// Backward, written with the adjoint of every step (pad* = crop, crop* = pad,
// (K^ .)* = conj(K^) ., r2c* / c2r* = the R2C_ADJOINT / C2R_ADJOINT plans):
//   G_b  = c2r*( pad_c g_b ) / (P Q)
//   gx_b = crop_0( r2c*( conj(K^) . G_b ) )
//   gk   = crop_0( r2c*( sum_b conj(X^_b) . G_b ) )
// The spectra X^_b and K^ are recomputed in the backward pass instead of being
// stored by the forward pass.  Inputs and outputs are float32 or float64; all
// arithmetic is FP64.
#include <algorithm>
#include <cstring>

#include "engine.h"
#include "fft_common.h"

namespace ifft {

// dst[b, i, j] (P x Q, f64) = src[b, i - oy, j - ox] for the in-range part, else 0
template <typename T>
__global__ void k_pad2d(const T *__restrict__ src, double *__restrict__ dst, int B, int H, int W, int P, int Q, int oy,
                        int ox)
{
    const long long total = (long long)B * P * Q;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e % Q);
        const long long r = e / Q;
        const int i = (int)(r % P);
        const int b = (int)(r / P);
        const int si = i - oy, sj = j - ox;
        double v = 0.0;
        if (si >= 0 && si < H && sj >= 0 && sj < W) v = (double)src[((long long)b * H + si) * W + sj];
        dst[e] = v;
    }
}

// dst[b, i, j] = scale * src[b, i + oy, j + ox], i < H, j < W
template <typename T>
__global__ void k_crop2d(const double *__restrict__ src, T *__restrict__ dst, int B, int H, int W, int P, int Q, int oy,
                         int ox, double scale)
{
    const long long total = (long long)B * H * W;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e % W);
        const long long r = e / W;
        const int i = (int)(r % H);
        const int b = (int)(r / H);
        dst[e] = (T)(scale * src[((long long)b * P + i + oy) * Q + j + ox]);
    }
}

// z[b, m] = a[b, m] * (conj ? conj(k[m]) : k[m]), m < n
__global__ void k_spec_mul(const double2 *__restrict__ a, const double2 *__restrict__ k, double2 *__restrict__ z, int B,
                           long long n, int conj)
{
    const long long total = (long long)B * n;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const double2 kv = k[e % n];
        z[e] = cmul(a[e], conj ? cconj(kv) : kv);
    }
}

// r[m] = sum_b conj(a[b, m]) * g[b, m]
__global__ void k_spec_corr_reduce(const double2 *__restrict__ a, const double2 *__restrict__ g, double2 *__restrict__ r,
                                   int B, long long n)
{
    for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (long long)gridDim.x * blockDim.x) {
        double2 acc = make_double2(0.0, 0.0);
        for (int b = 0; b < B; ++b) acc = cadd(acc, cmul(cconj(a[(long long)b * n + m]), g[(long long)b * n + m]));
        r[m] = acc;
    }
}

static unsigned blocks_for(long long total)
{
    long long b = (total + 255) / 256;
    return (unsigned)std::max<long long>(1, std::min<long long>(b, 1 << 16));
}

static bool smooth13_c(long long n)
{
    for (int p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}

static long long conv_len(long long need, bool even)
{
    for (long long n = need;; ++n)
        if ((!even || n % 2 == 0) && smooth13_c(n) && (!even || smooth13_c(n / 2))) return n;
}

struct Conv2d {
    int B = 0, H = 0, W = 0, KH = 0, KW = 0, dtype = 0;   // dtype 0: f64, 1: f32
    long long P = 0, Q = 0, nspec = 0;                     // grid and half-spectrum points per image
    std::unique_ptr<Plan> r2c_b, r2c_1, c2r_b, c2r_adj_b, r2c_adj_b, r2c_adj_1;
    DevBuf pad, spec_a, spec_b, kspec, kpad;
};

static int conv_plan(std::unique_ptr<Plan> &p, int kind, int batch, long long P, long long Q, int forward, Error &err)
{
    const size_t shp3[3] = {(size_t)batch, (size_t)P, (size_t)Q};
    const long long Qh = Q / 2 + 1;
    // real side C-contiguous [B, P, Q]; complex side C-contiguous [B, P, Q/2+1]
    const ptrdiff_t sr[3] = {(ptrdiff_t)(P * Q * 8), (ptrdiff_t)(Q * 8), 8};
    const ptrdiff_t sc[3] = {(ptrdiff_t)(P * Qh * 16), (ptrdiff_t)(Qh * 16), 16};
    const size_t axes[2] = {1, 2};
    const bool real_in = (kind == 1 || kind == 4);
    return plan_create(p, kind, 0, 3, shp3, real_in ? sr : sc, real_in ? sc : sr, 2, axes, forward, err);
}

int conv2d_create(Conv2d **out, int B, int H, int W, int KH, int KW, int dtype, Error &err)
{
    if (B < 1 || H < 1 || W < 1 || KH < 1 || KW < 1 || KH > H || KW > W || (dtype != 0 && dtype != 1)) {
        err = {E_INVALID, "invalid convolution shape or dtype"};
        return E_INVALID;
    }
    auto c = std::make_unique<Conv2d>();
    c->B = B; c->H = H; c->W = W; c->KH = KH; c->KW = KW; c->dtype = dtype;
    c->P = conv_len((long long)H + KH - 1, false);
    c->Q = conv_len((long long)W + KW - 1, true);
    c->nspec = c->P * (c->Q / 2 + 1);
    int rc;
    if ((rc = conv_plan(c->r2c_b, 1, B, c->P, c->Q, 1, err)) != OK) return rc;
    if ((rc = conv_plan(c->r2c_1, 1, 1, c->P, c->Q, 1, err)) != OK) return rc;
    if ((rc = conv_plan(c->c2r_b, 2, B, c->P, c->Q, 0, err)) != OK) return rc;
    if ((rc = conv_plan(c->c2r_adj_b, 4, B, c->P, c->Q, 0, err)) != OK) return rc;
    if ((rc = conv_plan(c->r2c_adj_b, 3, B, c->P, c->Q, 1, err)) != OK) return rc;
    if ((rc = conv_plan(c->r2c_adj_1, 3, 1, c->P, c->Q, 1, err)) != OK) return rc;
    const size_t real_b = (size_t)B * c->P * c->Q * 8, spec_b = (size_t)B * c->nspec * 16;
    if (c->pad.alloc(real_b) || c->spec_a.alloc(spec_b) || c->spec_b.alloc(spec_b) || c->kspec.alloc(c->nspec * 16) ||
        c->kpad.alloc((size_t)c->P * c->Q * 8)) {
        err = {E_NOMEM, "convolution workspace allocation failed"};
        return E_NOMEM;
    }
    *out = c.release();
    return OK;
}

static hipError_t pad(const Conv2d &c, const void *src, bool f32, double *dst, int B, int H, int W, int oy, int ox,
                      hipStream_t s)
{
    const unsigned g = blocks_for((long long)B * c.P * c.Q);
    if (f32) hipLaunchKernelGGL(k_pad2d<float>, dim3(g), dim3(256), 0, s, (const float *)src, dst, B, H, W, (int)c.P, (int)c.Q, oy, ox);
    else hipLaunchKernelGGL(k_pad2d<double>, dim3(g), dim3(256), 0, s, (const double *)src, dst, B, H, W, (int)c.P, (int)c.Q, oy, ox);
    return hipGetLastError();
}

static hipError_t crop(const Conv2d &c, const double *src, void *dst, bool f32, int B, int H, int W, int oy, int ox,
                       double scale, hipStream_t s)
{
    const unsigned g = blocks_for((long long)B * H * W);
    if (f32) hipLaunchKernelGGL(k_crop2d<float>, dim3(g), dim3(256), 0, s, src, (float *)dst, B, H, W, (int)c.P, (int)c.Q, oy, ox, scale);
    else hipLaunchKernelGGL(k_crop2d<double>, dim3(g), dim3(256), 0, s, src, (double *)dst, B, H, W, (int)c.P, (int)c.Q, oy, ox, scale);
    return hipGetLastError();
}

#define IFFT_TRY_HIP(expr)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            err = {E_HIP, std::string("convolution launch failed: ") + hipGetErrorString(e_)}; \
            return E_HIP;                                                                   \
        }                                                                                   \
    } while (0)
#define IFFT_TRY(expr)               \
    do {                             \
        int r_ = (expr);             \
        if (r_ != OK) return r_;     \
    } while (0)

// K^ = r2c(pad_0 k)
static int kernel_spectrum(Conv2d &c, const void *k, hipStream_t s, Error &err)
{
    IFFT_TRY_HIP(pad(c, k, c.dtype == 1, (double *)c.kpad.p, 1, c.KH, c.KW, 0, 0, s));
    return plan_execute(*c.r2c_1, c.kpad.p, c.kspec.p, 1.0, s, err);
}

int conv2d_forward(Conv2d &c, const void *x, const void *k, void *y, hipStream_t s, Error &err)
{
    const bool f32 = c.dtype == 1;
    IFFT_TRY(kernel_spectrum(c, k, s, err));
    IFFT_TRY_HIP(pad(c, x, f32, (double *)c.pad.p, c.B, c.H, c.W, 0, 0, s));
    IFFT_TRY(plan_execute(*c.r2c_b, c.pad.p, c.spec_a.p, 1.0, s, err));
    hipLaunchKernelGGL(k_spec_mul, dim3(blocks_for(c.B * c.nspec)), dim3(256), 0, s, (const double2 *)c.spec_a.p,
                       (const double2 *)c.kspec.p, (double2 *)c.spec_a.p, c.B, c.nspec, 0);
    IFFT_TRY_HIP(hipGetLastError());
    IFFT_TRY(plan_execute(*c.c2r_b, c.spec_a.p, c.pad.p, 1.0, s, err));
    const int cy = (c.KH - 1) / 2, cx = (c.KW - 1) / 2;
    IFFT_TRY_HIP(crop(c, (const double *)c.pad.p, y, f32, c.B, c.H, c.W, cy, cx, 1.0 / ((double)c.P * c.Q), s));
    return OK;
}

int conv2d_backward(Conv2d &c, const void *x, const void *k, const void *gy, void *gx, void *gk, hipStream_t s, Error &err)
{
    const bool f32 = c.dtype == 1;
    const int cy = (c.KH - 1) / 2, cx = (c.KW - 1) / 2;
    const double inv = 1.0 / ((double)c.P * c.Q);
    IFFT_TRY(kernel_spectrum(c, k, s, err));
    // G_b = c2r*(pad_c g_b) / (P Q)
    IFFT_TRY_HIP(pad(c, gy, f32, (double *)c.pad.p, c.B, c.H, c.W, cy, cx, s));
    IFFT_TRY(plan_execute(*c.c2r_adj_b, c.pad.p, c.spec_b.p, inv, s, err));
    if (gk) {
        // X^_b recomputed, then sum_b conj(X^_b) G_b -> r2c* -> crop
        IFFT_TRY_HIP(pad(c, x, f32, (double *)c.pad.p, c.B, c.H, c.W, 0, 0, s));
        IFFT_TRY(plan_execute(*c.r2c_b, c.pad.p, c.spec_a.p, 1.0, s, err));
        hipLaunchKernelGGL(k_spec_corr_reduce, dim3(blocks_for(c.nspec)), dim3(256), 0, s, (const double2 *)c.spec_a.p,
                           (const double2 *)c.spec_b.p, (double2 *)c.spec_a.p, c.B, c.nspec);
        IFFT_TRY_HIP(hipGetLastError());
        IFFT_TRY(plan_execute(*c.r2c_adj_1, c.spec_a.p, c.kpad.p, 1.0, s, err));
        IFFT_TRY_HIP(crop(c, (const double *)c.kpad.p, gk, f32, 1, c.KH, c.KW, 0, 0, 1.0, s));
    }
    if (gx) {
        hipLaunchKernelGGL(k_spec_mul, dim3(blocks_for(c.B * c.nspec)), dim3(256), 0, s, (const double2 *)c.spec_b.p,
                           (const double2 *)c.kspec.p, (double2 *)c.spec_b.p, c.B, c.nspec, 1);
        IFFT_TRY_HIP(hipGetLastError());
        IFFT_TRY(plan_execute(*c.r2c_adj_b, c.spec_b.p, c.pad.p, 1.0, s, err));
        IFFT_TRY_HIP(crop(c, (const double *)c.pad.p, gx, f32, c.B, c.H, c.W, 0, 0, 1.0, s));
    }
    return OK;
}

void conv2d_grid(const Conv2d &c, long long &P, long long &Q) { P = c.P; Q = c.Q; }
void conv2d_destroy(Conv2d *c) { delete c; }

}  // namespace ifft

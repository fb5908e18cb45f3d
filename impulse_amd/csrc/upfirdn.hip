// upfirdn: upsample by `up` (up-1 zeros after every sample), filter with the
// FIR taps h (full convolution), keep every `down`-th sample from the first
// (impulse/signal.nim:545-594, `t.upsample(up).convolve(h, full, down)`).
// With u the upsampled signal (length Lu = (nt-1) up + 1) and c = u * h
// (length Lc = Lu + nh - 1), the output is y[j] = c[j down], j < ceil(Lc / down).
//
// Two algorithms over FP64, chosen per plan by a cost model (or forced):
//   DIRECT  polyphase sums: c[m] = sum_q t[q] h[m - q up] over the q with
//           0 <= m - q up < nh, one output per thread, taps staged in LDS when
//           they fit.  About nh / up fused multiply-adds per output.
//   FFT     overlap-save over u: blocks of B = P - (nh-1) consecutive c values,
//           each from one length-P circular convolution
//             seg_k[i] = u[k B - (nh-1) + i]       (gather, upsampling fused)
//             seg_k    = c2r( r2c(seg_k) . H / P ) (the engine's batched plans)
//             y[j]     = seg_k[j down - k B + nh - 1], k = floor(j down / B)
//           with H = r2c(h zero-padded to P) computed once at plan creation.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "engine.h"
#include "fft_common.h"

namespace ifft {

constexpr int kUfdTileR = 4;             // outputs per thread of the tiled direct kernel
constexpr int kUfdLdsBytes = 64 * 1024;  // LDS per workgroup of the tiled kernel (two per CU)

// Polyphase sum of output j: with m = j down, qhi = floor(m / up) and phase
// r = m - qhi up, y[j] = sum_l t[qhi - l] h[r + l up] over the taps l with
// r + l up < nh and 0 <= qhi - l < nt.  `ts` holds t from index q0 on.
__device__ __forceinline__ double ufd_output(const double *ts, long long q0, const double *hs, long long j,
                                             long long nt, int nh, int up, int down)
{
    const long long m = j * down;
    long long qhi;
    int r;
    if (m <= 0xffffffffLL) {
        const unsigned mm = (unsigned)m, q = mm / (unsigned)up;
        qhi = q;
        r = (int)(mm - q * (unsigned)up);
    } else {
        qhi = m / up;
        r = (int)(m - qhi * up);
    }
    if (r >= nh) return 0.0;
    long long lmax = (nh - 1 - r) / up;                     // last tap of this phase
    if (lmax > qhi) lmax = qhi;                             // t[qhi - l] exists only for l <= qhi
    const long long l0 = qhi > nt - 1 ? qhi - (nt - 1) : 0; // taps past the end of t
    const double *tp = ts + (qhi - l0 - q0);
    const double *hp = hs + r + l0 * up;
    int c = (int)(lmax - l0 + 1);
    // four independent chains keep several loads in flight per lane
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (; c >= 4; c -= 4) {
        const double t0 = tp[0], t1 = tp[-1], t2 = tp[-2], t3 = tp[-3];
        const double h0 = hp[0], h1 = hp[up], h2 = hp[2 * up], h3 = hp[3 * up];
        a0 = fma(t0, h0, a0);
        a1 = fma(t1, h1, a1);
        a2 = fma(t2, h2, a2);
        a3 = fma(t3, h3, a3);
        tp -= 4;
        hp += 4 * up;
    }
    for (; c > 0; --c) {
        a0 = fma(*tp, *hp, a0);
        --tp;
        hp += up;
    }
    return (a0 + a1) + (a2 + a3);
}

// Tiled polyphase kernel: workgroup (blockIdx.x, row blockIdx.y) computes the
// outputs J0 .. J0 + 256 R - 1 from LDS copies of the taps and of the signal
// segment those outputs read (its length, seg_cap, is bounded on the host).
template <int R>
__global__ void __launch_bounds__(256) k_upfirdn_tile(const double *__restrict__ t, const double *__restrict__ h,
                                                      double *__restrict__ y, long long nt, int nh, int up, int down,
                                                      long long nout, int seg_cap)
{
    extern __shared__ double sm[];
    double *hs = sm;
    double *ts = sm + nh;
    const double *tb = t + (long long)blockIdx.y * nt;
    double *yb = y + (long long)blockIdx.y * nout;
    const long long J0 = (long long)blockIdx.x * (256 * R);
    const long long J1 = (J0 + 256 * R < nout ? J0 + 256 * R : nout) - 1;
    const long long first = J0 * down - (nh - 1);
    const long long q0 = first <= 0 ? 0 : first / up;
    long long q1 = J1 * down / up;
    if (q1 > nt - 1) q1 = nt - 1;
    const int nseg = (int)(q1 - q0 + 1);
    for (int i = threadIdx.x; i < nh; i += 256) hs[i] = h[i];
    for (int i = threadIdx.x; i < nseg && i < seg_cap; i += 256) ts[i] = tb[q0 + i];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const long long j = J0 + k * 256 + threadIdx.x;
        if (j < nout) yb[j] = ufd_output(ts, q0, hs, j, nt, nh, up, down);
    }
}

// Fallback for filters or rate ratios whose tile does not fit LDS: taps and
// signal read from global memory, one output per thread.
__global__ void __launch_bounds__(256) k_upfirdn_direct(const double *__restrict__ t, const double *__restrict__ h,
                                                        double *__restrict__ y, long long nt, int nh, int up, int down,
                                                        long long nout)
{
    const double *tb = t + (long long)blockIdx.y * nt;
    double *yb = y + (long long)blockIdx.y * nout;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < nout;
         j += (long long)gridDim.x * blockDim.x)
        yb[j] = ufd_output(tb, 0, h, j, nt, nh, up, down);
}

__global__ void k_upfirdn_gather(const double *__restrict__ t, double *__restrict__ seg, long long nt, int up,
                                 long long Lu, int nh, long long B, long long P, long long nblk, long long total)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long row = e / P, i = e - row * P;
        const long long b = row / nblk, k = row - b * nblk;
        const long long m = k * B - (nh - 1) + i;
        double v = 0.0;
        if (m >= 0 && m < Lu && m % up == 0) v = t[b * nt + m / up];
        seg[e] = v;
    }
}

__global__ void k_upfirdn_mul(double2 *__restrict__ spec, const double2 *__restrict__ hs, long long nspec,
                              double scale, long long total)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const double2 w = hs[e % nspec];
        spec[e] = cscale(cmul(spec[e], w), scale);
    }
}

__global__ void k_upfirdn_scatter(const double *__restrict__ seg, double *__restrict__ y, int down, int nh,
                                  long long B, long long P, long long nblk, long long nout, long long total)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long b = e / nout, j = e - b * nout;
        const long long m = j * down, k = m / B;
        y[e] = seg[(b * nblk + k) * P + (m - k * B) + nh - 1];
    }
}

static unsigned ufd_blocks(long long total)
{
    const long long b = (total + 255) / 256;
    return (unsigned)std::max<long long>(1, std::min<long long>(b, 1 << 20));
}

static bool smooth7(long long n)
{
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

static long long even_smooth7(long long need)
{
    for (long long n = std::max<long long>(need, 2);; ++n)
        if (n % 2 == 0 && smooth7(n)) return n;
}

struct Upfirdn {
    long long batch = 0, nt = 0, nh = 0, nout = 0, Lu = 0, Lc = 0;
    int up = 1, down = 1, algo = 1;
    DevBuf h;                                  // taps (direct)
    int seg_cap = 0;                           // signal points per tile in LDS (0: untiled direct kernel)
    long long P = 0, B = 0, nblk = 0, nspec = 0;
    std::unique_ptr<Plan> r2c, c2r;
    DevBuf hspec, seg, spec;
    std::mutex host_mtx;
    DevBuf stage_t, stage_y;
};

// modelled seconds of each algorithm (FP64 FMA rate and HBM rate discounted to what these kernels reach)
static double direct_cost(const Upfirdn &U)
{
    const double taps = std::ceil((double)U.nh / U.up);
    return 2.0 * U.batch * U.nout * taps / 30e12 + (double)U.batch * (U.nt + U.nout) * 8 / 5e12;
}

static void fft_shape(Upfirdn &U)
{
    const long long nh1 = U.nh - 1;
    const long long need = (U.nout - 1) * U.down + 1;   // c values the output reads
    long long target = 4096;
    while (target < 8 * nh1 && target < (1LL << 22)) target *= 2;
    if (target <= nh1) target = even_smooth7(2 * nh1 + 2);
    const long long single = even_smooth7(need + nh1);
    if (single <= target) {
        U.P = single;
        U.nblk = 1;
    } else {
        U.P = target;
        U.nblk = (need + (U.P - nh1) - 1) / (U.P - nh1);
    }
    U.B = U.P - nh1;
    U.nspec = U.P / 2 + 1;
}

static bool tiles_fit(const Upfirdn &U)
{
    const long long tiles = (U.nout + 256 * kUfdTileR - 1) / (256 * kUfdTileR);
    return tiles < (1LL << 31);
}

static double fft_cost(const Upfirdn &U)
{
    const double rows = (double)U.batch * U.nblk;
    return 10.0 * rows * U.P * 8 / 4e12 + 6 * 5e-6;
}

int upfirdn_create(Upfirdn **out, size_t batch, size_t nt, const double *h, size_t nh, int up, int down, int algo,
                   Error &err)
{
    if (batch < 1 || nt < 1 || nh < 1 || !h || up < 1 || down < 1 || algo < 0 || algo > 2) {
        err = {E_INVALID, "upfirdn needs batch, signal and filter lengths >= 1, up and down >= 1, algorithm 0-2"};
        return E_INVALID;
    }
    if (nh > (size_t)INT32_MAX || nt > (size_t)1 << 40 || batch > 65535) {
        err = {E_UNSUPPORTED, "upfirdn signal or filter too long"};
        return E_UNSUPPORTED;
    }
    int rc = device_check(err);
    if (rc != OK) return rc;
    auto U = std::make_unique<Upfirdn>();
    U->batch = (long long)batch;
    U->nt = (long long)nt;
    U->nh = (long long)nh;
    U->up = up;
    U->down = down;
    U->Lu = (U->nt - 1) * up + 1;
    U->Lc = U->Lu + U->nh - 1;
    U->nout = (U->Lc + down - 1) / down;
    fft_shape(*U);
    U->algo = algo ? algo : (fft_cost(*U) < direct_cost(*U) ? 2 : 1);
    {
        // signal points a tile of 256 R outputs reads: floor(J1 down / up) - floor((J0 down - nh + 1) / up) + 1
        const long long span = ((256LL * kUfdTileR - 1) * down + U->nh - 1) / up + 2;
        if ((U->nh + span) * 8 <= kUfdLdsBytes && tiles_fit(*U)) U->seg_cap = (int)span;
    }
    if (U->algo == 1) {
        if (U->h.alloc(nh * 8) || hipMemcpy(U->h.p, h, nh * 8, hipMemcpyHostToDevice) != hipSuccess) {
            err = {E_NOMEM, "upfirdn filter upload failed"};
            return E_NOMEM;
        }
    } else {
        const long long rows = U->batch * U->nblk;
        const size_t shp[2] = {(size_t)rows, (size_t)U->P};
        const ptrdiff_t sr[2] = {(ptrdiff_t)(U->P * 8), 8};
        const ptrdiff_t sc[2] = {(ptrdiff_t)(U->nspec * 16), 16};
        const size_t ax[1] = {1};
        if ((rc = plan_create(U->r2c, 1, 0, 2, shp, sr, sc, 1, ax, 1, err)) != OK) return rc;
        if ((rc = plan_create(U->c2r, 2, 0, 2, shp, sc, sr, 1, ax, 0, err)) != OK) return rc;
        if (U->seg.alloc((size_t)rows * U->P * 8) || U->spec.alloc((size_t)rows * U->nspec * 16) ||
            U->hspec.alloc((size_t)U->nspec * 16)) {
            err = {E_NOMEM, "upfirdn workspace allocation failed"};
            return E_NOMEM;
        }
        // H = r2c(h zero-padded to P), once
        std::unique_ptr<Plan> one;
        const size_t shp1[1] = {(size_t)U->P};
        const ptrdiff_t sr1[1] = {8}, sc1[1] = {16};
        const size_t ax1[1] = {0};
        if ((rc = plan_create(one, 1, 0, 1, shp1, sr1, sc1, 1, ax1, 1, err)) != OK) return rc;
        std::vector<double> hp((size_t)U->P, 0.0);
        std::memcpy(hp.data(), h, nh * 8);
        if (hipMemcpy(U->seg.p, hp.data(), (size_t)U->P * 8, hipMemcpyHostToDevice) != hipSuccess) {
            err = {E_HIP, "upfirdn filter upload failed"};
            return E_HIP;
        }
        if ((rc = plan_execute(*one, U->seg.p, U->hspec.p, 1.0, nullptr, err)) != OK) return rc;
        if (hipDeviceSynchronize() != hipSuccess) {
            err = {E_HIP, "upfirdn filter spectrum failed"};
            return E_HIP;
        }
    }
    *out = U.release();
    return OK;
}

#define UFD_TRY_HIP(expr)                                                                 \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            err = {E_HIP, std::string("upfirdn launch failed: ") + hipGetErrorString(e_)}; \
            return E_HIP;                                                                 \
        }                                                                                 \
    } while (0)

int upfirdn_execute(Upfirdn &U, const double *t, double *y, hipStream_t s, Error &err)
{
    if (!t || !y) {
        err = {E_INVALID, "upfirdn buffers must not be NULL"};
        return E_INVALID;
    }
    const long long total = U.batch * U.nout;
    if (U.algo == 1) {
        if (U.seg_cap > 0) {
            const long long tiles = (U.nout + 256 * kUfdTileR - 1) / (256 * kUfdTileR);
            const size_t lds = (size_t)(U.nh + U.seg_cap) * 8;
            hipLaunchKernelGGL(k_upfirdn_tile<kUfdTileR>, dim3((unsigned)tiles, (unsigned)U.batch), dim3(256), lds, s,
                               t, (const double *)U.h.p, y, U.nt, (int)U.nh, U.up, U.down, U.nout, U.seg_cap);
        } else {
            hipLaunchKernelGGL(k_upfirdn_direct, dim3(ufd_blocks(U.nout), (unsigned)U.batch), dim3(256), 0, s, t,
                               (const double *)U.h.p, y, U.nt, (int)U.nh, U.up, U.down, U.nout);
        }
        UFD_TRY_HIP(hipGetLastError());
        return OK;
    }
    const long long rows = U.batch * U.nblk;
    hipLaunchKernelGGL(k_upfirdn_gather, dim3(ufd_blocks(rows * U.P)), dim3(256), 0, s, t, (double *)U.seg.p, U.nt,
                       U.up, U.Lu, (int)U.nh, U.B, U.P, U.nblk, rows * U.P);
    UFD_TRY_HIP(hipGetLastError());
    int rc = plan_execute(*U.r2c, U.seg.p, U.spec.p, 1.0, s, err);
    if (rc != OK) return rc;
    hipLaunchKernelGGL(k_upfirdn_mul, dim3(ufd_blocks(rows * U.nspec)), dim3(256), 0, s, (double2 *)U.spec.p,
                       (const double2 *)U.hspec.p, U.nspec, 1.0 / (double)U.P, rows * U.nspec);
    UFD_TRY_HIP(hipGetLastError());
    rc = plan_execute(*U.c2r, U.spec.p, U.seg.p, 1.0, s, err);
    if (rc != OK) return rc;
    hipLaunchKernelGGL(k_upfirdn_scatter, dim3(ufd_blocks(total)), dim3(256), 0, s, (const double *)U.seg.p, y,
                       U.down, (int)U.nh, U.B, U.P, U.nblk, U.nout, total);
    UFD_TRY_HIP(hipGetLastError());
    return OK;
}

int upfirdn_execute_host(Upfirdn &U, const double *t, double *y, Error &err)
{
    if (!t || !y) {
        err = {E_INVALID, "upfirdn buffers must not be NULL"};
        return E_INVALID;
    }
    std::lock_guard<std::mutex> lk(U.host_mtx);
    const size_t tb = (size_t)(U.batch * U.nt) * 8, yb = (size_t)(U.batch * U.nout) * 8;
    if ((!U.stage_t.p && U.stage_t.alloc(tb)) || (!U.stage_y.p && U.stage_y.alloc(yb))) {
        err = {E_NOMEM, "upfirdn staging allocation failed"};
        return E_NOMEM;
    }
    UFD_TRY_HIP(hipMemcpy(U.stage_t.p, t, tb, hipMemcpyHostToDevice));
    int rc = upfirdn_execute(U, (const double *)U.stage_t.p, (double *)U.stage_y.p, nullptr, err);
    if (rc != OK) return rc;
    UFD_TRY_HIP(hipMemcpy(y, U.stage_y.p, yb, hipMemcpyDeviceToHost));
    return OK;
}

void upfirdn_info(const Upfirdn &U, size_t *nout, int *algo, size_t *fft_length, size_t *blocks)
{
    if (nout) *nout = (size_t)U.nout;
    if (algo) *algo = U.algo;
    if (fft_length) *fft_length = U.algo == 2 ? (size_t)U.P : 0;
    if (blocks) *blocks = U.algo == 2 ? (size_t)U.nblk : 0;
}

void upfirdn_destroy(Upfirdn *u) { delete u; }

}  // namespace ifft

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
//
// Fused path (7-smooth grids with Q/2 <= 2560 and P <= 2560): three kinds of LDS kernels
// over half spectra kept in row layout S[img][r][j] (j < Qh = Q/2 + 1), and only
// for the rows an image occupies:
//   rows   : real rows read straight from the caller's array (the padding is
//            generated, never stored), r2c of length Q by the half-length
//            complex transform, half spectrum out;
//   columns: each column group loads its image rows at their grid offset (the
//            other P - H rows are zeros in LDS), transforms, multiplies by K^
//            (conj K^ for the adjoint) and transforms back, keeping only the
//            rows the crop needs -- or, for the kernel gradient, sums
//            conj(X^_b) G^_b over the images of a chunk;
//   rows^-1: c2r of the kept rows, the cropped columns scaled and stored in the
//            caller's dtype.
// Three passes over half spectra instead of pad / r2c / multiply / c2r / crop
// over the whole padded grid.  Larger grids take the plan-based path.
#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"
#include "fft_common.h"
#include "lds_fft.h"

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

// ---------------------------------------------------------------------------
// fused path
// ---------------------------------------------------------------------------

// The LDS transforms are inlined (no call frames in scratch memory) and use
// radices 2-8 only, so the fused path runs on 7-smooth grids: kConvPPT points a
// thread at most, 512-thread row kernels (two workgroups a CU) and 1,024-thread
// column kernels, all within 128 registers.
constexpr int kConvPPT = 5;
constexpr int kConvOrder = 2;          // powers of two as radix-8 stages
constexpr int kConvRowThreads = 512;
constexpr int kConvColThreads = 1024;
constexpr int kConvColPoints = kConvPPT * kConvColThreads;   // points of one column LDS batch

// real rows -> half spectra: row q of src (W values, placed at grid columns
// [rx, rx + W) of a Q-point row, zeros elsewhere) -> S[q][0 .. Q/2]
struct ConvRows {
    const void *src;
    double2 *S;
    const double2 *tq;   // W_Q^j = exp(-2 pi i j / Q), j < Q
    long long rows;
    int W, Q, rx, f32, count, lc;
};

// half spectra -> real rows: S[q][0 .. Q/2] -> dst[q][c] = scale * x[ox + c], c < Wout,
// x the unnormalised c2r of length Q
struct ConvRowsInv {
    const double2 *S;
    void *dst;
    const double2 *tq;
    long long rows;
    int Wout, Q, ox, f32, count, lc;
    double scale;
};

// column groups of `count` spectrum columns (grid rows 0 .. P-1)
struct ConvCols {
    const double2 *Sa;   // image rows [img][Ha][Qh], placed at grid rows [ra, ra + Ha)
    const double2 *Sg;   // kernel gradient: gradient rows [img][Ha][Qh] at grid rows [rg, rg + Ha)
    double2 *Sout;       // kept rows [oy, oy + Ho) as [img][Ho][Qh]; spectrum mode: [Qh][P]
    const double2 *KT;   // kernel spectrum [Qh][P]
    double2 *acc;        // kernel-gradient partial sums [nchunk][Qh][P]
    const double2 *tp;   // W_P^j, j < P
    int P, Qh, Ha, ra, rg, oy, Ho, nimg, chunk, nchunk, count, lc, conjk;
};

enum ConvColMode { CC_SPECTRUM = 0, CC_CONV = 1, CC_GK = 2, CC_GK_FINAL = 3 };

template <int THREADS>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) k_conv_rows(const ConvRows P)
{
    extern __shared__ double2 zl[];
    const int M = P.Q >> 1, Qh = M + 1;
    const int NT = THREADS >> P.lc;
    const int t = (int)threadIdx.x / NT, i = (int)threadIdx.x - t * NT;
    for (long long r0 = (long long)blockIdx.x * P.count; r0 < P.rows; r0 += (long long)gridDim.x * P.count) {
        const long long row = r0 + t;
        const bool valid = row < P.rows;
        if (valid) {
            const long long base = row * P.W;
            for (int m = i; m < M; m += NT) {
                const int c0 = 2 * m - P.rx, c1 = c0 + 1;
                const double a = (c0 >= 0 && c0 < P.W) ? gld_r(P.src, P.f32, base + c0) : 0.0;
                const double b = (c1 >= 0 && c1 < P.W) ? gld_r(P.src, P.f32, base + c1) : 0.0;
                zl[lpad(t * M + m)] = make_double2(a, b);
            }
        }
        __syncthreads();
        lds_fft_inline<kConvPPT, true>(zl, M, P.count, -1.0, kConvOrder, TwTable{P.tq, P.Q});   // ends with a barrier
        if (valid) {
            // X[k] = E + W_Q^k O, E = (Z[k] + conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / 2i, k <= M
            double2 *out = P.S + row * Qh;
            for (int k = i; k <= M; k += NT) {
                const double2 zk = zl[lpad(t * M + (k == M ? 0 : k))];
                const double2 zm = zl[lpad(t * M + (k == 0 ? 0 : M - k))];
                const double2 e = cscale(cadd(zk, cconj(zm)), 0.5), d = csub(zk, cconj(zm));
                const double2 o = make_double2(0.5 * d.y, -0.5 * d.x);
                out[k] = cadd(e, cmul(o, P.tq[k]));
            }
        }
        __syncthreads();
    }
}

template <int THREADS>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) k_conv_rows_inv(const ConvRowsInv P)
{
    extern __shared__ double2 zl[];
    const int M = P.Q >> 1, Qh = M + 1;
    const int NT = THREADS >> P.lc;
    const int t = (int)threadIdx.x / NT, i = (int)threadIdx.x - t * NT;
    for (long long r0 = (long long)blockIdx.x * P.count; r0 < P.rows; r0 += (long long)gridDim.x * P.count) {
        const long long row = r0 + t;
        const bool valid = row < P.rows;
        if (valid) {
            // Z[k] = E + i O, E = (X[k] + conj X[M-k]) / 2, O = (X[k] - conj X[M-k]) conj(W_Q^k) / 2;
            // the imaginary parts of X[0] and X[M] are ignored, as c2r does
            const double2 *X = P.S + row * Qh;
            for (int k = i; k < M; k += NT) {
                double2 a = X[k], b = X[M - k];
                if (k == 0) {
                    a.y = 0.0;
                    b.y = 0.0;
                }
                const double2 e = cscale(cadd(a, cconj(b)), 0.5);
                const double2 o = cmul(cscale(csub(a, cconj(b)), 0.5), cconj(P.tq[k]));
                zl[lpad(t * M + k)] = make_double2(e.x - o.y, e.y + o.x);
            }
        }
        __syncthreads();
        lds_fft_inline<kConvPPT, true>(zl, M, P.count, 1.0, kConvOrder, TwTable{P.tq, P.Q});   // ends with a barrier
        if (valid) {
            // x[2m] = 2 Re z[m], x[2m+1] = 2 Im z[m] for the unnormalised length-Q transform
            const double f = 2.0 * P.scale;
            for (int c = i; c < P.Wout; c += NT) {
                const int q = c + P.ox;
                const double2 z = zl[lpad(t * M + (q >> 1))];
                gst_r(P.dst, P.f32, row * P.Wout + c, f * ((q & 1) ? z.y : z.x));
            }
        }
        __syncthreads();
    }
}

// One workgroup per column group (and, for the convolution and the backward
// pass, per chunk of images).  Column positions run with the column index
// fastest (e = m * count + t), so a wave reads and writes runs of `count`
// neighbouring bins of a spectrum row.
template <int MODE>
__global__ void __launch_bounds__(kConvColThreads) 
k_conv_cols(const ConvCols C)
{
    extern __shared__ double2 zl[];
    constexpr int T = kConvColThreads;
    const int P = C.P, Qh = C.Qh, cnt = C.count, lc = C.lc, n = cnt * P;
    const int ngroups = (Qh + cnt - 1) >> lc;
    const int logical = (int)xcd_block(blockIdx.x, gridDim.x);
    const int group = logical % ngroups, chunk = logical / ngroups;
    const int tid = (int)threadIdx.x;
    const TwTable tw{C.tp, P};

    // image rows of column group `group` at grid rows [r0, r0 + Ha), zeros elsewhere
    constexpr int NQ = kConvColPoints / T + 1;   // column-group positions a thread owns: e = tid + q T
    auto load = [&](double2 *buf, const double2 *S, long long img, int r0) {
        for (int e = tid; e < n; e += T) {
            const int t = e & (cnt - 1), m = e >> lc, j = group * cnt + t, r = m - r0;
            double2 v = make_double2(0.0, 0.0);
            if (j < Qh && r >= 0 && r < C.Ha) v = S[(img * C.Ha + r) * Qh + j];
            buf[lpad(t * P + m)] = v;
        }
        __syncthreads();
    };
    auto store = [&](const double2 *buf, long long img) {
        for (int e = tid; e < cnt * C.Ho; e += T) {
            const int t = e & (cnt - 1), r = e >> lc, j = group * cnt + t;
            if (j < Qh) C.Sout[(img * C.Ho + r) * Qh + j] = buf[lpad(t * P + C.oy + r)];
        }
        __syncthreads();
    };
    // the kernel spectrum at the thread's positions, read once for all images of the chunk
    auto kernel_values = [&](double2 (&kr)[NQ], bool conj) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tid + q * T, t = e & (cnt - 1), m = e >> lc, j = group * cnt + t;
            double2 k = make_double2(0.0, 0.0);
            if (e < n && j < Qh) k = C.KT[(long long)j * P + m];
            kr[q] = conj ? cconj(k) : k;
        }
    };

    if constexpr (MODE == CC_SPECTRUM) {
        load(zl, C.Sa, 0, C.ra);
        lds_fft_inline<kConvPPT, true>(zl, P, cnt, -1.0, kConvOrder, tw);
        for (int e = tid; e < n; e += T) {
            const int t = e & (cnt - 1), m = e >> lc, j = group * cnt + t;
            if (j < Qh) C.Sout[(long long)j * P + m] = zl[lpad(t * P + m)];
        }
    } else if constexpr (MODE == CC_CONV) {
        double2 kr[NQ];
        kernel_values(kr, C.conjk != 0);
        const int b1 = min(C.nimg, (chunk + 1) * C.chunk);
        for (int b = chunk * C.chunk; b < b1; ++b) {
            load(zl, C.Sa, b, C.ra);
            lds_fft_inline<kConvPPT, true>(zl, P, cnt, -1.0, kConvOrder, tw);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int e = tid + q * T;
                if (e < n) {
                    const int p = lpad((e & (cnt - 1)) * P + (e >> lc));
                    zl[p] = cmul(zl[p], kr[q]);
                }
            }
            __syncthreads();
            lds_fft_inline<kConvPPT, true>(zl, P, cnt, 1.0, kConvOrder, tw);
            store(zl, b);
        }
    } else if constexpr (MODE == CC_GK) {
        // x and g spectra of the column group in one LDS batch (lines 0 .. cnt-1 and cnt .. 2cnt-1),
        // transformed together; the chunk's sum of conj(X^) G^ in registers
        constexpr int NQG = kConvPPT * T / 2 / T + 1;
        double2 acc[NQG];
#pragma unroll
        for (int q = 0; q < NQG; ++q) acc[q] = make_double2(0.0, 0.0);
        const int b1 = min(C.nimg, (chunk + 1) * C.chunk);
        for (int b = chunk * C.chunk; b < b1; ++b) {
            for (int e = tid; e < 2 * n; e += T) {
                const int half = e >= n ? 1 : 0, f = e - half * n;
                const int t = f & (cnt - 1), m = f >> lc, j = group * cnt + t, r = m - (half ? C.rg : C.ra);
                double2 v = make_double2(0.0, 0.0);
                if (j < Qh && r >= 0 && r < C.Ha) v = (half ? C.Sg : C.Sa)[((long long)b * C.Ha + r) * Qh + j];
                zl[lpad((half * cnt + t) * P + m)] = v;
            }
            __syncthreads();
            lds_fft_inline<kConvPPT, true>(zl, P, 2 * cnt, -1.0, kConvOrder, tw);
#pragma unroll
            for (int q = 0; q < NQG; ++q) {
                const int e = tid + q * T;
                if (e < n) {
                    const int t = e & (cnt - 1), m = e >> lc;
                    acc[q] = cadd(acc[q], cmul(cconj(zl[lpad(t * P + m)]), zl[lpad((cnt + t) * P + m)]));
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < NQG; ++q) {
            const int e = tid + q * T, t = e & (cnt - 1), m = e >> lc, j = group * cnt + t;
            if (e < n && j < Qh) C.acc[((long long)chunk * Qh + j) * P + m] = acc[q];
        }
    } else {   // CC_GK_FINAL
        for (int e = tid; e < n; e += T) {
            const int t = e & (cnt - 1), m = e >> lc, j = group * cnt + t;
            double2 v = make_double2(0.0, 0.0);
            if (j < Qh)
                for (int c = 0; c < C.nchunk; ++c) v = cadd(v, C.acc[((long long)c * Qh + j) * P + m]);
            zl[lpad(t * P + m)] = v;
        }
        __syncthreads();
        lds_fft_inline<kConvPPT, true>(zl, P, cnt, 1.0, kConvOrder, tw);
        store(zl, 0);
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
    // plan-based path (grids beyond the fused kernels)
    std::unique_ptr<Plan> r2c_b, r2c_1, c2r_b, c2r_adj_b, r2c_adj_b, r2c_adj_1;
    DevBuf pad, spec_a, spec_b, kspec, kpad;
    // fused path
    bool fused = false;
    int row_count = 1, row_lc = 0, col_count = 1, col_lc = 0, gk_count = 1, gk_lc = 0, nchunk = 1, chunk = 1;
    DevBuf tq, tp;          // W_Q^j (j < Q), W_P^j (j < P)
    DevBuf sx, sg, sk, kt;  // half spectra [B][H][Qh] (two), kernel rows [KH][Qh], kernel spectrum [Qh][P]
    DevBuf acc;             // kernel-gradient partial sums [nchunk][Qh][P]
};

static int pow2_floor_int(long long v)
{
    int c = 1;
    while (2LL * c <= v) c *= 2;
    return c;
}

static int ilog2_int(int v)
{
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

// Fused path set-up for grids the LDS kernels hold: line counts per workgroup
// (powers of two, at most 8 points per thread), image chunks, tables and
// workspaces.  Leaves c.fused false for larger grids.
static long long smooth7_at_least(long long n)
{
    for (;; ++n) {
        long long r = n;
        for (int p : {2, 3, 5, 7})
            while (r % p == 0) r /= p;
        if (r == 1) return n;
    }
}

static int conv_fused_setup(Conv2d &c, Error &err)
{
    // A circular convolution of period P gives the 'same' rows [cy, cy + H) of the linear one when
    // P >= H + max(cy, KH - 1 - cy): the wrapped terms then only touch rows the crop drops.  The
    // correlations of the backward pass need the same bound (their wrapped terms land outside
    // [0, KH) and [0, H)).  Likewise for the columns.
    const int cy = (c.KH - 1) / 2, cx = (c.KW - 1) / 2;
    const long long P = smooth7_at_least((long long)c.H + std::max(cy, c.KH - 1 - cy));
    const long long M = smooth7_at_least(((long long)c.W + std::max(cx, c.KW - 1 - cx) + 1) / 2), Qh = M + 1;
    // rows of M complex points and, in the kernel-gradient pass, two columns of P points per LDS batch
    if (M > (long long)kConvPPT * kConvRowThreads || 2 * P > kConvColPoints) return OK;
    c.P = P;
    c.Q = 2 * M;
    c.nspec = c.P * Qh;
    c.row_count = std::min(64, pow2_floor_int((long long)kConvPPT * kConvRowThreads / M));
    c.row_lc = ilog2_int(c.row_count);
    c.col_count = std::min(64, pow2_floor_int(kConvColPoints / c.P));
    c.col_lc = ilog2_int(c.col_count);
    // the kernel-gradient pass transforms x and g columns together: 2 count P <= kConvColPoints
    c.gk_count = std::min(32, pow2_floor_int(kConvColPoints / (2 * c.P)));
    c.gk_lc = ilog2_int(c.gk_count);
    const long long ngroups = (Qh + c.col_count - 1) / c.col_count;
    // enough workgroups for four per CU of a 256-CU device, in chunks of images
    c.nchunk = (int)std::max<long long>(1, std::min<long long>(c.B, (4 * 256 + ngroups - 1) / ngroups));
    c.chunk = (c.B + c.nchunk - 1) / c.nchunk;
    c.nchunk = (c.B + c.chunk - 1) / c.chunk;
    std::vector<double2> h;
    twiddle_table(h, (uint64_t)c.Q, (uint64_t)c.Q, 1);
    if (c.tq.alloc(h.size() * sizeof(double2)) ||
        hipMemcpy(c.tq.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        err = {E_NOMEM, "convolution table allocation failed"};
        return E_NOMEM;
    }
    twiddle_table(h, (uint64_t)c.P, (uint64_t)c.P, 1);
    if (c.tp.alloc(h.size() * sizeof(double2)) ||
        hipMemcpy(c.tp.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        err = {E_NOMEM, "convolution table allocation failed"};
        return E_NOMEM;
    }
    const size_t spec = (size_t)c.B * c.H * Qh * sizeof(double2);
    if (c.sx.alloc(spec) || c.sg.alloc(spec) || c.sk.alloc((size_t)c.KH * Qh * sizeof(double2)) ||
        c.kt.alloc((size_t)Qh * c.P * sizeof(double2)) ||
        c.acc.alloc((size_t)c.nchunk * Qh * c.P * sizeof(double2))) {
        err = {E_NOMEM, "convolution workspace allocation failed"};
        return E_NOMEM;
    }
    static std::once_flag attr_once;
    std::call_once(attr_once, [] {
        for (const void *f : {(const void *)k_conv_cols<CC_SPECTRUM>, (const void *)k_conv_cols<CC_CONV>,
                              (const void *)k_conv_cols<CC_GK>, (const void *)k_conv_cols<CC_GK_FINAL>})
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024) != hipSuccess)
                (void)hipGetLastError();
    });
    c.fused = true;
    return OK;
}

static size_t row_lds(const Conv2d &c) { return (size_t)lds_points_padded(c.row_count * (int)(c.Q / 2)) * sizeof(double2); }

static hipError_t rows_fwd(const Conv2d &c, const void *src, bool f32, long long rows, int W, int rx, double2 *S,
                           hipStream_t s)
{
    ConvRows R{src, S, (const double2 *)c.tq.p, rows, W, (int)c.Q, rx, f32 ? 1 : 0, c.row_count, c.row_lc};
    const long long groups = (rows + c.row_count - 1) / c.row_count;
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(groups, 65536));
    hipLaunchKernelGGL(k_conv_rows<kConvRowThreads>, dim3(grid), dim3(kConvRowThreads), row_lds(c), s, R);
    return hipGetLastError();
}

static hipError_t rows_inv(const Conv2d &c, const double2 *S, long long rows, void *dst, bool f32, int Wout, int ox,
                           double scale, hipStream_t s)
{
    ConvRowsInv R{S, dst, (const double2 *)c.tq.p, rows, Wout, (int)c.Q, ox, f32 ? 1 : 0, c.row_count, c.row_lc,
                  scale};
    const long long groups = (rows + c.row_count - 1) / c.row_count;
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(groups, 65536));
    hipLaunchKernelGGL(k_conv_rows_inv<kConvRowThreads>, dim3(grid), dim3(kConvRowThreads), row_lds(c), s, R);
    return hipGetLastError();
}

static hipError_t cols(const Conv2d &c, int mode, ConvCols C, hipStream_t s)
{
    const bool gk = mode == CC_GK;
    C.P = (int)c.P;
    C.Qh = (int)(c.Q / 2 + 1);
    C.count = gk ? c.gk_count : c.col_count;
    C.lc = gk ? c.gk_lc : c.col_lc;
    C.tp = (const double2 *)c.tp.p;
    C.nchunk = c.nchunk;
    C.chunk = c.chunk;
    const int n = C.count * (int)c.P * (gk ? 2 : 1);   // points of one LDS batch
    const unsigned ngroups = (unsigned)((C.Qh + C.count - 1) / C.count);
    const bool chunked = mode == CC_CONV || gk;
    const dim3 grid(chunked ? ngroups * (unsigned)c.nchunk : ngroups), block(kConvColThreads);
    const size_t lds = (size_t)lds_points_padded(n) * sizeof(double2);
    switch (mode) {
        case CC_SPECTRUM: hipLaunchKernelGGL(k_conv_cols<CC_SPECTRUM>, grid, block, lds, s, C); break;
        case CC_CONV: hipLaunchKernelGGL(k_conv_cols<CC_CONV>, grid, block, lds, s, C); break;
        case CC_GK: hipLaunchKernelGGL(k_conv_cols<CC_GK>, grid, block, lds, s, C); break;
        default: hipLaunchKernelGGL(k_conv_cols<CC_GK_FINAL>, grid, block, lds, s, C); break;
    }
    return hipGetLastError();
}

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
    if (conv_fused_setup(*c, err) != OK) return err.code;
    if (c->fused) {
        *out = c.release();
        return OK;
    }
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

// fused path: K^ into kt ([Qh][P]) from the kernel's rows
static hipError_t fused_kernel_spectrum(Conv2d &c, const void *k, hipStream_t s)
{
    hipError_t e = rows_fwd(c, k, c.dtype == 1, c.KH, c.KW, 0, (double2 *)c.sk.p, s);
    if (e != hipSuccess) return e;
    ConvCols C{};
    C.Sa = (const double2 *)c.sk.p;
    C.Ha = c.KH;
    C.ra = 0;
    C.Sout = (double2 *)c.kt.p;
    return cols(c, CC_SPECTRUM, C, s);
}

// y = crop_c(c2r(K^ . r2c(pad_0 x))) / (P Q)
static int fused_forward(Conv2d &c, const void *x, const void *k, void *y, hipStream_t s, Error &err)
{
    const bool f32 = c.dtype == 1;
    const int cy = (c.KH - 1) / 2, cx = (c.KW - 1) / 2;
    IFFT_TRY_HIP(fused_kernel_spectrum(c, k, s));
    IFFT_TRY_HIP(rows_fwd(c, x, f32, (long long)c.B * c.H, c.W, 0, (double2 *)c.sx.p, s));
    ConvCols C{};
    C.Sa = (const double2 *)c.sx.p;
    C.Ha = c.H;
    C.ra = 0;
    C.Sout = (double2 *)c.sx.p;   // in place: a workgroup reads its columns before it writes them
    C.oy = cy;
    C.Ho = c.H;
    C.KT = (const double2 *)c.kt.p;
    C.nimg = c.B;
    C.conjk = 0;
    IFFT_TRY_HIP(cols(c, CC_CONV, C, s));
    IFFT_TRY_HIP(rows_inv(c, (const double2 *)c.sx.p, (long long)c.B * c.H, y, f32, c.W, cx,
                          1.0 / ((double)c.P * c.Q), s));
    return OK;
}

// gx = crop_0(c2r(conj(K^) . G^)) / (P Q), gk = crop_0(c2r(sum_b conj(X^_b) . G^_b)) / (P Q),
// G^_b = r2c(pad_c g_b): the correlations that are the adjoints of the forward map
static int fused_backward(Conv2d &c, const void *x, const void *k, const void *gy, void *gx, void *gk, hipStream_t s,
                          Error &err)
{
    const bool f32 = c.dtype == 1;
    const int cy = (c.KH - 1) / 2, cx = (c.KW - 1) / 2;
    const double inv = 1.0 / ((double)c.P * c.Q);
    if (gx) IFFT_TRY_HIP(fused_kernel_spectrum(c, k, s));   // K^ into kt before the kernel rows' workspace is reused
    IFFT_TRY_HIP(rows_fwd(c, gy, f32, (long long)c.B * c.H, c.W, cx, (double2 *)c.sg.p, s));
    if (gk) {
        // sum_b conj(X^_b) G^_b per chunk of images, then the chunks' sum transformed back
        IFFT_TRY_HIP(rows_fwd(c, x, f32, (long long)c.B * c.H, c.W, 0, (double2 *)c.sx.p, s));
        ConvCols C{};
        C.Sa = (const double2 *)c.sx.p;
        C.Sg = (const double2 *)c.sg.p;
        C.Ha = c.H;
        C.ra = 0;
        C.rg = cy;
        C.acc = (double2 *)c.acc.p;
        C.nimg = c.B;
        IFFT_TRY_HIP(cols(c, CC_GK, C, s));
        ConvCols F{};
        F.acc = (double2 *)c.acc.p;
        F.Sout = (double2 *)c.sk.p;
        F.oy = 0;
        F.Ho = c.KH;
        IFFT_TRY_HIP(cols(c, CC_GK_FINAL, F, s));
        IFFT_TRY_HIP(rows_inv(c, (const double2 *)c.sk.p, c.KH, gk, f32, c.KW, 0, inv, s));
    }
    if (gx) {
        // the convolution pass with conj(K^), G^'s rows at cy, result rows from 0, in place
        ConvCols C{};
        C.Sa = (const double2 *)c.sg.p;
        C.Ha = c.H;
        C.ra = cy;
        C.Sout = (double2 *)c.sg.p;
        C.oy = 0;
        C.Ho = c.H;
        C.KT = (const double2 *)c.kt.p;
        C.nimg = c.B;
        C.conjk = 1;
        IFFT_TRY_HIP(cols(c, CC_CONV, C, s));
        IFFT_TRY_HIP(rows_inv(c, (const double2 *)c.sg.p, (long long)c.B * c.H, gx, f32, c.W, 0, inv, s));
    }
    return OK;
}

int conv2d_forward(Conv2d &c, const void *x, const void *k, void *y, hipStream_t s, Error &err)
{
    if (c.fused) return fused_forward(c, x, k, y, s, err);
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
    if (c.fused) return fused_backward(c, x, k, gy, gx, gk, s, err);
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

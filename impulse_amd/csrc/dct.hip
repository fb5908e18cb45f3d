// DCT and DST of types 1-4 over any set of axes, FP64, with the semantics of
// pocketfft::dct / pocketfft::dst (pocketfft_hdronly.h:3284-3318): unnormalised
// sums, `ortho` applying only the end-point scalings the reference applies
// (the overall 1/sqrt factor is the caller's fct), fct on the first axis.
//
// Each axis of length N runs one c2c transform per line through the engine's
// own FP64 plan over a workspace, between a pre-pass that builds the complex
// sequence from the strided input and a post-pass that forms the outputs.
//
// Types 2-4 at even N (M = N/2 points per line, after Makhoul): with x' the
// input as the DCT sees it -- x'[n] = (-1)^n x[n] for DST-II, x'[n] = x[N-1-n]
// for DST-III / DST-IV (DST-II(x)[k] = DCT-II(x')[N-1-k], DST-III(x)[k] =
// (-1)^k DCT-III(x')[k], DST-IV(x)[k] = (-1)^k DCT-IV(x')[k]) --
//   DCT-II   z[p] = v[2p] + i v[2p+1], v[q] = x'[2q] (q < M), x'[2N-2q-1] (q >= M)
//            forward; V[k] = E[k] + e^{-2 pi i k/N} O[k] from Z[k], conj Z[M-k];
//            W = e^{-i pi k/2N} V[k]; y[k] = 2 Re W, y[N-k] = -2 Im W (k = 0..M)
//   DCT-III  V(k) = e^{i pi k/2N} (x'[k] - i x'[N-k]) (x'[N] = 0), k = 0..M;
//            z[k] = (V(k) + conj V(M-k)) + i e^{2 pi i k/N} (V(k) - conj V(M-k));
//            backward; v[2n] = Re Z[n], v[2n+1] = Im Z[n]; y[2q] = v[q],
//            y[2q+1] = v[N-1-q]
//   DCT-IV   z[p] = (x'[2p] + i x'[N-1-2p]) e^{-i pi p/N}; forward;
//            S = e^{-i pi (4k+1)/4N} Z[k]; y[2k] = 2 Re S, y[N-1-2k] = -2 Im S
// Types 1 and odd N (L points per line):
//   DCT-I    z = even extension of x, L = 2(N-1)  forward   y[k] = Re Z[k]
//   DST-I    z = (0, x, 0, -x~), L = 2(N+1)       forward   y[k] = -Im Z[k+1]
//   DCT-II   z = x, zero-padded, L = 2N           forward   y[k] = 2 Re(e^{-i pi k/2N} Z[k])
//   DST-II   z = x, zero-padded                   forward   y[k] = -2 Im(e^{-i pi (k+1)/2N} Z[k+1])
//   DCT-III  z[n] = x''[n] e^{i pi n/2N}, x''[0] = x[0], x''[n] = 2 x[n]
//                                                 backward  y[k] = Re c[k]
//   DST-III  z[m] = x''[m] e^{i pi m/2N}, m = 1..N, x''[m] = 2 x[m-1] (m < N), x''[N] = x[N-1]
//                                                 backward  y[k] = Im c[k]
//   DCT-IV   z[n] = x[n] e^{-i pi (2n+1)/4N}      forward   y[k] = 2 Re(e^{-i pi k/2N} Z[k])
//   DST-IV   as DCT-IV                            forward   y[k] = -2 Im(e^{-i pi k/2N} Z[k])
// (each follows from writing the cosine / sine as the real / imaginary part of
// one exponential and collecting the DFT kernel).  The half-length forms take
// their pre / post twiddles, and the fused kernel its LDS-stage roots, from
// per-axis tables written by the long-double host generator (one length per
// axis, so one table serves every line); the full-length forms use FP64
// sincospi.
//
// Workspace layout per axis: [lines, L] when the axis is the input's unit-
// stride dimension (threads walk along a line), else [L, lines] (threads walk
// across lines, which keeps the strided gathers and scatters coalesced; the
// c2c then runs over the workspace's strided axis).
#include <cmath>
#include <vector>

#include "butterflies.h"
#include "engine.h"
#include "lds_fft.h"

namespace ifft {

constexpr int kMaxDims = kMaxBatchDims + 1;

// element offsets of the lines along one axis: the other dimensions, last fastest
struct LineMap {
    int nd;
    long long cnt[kMaxDims];
    long long st[kMaxDims];
};

struct R2rParams {
    const double *src;
    double *dst;
    double2 *work;
    LineMap in, out;
    long long sa_in, sa_out;   // element stride along the axis
    long long N, L, lines;
    int cosine, type, ortho;
    int half;                  // types 2-4 at even N: L = N/2 points per line
    int col;                   // workspace [L, lines] (line fastest) instead of [lines, L]
    double fct;
    // half-length lines: the twiddles of the pre / post steps, L + 1 entries each
    // (types 2, 3: e1[k] = exp(i pi 2k / N), e2[k] = exp(i pi k / 2N);
    //  type 4: e1[k] = exp(i pi k / N), e2[k] = exp(i pi (4k + 1) / 4N)),
    // and W_L^j (j < L) for the fused kernel's LDS stages
    const double2 *e1, *e2, *ftw;
};

// (line, point) of work item e and the workspace index of that point
__device__ __forceinline__ void r2r_item(const R2rParams &P, long long e, long long cnt, long long &line, long long &m)
{
    if (P.col) {
        m = e / P.lines;
        line = e - m * P.lines;
    } else {
        line = e / cnt;
        m = e - line * cnt;
    }
}

__device__ __forceinline__ long long r2r_widx(const R2rParams &P, long long line, long long m)
{
    return P.col ? m * P.lines + line : line * P.L + m;
}

__device__ __forceinline__ long long line_off(const LineMap &m, long long line)
{
    long long off = 0;
    for (int d = m.nd - 1; d >= 0; --d) {
        const long long c = m.cnt[d];
        off += (line % c) * m.st[d];
        line /= c;
    }
    return off;
}

// exp(i pi a / b)
__device__ __forceinline__ double2 expipi(double a, double b)
{
    double s, c;
    sincospi(a / b, &s, &c);
    return make_double2(c, s);
}

constexpr double kSqrt2 = 1.4142135623730951;

// input point n as the half-length DCT sees it (DST sign / reversal, type-3 ortho)
template <int TYPE>
__device__ __forceinline__ double r2r_xp(const R2rParams &P, const double *x, long long n)
{
    const long long N = P.N;
    const long long orig = (!P.cosine && TYPE != 2) ? N - 1 - n : n;
    double v = x[orig * P.sa_in];
    if (TYPE == 3 && P.ortho && orig == 0) v *= kSqrt2;
    if (!P.cosine && TYPE == 2 && (n & 1)) v = -v;
    return v;
}

// point m of the half-length complex sequence of one line (types 2-4, even N)
template <int TYPE>
__device__ __forceinline__ double2 half_pre(const R2rParams &P, const double *x, long long m)
{
    const long long N = P.N, M = P.L;
    if constexpr (TYPE == 2) {
        auto v = [&](long long q) {
            return q < M ? r2r_xp<2>(P, x, 2 * q) : r2r_xp<2>(P, x, 2 * N - 2 * q - 1);
        };
        return make_double2(v(2 * m), v(2 * m + 1));
    } else if constexpr (TYPE == 3) {
        auto xe = [&](long long k) { return k < N ? r2r_xp<3>(P, x, k) : 0.0; };
        auto V = [&](long long k) { return cmul(P.e2[k], make_double2(xe(k), -xe(N - k))); };
        const double2 vk = V(m), vm = cconj(V(M - m));
        const double2 sum = cadd(vk, vm);
        const double2 dif = cmul(P.e1[m], csub(vk, vm));
        return make_double2(sum.x - dif.y, sum.y + dif.x);
    } else {
        return cmul(make_double2(r2r_xp<4>(P, x, 2 * m), r2r_xp<4>(P, x, N - 1 - 2 * m)), cconj(P.e1[m]));
    }
}

__global__ void k_r2r_pre(const R2rParams P)
{
    const long long total = P.lines * P.L, N = P.N, L = P.L;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long line, m;
        r2r_item(P, e, L, line, m);
        const double *x = P.src + line_off(P.in, line);
        auto X = [&](long long j) { return x[j * P.sa_in]; };
        double2 z = make_double2(0.0, 0.0);
        if (P.half) {
            z = P.type == 2 ? half_pre<2>(P, x, m) : P.type == 3 ? half_pre<3>(P, x, m) : half_pre<4>(P, x, m);
        } else if (P.cosine) {
            switch (P.type) {
                case 1: {
                    const long long j = m < N ? m : L - m;
                    double v = X(j);
                    if (P.ortho && (j == 0 || j == N - 1)) v *= kSqrt2;
                    z = make_double2(v, 0.0);
                    break;
                }
                case 2:
                    if (m < N) z = make_double2(X(m), 0.0);
                    break;
                case 3:
                    if (m < N) {
                        const double v = m == 0 ? X(0) * (P.ortho ? kSqrt2 : 1.0) : 2.0 * X(m);
                        z = cscale(expipi((double)m, 2.0 * N), v);
                    }
                    break;
                default:
                    if (m < N) z = cscale(cconj(expipi((double)(2 * m + 1), 4.0 * N)), X(m));
            }
        } else {
            switch (P.type) {
                case 1:
                    if (m >= 1 && m <= N) z = make_double2(X(m - 1), 0.0);
                    else if (m >= N + 2) z = make_double2(-X(L - 1 - m), 0.0);
                    break;
                case 2:
                    if (m < N) z = make_double2(X(m), 0.0);
                    break;
                case 3:
                    if (m >= 1 && m <= N) {
                        double v = m < N ? 2.0 * X(m - 1) : X(N - 1);
                        if (P.ortho && m == 1) v *= kSqrt2;   // the reference scales input x[0]
                        z = cscale(expipi((double)m, 2.0 * N), v);
                    }
                    break;
                default:
                    if (m < N) z = cscale(cconj(expipi((double)(2 * m + 1), 4.0 * N)), X(m));
            }
        }
        P.work[r2r_widx(P, line, m)] = z;
    }
}

// output i (as the DCT forms it) of a half-length line: DST reversal / signs, scale
template <int TYPE>
__device__ __forceinline__ void r2r_put(const R2rParams &P, double *y, long long i, double v)
{
    const long long N = P.N;
    long long o = i;
    if (!P.cosine) {
        if (TYPE == 2) o = N - 1 - i;
        else if (i & 1) v = -v;
    }
    y[o * P.sa_out] = v * P.fct;
}

// work items per line of the post-pass
__host__ __device__ inline long long r2r_post_count(int half, int type, long long N)
{
    return half ? (type == 2 ? N / 2 + 1 : N / 2) : N;
}

// outputs of post item k of one half-length line from its transformed points Zat(m)
template <int TYPE, typename ZAT>
__device__ __forceinline__ void half_post(const R2rParams &P, double *yl, long long k, ZAT Zat)
{
    const long long N = P.N, M = P.L;
    if constexpr (TYPE == 2) {
        const double2 zk = Zat(k == M ? 0 : k), zm = cconj(Zat(k == 0 ? 0 : M - k));
        const double2 ev = cscale(cadd(zk, zm), 0.5);
        const double2 d = csub(zk, zm);
        const double2 od = make_double2(0.5 * d.y, -0.5 * d.x);
        const double2 V = cadd(ev, cmul(cconj(P.e1[k]), od));
        const double2 W = cmul(cconj(P.e2[k]), V);
        const long long oi = P.cosine ? 0 : N - 1;   // the ortho end point as the DCT indexes it
        double lo = 2.0 * W.x;
        if (P.ortho && k == oi) lo *= 0.5 * kSqrt2;
        r2r_put<2>(P, yl, k, lo);
        if (k > 0 && k < M) {
            double hi = -2.0 * W.y;
            if (P.ortho && N - k == oi) hi *= 0.5 * kSqrt2;
            r2r_put<2>(P, yl, N - k, hi);
        }
    } else if constexpr (TYPE == 3) {
        const double2 z = Zat(k);
        auto sigma = [&](long long q) { return q < M ? 2 * q : 2 * N - 2 * q - 1; };
        r2r_put<3>(P, yl, sigma(2 * k), z.x);
        r2r_put<3>(P, yl, sigma(2 * k + 1), z.y);
    } else {
        const double2 S = cmul(Zat(k), cconj(P.e2[k]));
        r2r_put<4>(P, yl, 2 * k, 2.0 * S.x);
        r2r_put<4>(P, yl, N - 1 - 2 * k, -2.0 * S.y);
    }
}

__global__ void k_r2r_post(const R2rParams P)
{
    const long long N = P.N;
    const long long cnt = r2r_post_count(P.half, P.type, N), total = P.lines * cnt;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long line, k;
        r2r_item(P, e, cnt, line, k);
        double *yl = P.dst + line_off(P.out, line);
        auto Zat = [&](long long m) { return P.work[r2r_widx(P, line, m)]; };
        if (P.half) {
            if (P.type == 2) half_post<2>(P, yl, k, Zat);
            else if (P.type == 3) half_post<3>(P, yl, k, Zat);
            else half_post<4>(P, yl, k, Zat);
            continue;
        }
        double y;
        if (P.cosine) {
            switch (P.type) {
                case 1:
                    y = Zat(k).x;
                    if (P.ortho && (k == 0 || k == N - 1)) y *= 0.5 * kSqrt2;
                    break;
                case 2:
                    y = 2.0 * cmul(cconj(expipi((double)k, 2.0 * N)), Zat(k)).x;
                    if (P.ortho && k == 0) y *= 0.5 * kSqrt2;
                    break;
                case 3:
                    y = Zat(k).x;
                    break;
                default:
                    y = 2.0 * cmul(cconj(expipi((double)k, 2.0 * N)), Zat(k)).x;
            }
        } else {
            switch (P.type) {
                case 1:
                    y = -Zat(k + 1).y;
                    break;
                case 2:
                    y = -2.0 * cmul(cconj(expipi((double)(k + 1), 2.0 * N)), Zat(k + 1)).y;
                    if (P.ortho && k == 0) y *= 0.5 * kSqrt2;
                    break;
                case 3:
                    y = Zat(k).y;
                    break;
                default:
                    y = -2.0 * cmul(cconj(expipi((double)k, 2.0 * N)), Zat(k)).y;
            }
        }
        yl[k * P.sa_out] = y * P.fct;
    }
}

// Types 2-4 at even N with the N/2-point transform in LDS: `count` lines per
// workgroup (count M <= 8 THREADS points), each loaded, transformed and stored
// by THREADS / count threads, so HBM sees each line read once and written once.
// One instance per type keeps only that type's pre- and post-processing in
// the kernel; three waves per SIMD (168 registers) let three 256-thread
// workgroups share a CU.
constexpr int kR2rFusedMax = 4096;   // longest N/2 the fused kernel takes (512 threads)
#ifndef DCT_FUSED_WPE
#define DCT_FUSED_WPE 3
#endif

// SMALL: a 7-smooth M, transformed by inlined radix-2..8 stages (no call frames in
// scratch memory); other lengths call the per-radix stage functions.
template <int TYPE, int THREADS, bool SMALL>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(DCT_FUSED_WPE)))
k_r2r_fused(const R2rParams P, int count)
{
    extern __shared__ double2 zl[];
    const int M = (int)P.L;
    const int NT = THREADS / count;
    const int t = threadIdx.x / NT, i = threadIdx.x % NT;
    const long long cnt = r2r_post_count(1, TYPE, P.N);
    const double s = TYPE == 3 ? 1.0 : -1.0;
    for (long long l0 = (long long)blockIdx.x * count; l0 < P.lines; l0 += (long long)gridDim.x * count) {
        const long long line = l0 + t;
        const bool valid = line < P.lines;
        if (valid) {
            const double *x = P.src + line_off(P.in, line);
            for (int m = i; m < M; m += NT) zl[lpad(t * M + m)] = half_pre<TYPE>(P, x, m);
        }
        __syncthreads();
        if constexpr (SMALL) lds_fft_inline<kPointsPerThread, true>(zl, M, count, s, 2, TwTable{P.ftw, M});
        else lds_fft(zl, M, count, s, 0, TwTable{P.ftw, M});   // both end with a barrier
        if (valid) {
            double *yl = P.dst + line_off(P.out, line);
            auto Zat = [&](long long m) { return zl[lpad(t * M + (int)m)]; };
            for (long long k = i; k < cnt; k += NT) half_post<TYPE>(P, yl, k, Zat);
        }
        __syncthreads();
    }
}

static bool smooth13_r2r(long long n)
{
    for (int p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}

static bool smooth7_r2r(long long n)
{
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

// workgroup size of the fused kernel for N/2 = M, and lines per workgroup: the
// largest power of two <= 64 with count M <= 8 threads
static int fused_threads(long long M) { return M <= 2048 ? 256 : 512; }
static int fused_count(long long M)
{
    const long long budget = 8LL * fused_threads(M);
    int c = 1;
    while (2 * c <= 64 && 2 * c * M <= budget) c *= 2;
    return c;
}

template <int THREADS, bool SMALL>
static hipError_t launch_fused_ts(int type, const R2rParams &Q, unsigned grid, size_t lds, int count, hipStream_t s)
{
    switch (type) {
        case 2: hipLaunchKernelGGL((k_r2r_fused<2, THREADS, SMALL>), dim3(grid), dim3(THREADS), lds, s, Q, count); break;
        case 3: hipLaunchKernelGGL((k_r2r_fused<3, THREADS, SMALL>), dim3(grid), dim3(THREADS), lds, s, Q, count); break;
        default: hipLaunchKernelGGL((k_r2r_fused<4, THREADS, SMALL>), dim3(grid), dim3(THREADS), lds, s, Q, count); break;
    }
    return hipGetLastError();
}

template <int THREADS>
static hipError_t launch_fused_t(int type, bool small, const R2rParams &Q, unsigned grid, size_t lds, int count,
                                 hipStream_t s)
{
    static std::once_flag attr_once;
    std::call_once(attr_once, [] {
        for (const void *f : {(const void *)k_r2r_fused<2, THREADS, false>, (const void *)k_r2r_fused<3, THREADS, false>,
                              (const void *)k_r2r_fused<4, THREADS, false>, (const void *)k_r2r_fused<2, THREADS, true>,
                              (const void *)k_r2r_fused<3, THREADS, true>, (const void *)k_r2r_fused<4, THREADS, true>})
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024) != hipSuccess)
                (void)hipGetLastError();
    });
    return small ? launch_fused_ts<THREADS, true>(type, Q, grid, lds, count, s)
                 : launch_fused_ts<THREADS, false>(type, Q, grid, lds, count, s);
}

struct R2r {
    int cosine = 1, type = 2, ortho = 0;
    struct Axis {
        int axis;
        long long N, L, lines;
        int forward, half, col;
        int fused;                        // one LDS kernel per line batch (no sub-plan)
        std::shared_ptr<DevBuf> tab;      // half lines: e1 | e2 (L + 1 each) | W_L^j (fused, L)
        LineMap in_first, in_next, out;
        long long sa_in_first, sa_in_next, sa_out;
        std::unique_ptr<Plan> fft;
    };
    std::vector<Axis> ax;
    DevBuf work;
};

void r2r_destroy(R2r *r) { delete r; }

static LineMap line_map(const std::vector<size_t> &shape, const std::vector<ptrdiff_t> &st, size_t axis)
{
    LineMap m{};
    for (size_t d = 0; d < shape.size(); ++d) {
        if (d == axis) continue;
        m.cnt[m.nd] = (long long)shape[d];
        m.st[m.nd] = (long long)(st[d] / (ptrdiff_t)sizeof(double));
        ++m.nd;
    }
    return m;
}

// the twiddle tables of a half-length axis, from the long-double host generator
static int half_tables(R2r::Axis &A, int type, Error &err)
{
    const long long N = A.N, L = A.L;
    std::vector<double2> t((size_t)(2 * (L + 1) + (A.fused ? L : 0)));
    auto cj = [](double2 w) { return make_double2(w.x, -w.y); };   // exp(+i..) from W = exp(-i..)
    for (long long k = 0; k <= L; ++k) {
        if (type == 4) {
            t[(size_t)k] = cj(twiddle_at((uint64_t)k, (uint64_t)(2 * N)));
            t[(size_t)(L + 1 + k)] = cj(twiddle_at((uint64_t)(4 * k + 1), (uint64_t)(8 * N)));
        } else {
            t[(size_t)k] = cj(twiddle_at((uint64_t)k, (uint64_t)N));
            t[(size_t)(L + 1 + k)] = cj(twiddle_at((uint64_t)k, (uint64_t)(4 * N)));
        }
    }
    if (A.fused)
        for (long long j = 0; j < L; ++j) t[(size_t)(2 * (L + 1) + j)] = twiddle_at((uint64_t)j, (uint64_t)L);
    A.tab = std::make_shared<DevBuf>();
    if (A.tab->alloc(t.size() * sizeof(double2))) { err = {E_NOMEM, "DCT/DST table allocation failed"}; return E_NOMEM; }
    if (hipMemcpy(A.tab->p, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        err = {E_HIP, "hipMemcpy of a DCT/DST table failed"};
        return E_HIP;
    }
    return OK;
}

int r2r_plan_create(std::unique_ptr<Plan> &out, int kind, int type, int ortho, int dtype, size_t ndim,
                    const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out, size_t naxes,
                    const size_t *axes, Error &err)
{
    if (kind != 5 && kind != 6) { err = {E_INVALID, "unknown transform kind"}; return E_INVALID; }
    if (type < 1 || type > 4) { err = {E_INVALID, kind == 5 ? "invalid DCT type" : "invalid DST type"}; return E_INVALID; }
    if (dtype != 0) { err = {E_UNSUPPORTED, "only IMPULSE_FFT_F64 is supported"}; return E_UNSUPPORTED; }
    if (ndim < 1 || !shape) { err = {E_INVALID, "ndim must be >= 1"}; return E_INVALID; }
    if (ndim > (size_t)kMaxDims) { err = {E_UNSUPPORTED, "too many dimensions"}; return E_UNSUPPORTED; }
    if (!stride_in || !stride_out) { err = {E_INVALID, "stride dimension mismatch"}; return E_INVALID; }
    if (naxes < 1 || !axes) { err = {E_INVALID, "no axes given"}; return E_INVALID; }
    std::vector<int> seen(ndim, 0);
    for (size_t a = 0; a < naxes; ++a) {
        if (axes[a] >= ndim) { err = {E_INVALID, "bad axis number"}; return E_INVALID; }
        if (++seen[axes[a]] > 1) { err = {E_INVALID, "axis specified repeatedly"}; return E_INVALID; }
    }
    int rc = device_check(err);
    if (rc != OK) return rc;
    auto p = std::make_unique<Plan>();
    Plan &P = *p;
    P.kind = kind;
    P.dtype = dtype;
    P.forward = 1;
    (void)hipGetDevice(&P.device);
    P.shape.assign(shape, shape + ndim);
    P.sin.assign(stride_in, stride_in + ndim);
    P.sout.assign(stride_out, stride_out + ndim);
    P.axes.assign(axes, axes + naxes);
    for (size_t d = 0; d < ndim; ++d)
        if (shape[d] == 0) P.empty = true;
    if (P.empty) {
        out = std::move(p);
        return OK;
    }
    for (size_t d = 0; d < ndim; ++d)
        if (P.sin[d] % 8 || P.sout[d] % 8) {
            err = {E_UNSUPPORTED, "strides must be multiples of the element size"};
            return E_UNSUPPORTED;
        }
    plan_spans(P, 8, 8);
    auto R = std::make_unique<R2r>();
    R->cosine = kind == 5;
    R->type = type;
    R->ortho = ortho ? 1 : 0;
    size_t work = 0;
    for (size_t a = 0; a < naxes; ++a) {
        R2r::Axis A;
        A.axis = (int)axes[a];
        A.N = (long long)shape[A.axis];
        if (R->cosine && type == 1 && A.N < 2) {
            err = {E_INVALID, "DCT-I needs at least 2 points along every axis"};
            return E_INVALID;
        }
        A.half = type != 1 && A.N % 2 == 0;
        A.L = A.half ? A.N / 2
                     : (R->cosine && type == 1) ? 2 * (A.N - 1) : (!R->cosine && type == 1) ? 2 * (A.N + 1) : 2 * A.N;
        A.lines = 1;
        for (size_t d = 0; d < ndim; ++d)
            if ((int)d != A.axis) A.lines *= (long long)shape[d];
        A.forward = type != 3;
        A.in_first = line_map(P.shape, P.sin, A.axis);
        A.in_next = line_map(P.shape, P.sout, A.axis);
        A.out = A.in_next;
        A.sa_in_first = P.sin[A.axis] / 8;
        A.sa_in_next = A.sa_out = P.sout[A.axis] / 8;
        // threads walk along lines when the axis is the input's unit-stride dimension, else across lines
        A.col = A.lines > 1 && std::llabs(a == 0 ? A.sa_in_first : A.sa_in_next) != 1;
        // lines along the unit-stride axis, N/2 13-smooth and within the LDS budget: one fused kernel
        A.fused = A.half && !A.col && A.L <= kR2rFusedMax && smooth13_r2r(A.L);
        if (A.half && (rc = half_tables(A, type, err)) != OK) return rc;
        if (A.fused) {
            R->ax.push_back(std::move(A));
            continue;
        }
        // c2c over the workspace, in place: [lines, L] along axis 1, or [L, lines] along axis 0
        const size_t fshape_r[2] = {(size_t)A.lines, (size_t)A.L}, fshape_c[2] = {(size_t)A.L, (size_t)A.lines};
        const ptrdiff_t fstr_r[2] = {(ptrdiff_t)(A.L * 16), 16}, fstr_c[2] = {(ptrdiff_t)(A.lines * 16), 16};
        const size_t fax_r[1] = {1}, fax_c[1] = {0};
        rc = A.col ? plan_create(A.fft, 0, 0, 2, fshape_c, fstr_c, fstr_c, 1, fax_c, A.forward, err)
                   : plan_create(A.fft, 0, 0, 2, fshape_r, fstr_r, fstr_r, 1, fax_r, A.forward, err);
        if (rc != OK) return rc;
        work = std::max(work, (size_t)(A.lines * A.L) * sizeof(double2));
        R->ax.push_back(std::move(A));
    }
    if (work && R->work.alloc(work)) { err = {E_NOMEM, "DCT/DST workspace allocation failed"}; return E_NOMEM; }
    P.r2r = R.release();
    out = std::move(p);
    return OK;
}

int r2r_num_launches(const R2r &R)
{
    int n = 0;
    for (const auto &A : R.ax) n += A.fused ? 1 : 2 + (int)A.fft->passes.size();
    return n;
}

int r2r_execute(Plan &P, const void *in, void *out, double fct, hipStream_t s, Error &err)
{
    R2r &R = *P.r2r;
    for (size_t a = 0; a < R.ax.size(); ++a) {
        const auto &A = R.ax[a];
        R2rParams Q;
        Q.src = a == 0 ? (const double *)in : (const double *)out;
        Q.dst = (double *)out;
        Q.work = (double2 *)R.work.p;
        Q.in = a == 0 ? A.in_first : A.in_next;
        Q.out = A.out;
        Q.sa_in = a == 0 ? A.sa_in_first : A.sa_in_next;
        Q.sa_out = A.sa_out;
        Q.N = A.N;
        Q.L = A.L;
        Q.lines = A.lines;
        Q.cosine = R.cosine;
        Q.type = R.type;
        Q.ortho = R.ortho;
        Q.half = A.half;
        Q.col = A.col;
        Q.fct = a == 0 ? fct : 1.0;
        const double2 *tb = A.tab ? (const double2 *)A.tab->p : nullptr;
        Q.e1 = tb;
        Q.e2 = tb ? tb + (A.L + 1) : nullptr;
        Q.ftw = tb && A.fused ? tb + 2 * (A.L + 1) : nullptr;
        if (A.fused) {
            const int count = fused_count(A.L);
            const long long groups = (A.lines + count - 1) / count;
            const size_t lds = (size_t)lds_points_padded((int)(count * A.L)) * sizeof(double2);
            const unsigned grid = (unsigned)std::min<long long>(groups, 1 << 20);
            const bool small = smooth7_r2r(A.L);
            hipError_t e = fused_threads(A.L) == 256 ? launch_fused_t<256>(R.type, small, Q, grid, lds, count, s)
                                                     : launch_fused_t<512>(R.type, small, Q, grid, lds, count, s);
            if (e != hipSuccess) { err = {E_HIP, std::string("DCT/DST launch: ") + hipGetErrorString(e)}; return E_HIP; }
            continue;
        }
        const long long pre = A.lines * A.L, post = A.lines * r2r_post_count(A.half, R.type, A.N);
        const unsigned gpre = (unsigned)std::min<long long>((pre + 255) / 256, 1 << 20);
        const unsigned gpost = (unsigned)std::min<long long>((post + 255) / 256, 1 << 20);
        hipLaunchKernelGGL(k_r2r_pre, dim3(gpre), dim3(256), 0, s, Q);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { err = {E_HIP, std::string("DCT/DST launch: ") + hipGetErrorString(e)}; return E_HIP; }
        int rc = plan_execute(*A.fft, R.work.p, R.work.p, 1.0, s, err);
        if (rc != OK) return rc;
        hipLaunchKernelGGL(k_r2r_post, dim3(gpost), dim3(256), 0, s, Q);
        e = hipGetLastError();
        if (e != hipSuccess) { err = {E_HIP, std::string("DCT/DST launch: ") + hipGetErrorString(e)}; return E_HIP; }
    }
    return OK;
}

}  // namespace ifft

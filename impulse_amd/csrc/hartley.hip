// The real-to-real transforms of the FFTPACK family over any set of axes, FP64,
// with the semantics of pocketfft (pocketfft_hdronly.h:3392-3445):
//
//   r2r_fftpack           per axis, real <-> FFTPACK halfcomplex
//                         [r0, r1, i1, r2, i2, ... (, r_{n/2})] (ExecR2R, :3252-3270).
//                         The direction picks the operation: forward writes
//                         r_k + i i_k = X[k] = sum_n x[n] e^{-2 pi i k n / N}; backward
//                         reads a halfcomplex spectrum and writes the unnormalised
//                         inverse x[n] = sum_k X[k] e^{+2 pi i k n / N} of its Hermitian
//                         extension.  real2hermitian only adds sign flips of positions
//                         2, 4, ... (the imaginary slots): of the input before a forward
//                         transform when it is false, of the output after a backward
//                         transform when it is true (the reference passes `forward` as
//                         the plan's direction, :3264)
//   r2r_separable_hartley per axis, H[k] = Re X[k] + Im X[k] (copy_hartley, :3066-3095)
//   r2r_genuine_hartley   over all axes at once: X = the n-D r2c, H[m] = Re X[m] + Im X[m]
//                         and H[-m] = Re X[m] - Im X[m] (:3417-3445); one axis is separable
//
// fct scales once (the first axis, or the whole genuine transform).  Each axis
// runs the engine's own r2c / c2r plan between the strided real array and a
// contiguous half-spectrum workspace [lines, N/2 + 1] (lines = the other
// dimensions, last fastest), and one pass between that workspace and the real
// array: pack to halfcomplex, unpack from it, or fold into the Hartley values.
#include <cmath>
#include <vector>

#include "butterflies.h"
#include "engine.h"
#include "lds_fft.h"

namespace ifft {

constexpr int kRfpMaxDims = kMaxBatchDims + 1;

// real-array element offsets of the lines along one axis (the other dims, last fastest)
struct RfpLines {
    int nd;
    long long cnt[kRfpMaxDims];
    long long st[kRfpMaxDims];
};

__device__ __forceinline__ long long rfp_line_off(const RfpLines &m, long long line)
{
    long long off = 0;
    for (int d = m.nd - 1; d >= 0; --d) {
        const long long c = m.cnt[d];
        off += (line % c) * m.st[d];
        line /= c;
    }
    return off;
}

struct RfpParams {
    const double *rsrc;   // real array read (unpack: halfcomplex input)
    double *rdst;         // real array written (pack / Hartley output)
    double2 *spec;        // workspace [lines, H], H = N/2 + 1
    RfpLines lin, lout;
    long long sa_in, sa_out;
    long long N, H, lines;
    const double2 *ftw;   // fused lines: W_M^j (j < M = N/2), then W_N^k (k <= M/2)
};

// positions 2, 4, ... (the halfcomplex imaginary slots) of every line negated:
// from the strided source into the contiguous real workspace [lines, N]
__global__ void k_rfp_negcopy(const RfpParams P, double *rw)
{
    const long long total = P.lines * P.N, N = P.N;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long line = e / N, i = e - line * N;
        const double v = P.rsrc[rfp_line_off(P.lin, line) + i * P.sa_in];
        rw[e] = (i >= 2 && !(i & 1)) ? -v : v;
    }
}

// the same in place on the output array
__global__ void k_rfp_negate(const RfpParams P)
{
    const long long half = P.N / 2, total = P.lines * half;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long line = e / half, i = 2 * (e - line * half) + 2;
        if (i < P.N) {
            double *y = P.rdst + rfp_line_off(P.lout, line);
            y[i * P.sa_out] = -y[i * P.sa_out];
        }
    }
}

// halfcomplex position of the real part of bin k (k >= 1), its imaginary part follows
__device__ __forceinline__ long long hc_re(long long k) { return 2 * k - 1; }

// workspace bins -> halfcomplex, one thread per output element
__global__ void k_rfp_pack(const RfpParams P)
{
    const long long total = P.lines * P.N, N = P.N;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long line = e / N, i = e - line * N;
        const double2 *X = P.spec + line * P.H;
        double v;
        if (i == 0) v = X[0].x;
        else {
            const long long k = (i + 1) / 2;
            v = (i & 1) ? X[k].x : X[k].y;
        }
        P.rdst[rfp_line_off(P.lout, line) + i * P.sa_out] = v;
    }
}

// halfcomplex -> workspace bins (bin 0 and, for even N, bin N/2 real), one thread per bin
__global__ void k_rfp_unpack(const RfpParams P)
{
    const long long total = P.lines * P.H, N = P.N, H = P.H;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long line = e / H, k = e - line * H;
        const double *x = P.rsrc + rfp_line_off(P.lin, line);
        double2 z;
        if (k == 0) z = make_double2(x[0], 0.0);
        else if (2 * k == N) z = make_double2(x[(N - 1) * P.sa_in], 0.0);
        else {
            z = make_double2(x[hc_re(k) * P.sa_in], x[(hc_re(k) + 1) * P.sa_in]);
        }
        P.spec[e] = z;
    }
}

// workspace bins -> separable Hartley values along the axis, one thread per bin k <= N/2:
// out[k] = Re X[k] + Im X[k], out[N-k] = Re X[k] - Im X[k]
__global__ void k_rfp_hartley(const RfpParams P)
{
    const long long total = P.lines * P.H, N = P.N, H = P.H;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long line = e / H, k = e - line * H;
        const double2 X = P.spec[e];
        double *y = P.rdst + rfp_line_off(P.lout, line);
        if (k == 0 || 2 * k == N) {
            y[k * P.sa_out] = X.x;
        } else {
            y[k * P.sa_out] = X.x + X.y;
            y[(N - k) * P.sa_out] = X.x - X.y;
        }
    }
}

// genuine n-D Hartley: the half spectrum (C-contiguous, shape hshp, the halved axis
// `hax`) folded into the full real array: out[m] = Re + Im, out[rev(m)] = Re - Im,
// rev negating the index along every transformed axis (isax[d] = 1)
struct GenParams {
    const double2 *spec;
    double *dst;
    int nd, hax;
    long long hshp[kRfpMaxDims], shape[kRfpMaxDims], ost[kRfpMaxDims];
    int isax[kRfpMaxDims];
    long long total;
};

__global__ void k_rfp_genuine(const GenParams G)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < G.total;
         e += (long long)gridDim.x * blockDim.x) {
        long long rem = e, off = 0, roff = 0;
        for (int d = G.nd - 1; d >= 0; --d) {
            const long long c = G.hshp[d];
            const long long i = rem % c;
            rem /= c;
            off += i * G.ost[d];
            const long long ri = (G.isax[d] && i) ? G.shape[d] - i : i;
            roff += ri * G.ost[d];
        }
        const double2 X = G.spec[e];
        G.dst[off] = X.x + X.y;
        G.dst[roff] = X.x - X.y;
    }
}

// Forward r2r_fftpack (MODE 0) and the separable Hartley transform (MODE 1) of
// even-length lines along the input's unit-stride axis in one kernel: a line's
// z[m] = x[2m] + i x[2m+1] (M = N/2 points, the sign flips of positions 2, 4,
// ... fused into the load when `flip`) is transformed in LDS, split into
// X[k], X[M-k] = (Z[k] + conj Z[M-k]) / 2 -+ i/2 W_N^{+-k} (Z[k] - conj Z[M-k]),
// and written as halfcomplex or Hartley values; `count` lines per workgroup.
constexpr int kRfpFusedMax = 4096;
// SMALL: a 7-smooth M, transformed by inlined radix-2..8 stages (no call frames in scratch
// memory); other lengths call the per-radix stage functions.
template <int MODE, int THREADS, bool SMALL>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(3)))
k_rfp_fused(const RfpParams P, int count, int flip, double fct)
{
    extern __shared__ double2 zl[];
    const int M = (int)(P.N / 2);
    const long long N = P.N;
    const int NT = THREADS / count;
    const int t = threadIdx.x / NT, i = threadIdx.x % NT;
    for (long long l0 = (long long)blockIdx.x * count; l0 < P.lines; l0 += (long long)gridDim.x * count) {
        const long long line = l0 + t;
        const bool valid = line < P.lines;
        if (valid) {
            const double *x = P.rsrc + rfp_line_off(P.lin, line);
            for (int m = i; m < M; m += NT) {
                double re = x[(2LL * m) * P.sa_in];
                const double im = x[(2LL * m + 1) * P.sa_in];
                if (flip && m > 0) re = -re;
                zl[lpad(t * M + m)] = make_double2(re, im);
            }
        }
        __syncthreads();
        if constexpr (SMALL) lds_fft_inline<kPointsPerThread, true>(zl, M, count, -1.0, 2, TwTable{P.ftw, M});
        else lds_fft(zl, M, count, -1.0, 0, TwTable{P.ftw, M});   // both end with a barrier
        if (valid) {
            double *y = P.rdst + rfp_line_off(P.lout, line);
            const long long sa = P.sa_out;
            for (int k = i; 2 * k <= M; k += NT) {
                const double2 zk = zl[lpad(t * M + k)], zm = zl[lpad(t * M + (k == 0 ? 0 : M - k))];
                const double2 w = P.ftw[M + k];   // W_N^k
                const double2 e1 = cscale(cadd(zk, cconj(zm)), 0.5), d1 = csub(zk, cconj(zm));
                const double2 o1 = make_double2(0.5 * d1.y, -0.5 * d1.x);
                const double2 e2 = cscale(cadd(zm, cconj(zk)), 0.5), d2 = csub(zm, cconj(zk));
                const double2 o2 = make_double2(0.5 * d2.y, -0.5 * d2.x);
                const double2 Xk = cscale(cadd(e1, cmul(o1, w)), fct);
                const double2 Xm = cscale(cadd(e2, cmul(o2, make_double2(-w.x, w.y))), fct);
                auto put = [&](int q, double2 X) {
                    if (MODE == 0) {
                        if (q == 0) y[0] = X.x;
                        else if (q == M) y[(N - 1) * sa] = X.x;
                        else {
                            y[(2LL * q - 1) * sa] = X.x;
                            y[(2LL * q) * sa] = X.y;
                        }
                    } else {
                        if (q == 0 || q == M) y[(long long)q * sa] = X.x;
                        else {
                            y[(long long)q * sa] = X.x + X.y;
                            y[(N - q) * sa] = X.x - X.y;
                        }
                    }
                };
                put(k, Xk);
                if (M - k != k) put(M - k, Xm);
            }
        }
        __syncthreads();
    }
}

static bool rfp_smooth13(long long n)
{
    for (int p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}
static int rfp_fused_threads(long long M) { return M <= 2048 ? 256 : 512; }
static int rfp_fused_count(long long M)
{
    const long long budget = 8LL * rfp_fused_threads(M);
    int c = 1;
    while (2 * c <= 64 && 2 * c * M <= budget) c *= 2;
    return c;
}

template <int THREADS, bool SMALL>
static hipError_t launch_rfp_fused_s(int mode, const RfpParams &Q, unsigned grid, size_t lds, int count, int flip,
                                     double fct, hipStream_t s)
{
    if (mode == 0)
        hipLaunchKernelGGL((k_rfp_fused<0, THREADS, SMALL>), dim3(grid), dim3(THREADS), lds, s, Q, count, flip, fct);
    else
        hipLaunchKernelGGL((k_rfp_fused<1, THREADS, SMALL>), dim3(grid), dim3(THREADS), lds, s, Q, count, flip, fct);
    return hipGetLastError();
}

template <int THREADS>
static hipError_t launch_rfp_fused(int mode, bool small, const RfpParams &Q, unsigned grid, size_t lds, int count,
                                   int flip, double fct, hipStream_t s)
{
    static std::once_flag attr_once;
    std::call_once(attr_once, [] {
        for (const void *f : {(const void *)k_rfp_fused<0, THREADS, false>, (const void *)k_rfp_fused<1, THREADS, false>,
                              (const void *)k_rfp_fused<0, THREADS, true>, (const void *)k_rfp_fused<1, THREADS, true>})
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024) != hipSuccess)
                (void)hipGetLastError();
    });
    return small ? launch_rfp_fused_s<THREADS, true>(mode, Q, grid, lds, count, flip, fct, s)
                 : launch_rfp_fused_s<THREADS, false>(mode, Q, grid, lds, count, flip, fct, s);
}

static bool smooth7_rfp(long long n)
{
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

struct Rfp {
    int kind = 7;                 // 7 r2r_fftpack, 8 separable Hartley, 9 genuine Hartley
    int r2h = 1, forward = 1;
    struct Axis {
        int axis;
        long long N, H, lines;
        RfpLines lin_first, lin_next, lout;
        long long sa_in_first, sa_in_next, sa_out;
        std::unique_ptr<Plan> plan;        // r2c (to the workspace) or c2r (from it)
        bool negcopy = false;              // forward r2r_fftpack without real2hermitian
        bool fused = false;                // one LDS kernel per line batch (no sub-plan)
        std::shared_ptr<DevBuf> tab;       // fused: W_M^j (j < M), then W_N^k (k <= M/2)
    };
    std::vector<Axis> ax;
    std::unique_ptr<Plan> gen;             // genuine: n-D r2c into the workspace
    GenParams gp{};
    DevBuf work, rwork;                    // half spectra; sign-flipped real input
};

void rfp_destroy(Rfp *r) { delete r; }

static RfpLines rfp_lines(const std::vector<size_t> &shape, const std::vector<ptrdiff_t> &st, size_t axis)
{
    RfpLines m{};
    for (size_t d = 0; d < shape.size(); ++d) {
        if (d == axis) continue;
        m.cnt[m.nd] = (long long)shape[d];
        m.st[m.nd] = (long long)(st[d] / (ptrdiff_t)sizeof(double));
        ++m.nd;
    }
    return m;
}

// byte strides of the workspace [lines, H] seen as an array of the full rank
// (the axis with H points, step 16; the other dims C-contiguous around it)
static std::vector<ptrdiff_t> work_strides(const std::vector<size_t> &shape, size_t axis, long long H)
{
    std::vector<ptrdiff_t> st(shape.size());
    ptrdiff_t acc = (ptrdiff_t)(H * 16);
    for (size_t d = shape.size(); d-- > 0;) {
        if (d == axis) continue;
        st[d] = acc;
        acc *= (ptrdiff_t)shape[d];
    }
    st[axis] = 16;
    return st;
}

int rfp_plan_create(std::unique_ptr<Plan> &out, int kind, int r2h, int forward, int dtype, size_t ndim,
                    const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out, size_t naxes,
                    const size_t *axes, Error &err)
{
    if (kind < 7 || kind > 9) { err = {E_INVALID, "unknown transform kind"}; return E_INVALID; }
    if (dtype != 0) { err = {E_UNSUPPORTED, "only IMPULSE_FFT_F64 is supported"}; return E_UNSUPPORTED; }
    if (ndim < 1 || !shape) { err = {E_INVALID, "ndim must be >= 1"}; return E_INVALID; }
    if (ndim > (size_t)kRfpMaxDims) { err = {E_UNSUPPORTED, "too many dimensions"}; return E_UNSUPPORTED; }
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
    P.forward = forward ? 1 : 0;
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
    auto R = std::make_unique<Rfp>();
    R->kind = kind;
    R->r2h = r2h ? 1 : 0;
    R->forward = forward ? 1 : 0;
    const bool genuine = kind == 9 && naxes > 1;
    size_t work = 0, rwork = 0;
    if (genuine) {
        // n-D r2c (forward) into a C-contiguous half spectrum
        std::vector<size_t> hshp(P.shape);
        const size_t hax = axes[naxes - 1];
        hshp[hax] = hshp[hax] / 2 + 1;
        std::vector<ptrdiff_t> hst(ndim);
        ptrdiff_t acc = 16;
        for (size_t d = ndim; d-- > 0;) {
            hst[d] = acc;
            acc *= (ptrdiff_t)hshp[d];
        }
        rc = plan_create(R->gen, 1, 0, ndim, shape, stride_in, hst.data(), naxes, axes, 1, err);
        if (rc != OK) return rc;
        GenParams &G = R->gp;
        G.nd = (int)ndim;
        G.hax = (int)hax;
        G.total = 1;
        for (size_t d = 0; d < ndim; ++d) {
            G.hshp[d] = (long long)hshp[d];
            G.shape[d] = (long long)shape[d];
            G.ost[d] = (long long)(stride_out[d] / 8);
            G.isax[d] = seen[d];
            G.total *= (long long)hshp[d];
        }
        work = (size_t)G.total * 16;
    } else {
        for (size_t a = 0; a < naxes; ++a) {
            Rfp::Axis A;
            A.axis = (int)axes[a];
            A.N = (long long)shape[A.axis];
            A.H = A.N / 2 + 1;
            A.lines = 1;
            for (size_t d = 0; d < ndim; ++d)
                if ((int)d != A.axis) A.lines *= (long long)shape[d];
            A.lin_first = rfp_lines(P.shape, P.sin, A.axis);
            A.lin_next = rfp_lines(P.shape, P.sout, A.axis);
            A.lout = A.lin_next;
            A.sa_in_first = P.sin[A.axis] / 8;
            A.sa_in_next = A.sa_out = P.sout[A.axis] / 8;
            const std::vector<ptrdiff_t> wst = work_strides(P.shape, A.axis, A.H);
            const size_t ax1[1] = {(size_t)A.axis};
            const bool to_spec = !(kind == 7 && !forward);   // forward and Hartley run r2c; backward runs c2r
            // even lines along the input's unit-stride axis, N/2 13-smooth within the LDS budget: fused
            const long long sa_src = a == 0 ? A.sa_in_first : A.sa_in_next;
            A.fused = to_spec && A.N % 2 == 0 && A.N / 2 <= kRfpFusedMax && rfp_smooth13(A.N / 2) &&
                      (sa_src == 1 || A.lines == 1);
            if (A.fused) {
                A.negcopy = false;
                const long long M = A.N / 2;
                std::vector<double2> t((size_t)(M + M / 2 + 1));
                for (long long j = 0; j < M; ++j) t[(size_t)j] = twiddle_at((uint64_t)j, (uint64_t)M);
                for (long long k = 0; k <= M / 2; ++k) t[(size_t)(M + k)] = twiddle_at((uint64_t)k, (uint64_t)A.N);
                A.tab = std::make_shared<DevBuf>();
                if (A.tab->alloc(t.size() * sizeof(double2))) {
                    err = {E_NOMEM, "r2r_fftpack table allocation failed"};
                    return E_NOMEM;
                }
                if (hipMemcpy(A.tab->p, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
                    (void)hipGetLastError();
                    err = {E_HIP, "hipMemcpy of an r2r_fftpack table failed"};
                    return E_HIP;
                }
                R->ax.push_back(std::move(A));
                continue;
            }
            A.negcopy = kind == 7 && forward && !r2h;
            if (A.negcopy) {
                // r2c from the sign-flipped copy [lines, N]
                std::vector<ptrdiff_t> rst = work_strides(P.shape, A.axis, A.N);
                for (auto &v : rst) v /= 2;   // doubles, not complex words
                rc = plan_create(A.plan, 1, 0, ndim, shape, rst.data(), wst.data(), 1, ax1, 1, err);
                rwork = std::max(rwork, (size_t)(A.lines * A.N) * 8);
            } else if (to_spec) {
                // r2c from the real array (input strides for the first axis, output strides after it)
                const ptrdiff_t *src_st = a == 0 ? stride_in : stride_out;
                rc = plan_create(A.plan, 1, 0, ndim, shape, src_st, wst.data(), 1, ax1, 1, err);
            } else {
                // c2r from the workspace into the output array (backward sign)
                rc = plan_create(A.plan, 2, 0, ndim, shape, wst.data(), stride_out, 1, ax1, 0, err);
            }
            if (rc != OK) return rc;
            work = std::max(work, (size_t)(A.lines * A.H) * 16);
            R->ax.push_back(std::move(A));
        }
    }
    if (R->work.alloc(std::max<size_t>(work, 16)) || (rwork && R->rwork.alloc(rwork))) {
        err = {E_NOMEM, "workspace allocation failed"};
        return E_NOMEM;
    }
    P.rfp = R.release();
    out = std::move(p);
    return OK;
}

int rfp_num_launches(const Rfp &R)
{
    if (R.gen) return (int)R.gen->passes.size() + 1;
    int n = 0;
    for (const auto &A : R.ax) n += A.fused ? 1 : 1 + (int)A.plan->passes.size() + (A.negcopy ? 1 : 0);
    return n;
}

static unsigned rfp_grid(long long total)
{
    return (unsigned)std::max<long long>(1, std::min<long long>((total + 255) / 256, 1 << 20));
}

#define RFP_TRY_HIP(expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            err = {E_HIP, std::string("r2r launch failed: ") + hipGetErrorString(e_)};          \
            return E_HIP;                                                                       \
        }                                                                                       \
    } while (0)

int rfp_execute(Plan &P, const void *in, void *out, double fct, hipStream_t s, Error &err)
{
    Rfp &R = *P.rfp;
    double2 *spec = (double2 *)R.work.p;
    if (R.gen) {
        int rc = plan_execute(*R.gen, in, spec, fct, s, err);
        if (rc != OK) return rc;
        GenParams G = R.gp;
        G.spec = spec;
        G.dst = (double *)out;
        hipLaunchKernelGGL(k_rfp_genuine, dim3(rfp_grid(G.total)), dim3(256), 0, s, G);
        RFP_TRY_HIP(hipGetLastError());
        return OK;
    }
    for (size_t a = 0; a < R.ax.size(); ++a) {
        const auto &A = R.ax[a];
        RfpParams Q{};
        Q.spec = spec;
        Q.N = A.N;
        Q.H = A.H;
        Q.lines = A.lines;
        Q.lout = A.lout;
        Q.sa_out = A.sa_out;
        const double f = a == 0 ? fct : 1.0;
        const void *src = a == 0 ? in : out;
        if (R.kind == 7 && !R.forward) {
            // halfcomplex -> bins -> c2r into the output (sign flips after, with real2hermitian)
            Q.rsrc = (const double *)src;
            Q.lin = a == 0 ? A.lin_first : A.lin_next;
            Q.sa_in = a == 0 ? A.sa_in_first : A.sa_in_next;
            hipLaunchKernelGGL(k_rfp_unpack, dim3(rfp_grid(A.lines * A.H)), dim3(256), 0, s, Q);
            RFP_TRY_HIP(hipGetLastError());
            int rc = plan_execute(*A.plan, spec, out, f, s, err);
            if (rc != OK) return rc;
            if (R.r2h) {
                Q.rdst = (double *)out;
                hipLaunchKernelGGL(k_rfp_negate, dim3(rfp_grid(A.lines * (A.N / 2 + 1))), dim3(256), 0, s, Q);
                RFP_TRY_HIP(hipGetLastError());
            }
            continue;
        }
        if (A.fused) {
            const long long M = A.N / 2;
            const int count = rfp_fused_count(M);
            const long long groups = (A.lines + count - 1) / count;
            const size_t lds = (size_t)lds_points_padded((int)(count * M)) * sizeof(double2);
            const unsigned grid = (unsigned)std::min<long long>(groups, 1 << 20);
            Q.rsrc = (const double *)src;
            Q.lin = a == 0 ? A.lin_first : A.lin_next;
            Q.sa_in = a == 0 ? A.sa_in_first : A.sa_in_next;
            Q.rdst = (double *)out;
            Q.ftw = (const double2 *)A.tab->p;
            const int mode = R.kind == 7 ? 0 : 1;
            const int flip = R.kind == 7 && R.forward && !R.r2h;
            const bool small = smooth7_rfp(M);
            hipError_t e = rfp_fused_threads(M) == 256
                               ? launch_rfp_fused<256>(mode, small, Q, grid, lds, count, flip, f, s)
                               : launch_rfp_fused<512>(mode, small, Q, grid, lds, count, flip, f, s);
            RFP_TRY_HIP(e);
            continue;
        }
        const void *rsrc = src;
        if (A.negcopy) {
            Q.rsrc = (const double *)src;
            Q.lin = a == 0 ? A.lin_first : A.lin_next;
            Q.sa_in = a == 0 ? A.sa_in_first : A.sa_in_next;
            hipLaunchKernelGGL(k_rfp_negcopy, dim3(rfp_grid(A.lines * A.N)), dim3(256), 0, s, Q, (double *)R.rwork.p);
            RFP_TRY_HIP(hipGetLastError());
            rsrc = R.rwork.p;
        }
        // the r2c plan of the first axis reads the input strides; later axes read the output
        int rc = plan_execute(*A.plan, rsrc, spec, f, s, err);
        if (rc != OK) return rc;
        Q.rdst = (double *)out;
        if (R.kind == 7)
            hipLaunchKernelGGL(k_rfp_pack, dim3(rfp_grid(A.lines * A.N)), dim3(256), 0, s, Q);
        else
            hipLaunchKernelGGL(k_rfp_hartley, dim3(rfp_grid(A.lines * A.H)), dim3(256), 0, s, Q);
        RFP_TRY_HIP(hipGetLastError());
    }
    return OK;
}

}  // namespace ifft

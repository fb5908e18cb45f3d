// DCT and DST of types 1-4 over any set of axes, FP64, with the semantics of
// pocketfft::dct / pocketfft::dst (pocketfft_hdronly.h:3284-3318): unnormalised
// sums, `ortho` applying only the end-point scalings the reference applies
// (the overall 1/sqrt factor is the caller's fct), fct on the first axis.
//
// Each axis of length N is one c2c transform per line, of length L = 2N (types
// 2-4) or of the even / odd extension (DCT-I: 2(N-1), DST-I: 2(N+1)), run by
// the engine's own FP64 plan over a contiguous workspace, between a pre-pass
// that builds the complex sequence from the strided input and a post-pass that
// forms the outputs:
//   DCT-I    z = even extension of x            forward   y[k] = Re Z[k]
//   DST-I    z = odd extension (0, x, 0, -x~)   forward   y[k] = -Im Z[k+1]
//   DCT-II   z = x, zero-padded                 forward   y[k] = 2 Re(e^{-i pi k/2N} Z[k])
//   DST-II   z = x, zero-padded                 forward   y[k] = -2 Im(e^{-i pi (k+1)/2N} Z[k+1])
//   DCT-III  z[n] = x'[n] e^{i pi n/2N}, x'[0] = x[0], x'[n] = 2 x[n]
//                                               backward  y[k] = Re c[k]
//   DST-III  z[m] = x''[m] e^{i pi m/2N}, m = 1..N, x''[m] = 2 x[m-1] (m < N), x''[N] = x[N-1]
//                                               backward  y[k] = Im c[k]
//   DCT-IV   z[n] = x[n] e^{-i pi (2n+1)/4N}    forward   y[k] = 2 Re(e^{-i pi k/2N} Z[k])
//   DST-IV   as DCT-IV                          forward   y[k] = -2 Im(e^{-i pi k/2N} Z[k])
// (each follows from writing the cosine / sine as the real / imaginary part of
// one exponential and collecting the length-2N DFT kernel).  Twiddles come from
// FP64 sincospi.
#include <cmath>
#include <vector>

#include "butterflies.h"
#include "engine.h"

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
    double fct;
};

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

__global__ void k_r2r_pre(const R2rParams P)
{
    const long long total = P.lines * P.L, N = P.N, L = P.L;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long line = e / L, m = e - line * L;
        const double *x = P.src + line_off(P.in, line);
        auto X = [&](long long j) { return x[j * P.sa_in]; };
        double2 z = make_double2(0.0, 0.0);
        if (P.cosine) {
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
        P.work[e] = z;
    }
}

__global__ void k_r2r_post(const R2rParams P)
{
    const long long total = P.lines * P.N, N = P.N, L = P.L;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long line = e / N, k = e - line * N;
        const double2 *Z = P.work + line * L;
        double y;
        if (P.cosine) {
            switch (P.type) {
                case 1:
                    y = Z[k].x;
                    if (P.ortho && (k == 0 || k == N - 1)) y *= 0.5 * kSqrt2;
                    break;
                case 2:
                    y = 2.0 * cmul(cconj(expipi((double)k, 2.0 * N)), Z[k]).x;
                    if (P.ortho && k == 0) y *= 0.5 * kSqrt2;
                    break;
                case 3:
                    y = Z[k].x;
                    break;
                default:
                    y = 2.0 * cmul(cconj(expipi((double)k, 2.0 * N)), Z[k]).x;
            }
        } else {
            switch (P.type) {
                case 1:
                    y = -Z[k + 1].y;
                    break;
                case 2:
                    y = -2.0 * cmul(cconj(expipi((double)(k + 1), 2.0 * N)), Z[k + 1]).y;
                    if (P.ortho && k == 0) y *= 0.5 * kSqrt2;
                    break;
                case 3:
                    y = Z[k].y;
                    break;
                default:
                    y = -2.0 * cmul(cconj(expipi((double)k, 2.0 * N)), Z[k]).y;
            }
        }
        P.dst[line_off(P.out, line) + k * P.sa_out] = y * P.fct;
    }
}

struct R2r {
    int cosine = 1, type = 2, ortho = 0;
    struct Axis {
        int axis;
        long long N, L, lines;
        int forward;
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
        A.L = (R->cosine && type == 1) ? 2 * (A.N - 1) : (!R->cosine && type == 1) ? 2 * (A.N + 1) : 2 * A.N;
        A.lines = 1;
        for (size_t d = 0; d < ndim; ++d)
            if ((int)d != A.axis) A.lines *= (long long)shape[d];
        A.forward = type != 3;
        A.in_first = line_map(P.shape, P.sin, A.axis);
        A.in_next = line_map(P.shape, P.sout, A.axis);
        A.out = A.in_next;
        A.sa_in_first = P.sin[A.axis] / 8;
        A.sa_in_next = A.sa_out = P.sout[A.axis] / 8;
        // c2c over the workspace [lines, L], in place
        const size_t fshape[2] = {(size_t)A.lines, (size_t)A.L};
        const ptrdiff_t fstr[2] = {(ptrdiff_t)(A.L * 16), 16};
        const size_t fax[1] = {1};
        rc = plan_create(A.fft, 0, 0, 2, fshape, fstr, fstr, 1, fax, A.forward, err);
        if (rc != OK) return rc;
        work = std::max(work, (size_t)(A.lines * A.L) * sizeof(double2));
        R->ax.push_back(std::move(A));
    }
    if (R->work.alloc(work)) { err = {E_NOMEM, "DCT/DST workspace allocation failed"}; return E_NOMEM; }
    P.r2r = R.release();
    out = std::move(p);
    return OK;
}

int r2r_num_launches(const R2r &R)
{
    int n = 0;
    for (const auto &A : R.ax) n += 2 + (int)A.fft->passes.size();
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
        Q.fct = a == 0 ? fct : 1.0;
        const long long pre = A.lines * A.L, post = A.lines * A.N;
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

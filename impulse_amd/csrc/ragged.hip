// Ragged batches: thousands of real transforms of unequal lengths in one
// persistent launch.
//
// Every workgroup takes whole signals from a queue (one vector atomic per
// claim; the host orders the queue by decreasing cost, so the tail of the
// launch is short signals).  For a real signal of length n the workgroup forms
// a complex signal z of length M (M = n/2, z[m] = x[2m] + i x[2m+1] for even n;
// for odd n, M = n and two signals of the same length share one transform as
// z = x + i x2, separated as X = (Z[k] + conj Z[n-k]) / 2 and
// X2 = (Z[k] - conj Z[n-k]) / 2i), transforms it, and writes the n/2+1 bins.  The host plans each signal (task table):
//   M 13-smooth, fits the LDS budget : Stockham stages in LDS;
//   M 13-smooth, longer              : four-step M = L1 L2 through the
//                                      workgroup's global scratch (batches of
//                                      columns, then batches of rows, in LDS);
//   otherwise                        : Bluestein with a 7-smooth n2 >= M+K-1,
//                                      where K = M outputs are needed for even n
//                                      and K = (n+1)/2 for odd n; the chirp
//                                      kernel's spectrum is computed once per
//                                      distinct length when the batch is
//                                      created and kept in a device table, so an
//                                      execute runs two transforms of length n2
//                                      per Bluestein signal.
// Twiddles and the chirp exp(-i pi j^2 / M) are generated in FP64 with
// sincospi (the chirp index j^2 mod 2M in 64-bit integers).  Phases that hand
// data to other threads of the workgroup through global scratch are separated
// by a workgroup barrier.  Input is float32; arithmetic and the
// output bins (complex128) are FP64.
//
// The autotuner chooses the workgroup size, the LDS budget per transform
// (which decides which lengths stay in LDS and how four-step lengths split)
// and the radix order of the LDS stages.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"
#include "fft_common.h"
#include "lds_fft.h"

namespace ifft {

struct RaggedTask {
    long long in_off;    // first float of the signal
    long long out_off;   // first complex bin of the output
    int n;               // real length
    int M;               // complex length (n/2 or n)
    int L;               // transform length: M, or n2 for Bluestein
    int L1, L2;          // four-step split of L (0: L stays in LDS)
    int ca, cb;          // columns per LDS batch in pass A (length L2) / rows in pass B (length L1)
    int blue;            // Bluestein
    long long spec_off;  // Bluestein: first point of the chirp kernel's spectrum in the table
    long long in_off2;   // pair: the second odd signal of the same length (z = x + i x2)
    long long out_off2;
    int pair;
    int pad_;
};

struct RaggedParams {
    const float *x;
    double2 *y;
    const RaggedTask *tasks;
    int ntasks;
    int *counter;
    double2 *scratch;          // per workgroup: 2 buffers of `slot` complex points
    long long slot;
    double2 *spec;             // table of chirp-kernel spectra (unnormalized)
    int mode;                  // 0: transform the signals; 1: fill the spectrum table
    int order;                 // radix order of LDS stages (radix_codes): bit 0 odd radices first, bit 1 radix-8 powers of two
    double fct;                // output scale
    double osign;              // +1: forward bins; -1: conjugated (backward sign)
};

// ---------------------------------------------------------------------------
// shared host/device planning helpers
// ---------------------------------------------------------------------------

__host__ __device__ inline bool rg_smooth13(long long n)
{
    if (n < 1) return false;
    const int ps[6] = {2, 3, 5, 7, 11, 13};
    for (int k = 0; k < 6; ++k)
        while (n % ps[k] == 0) n /= ps[k];
    return n == 1;
}

// transforms of length L per LDS batch: largest power of two c <= 64 with c L <= lds_max
__host__ __device__ inline int rg_batch(long long L, int lds_max)
{
    int c = 1;
    while (2 * c <= 64 && 2 * c * L <= lds_max) c *= 2;
    return c;
}

static long long rg_good7(long long n)
{
    long long best = 1LL << 62;
    for (long long f7 = 1; f7 < best; f7 *= 7)
        for (long long f5 = f7; f5 < best; f5 *= 5)
            for (long long f3 = f5; f3 < best; f3 *= 3) {
                long long v = f3;
                while (v < n) v *= 2;
                if (v < best) best = v;
                if (f3 >= n) break;
            }
    return best;
}

// ---------------------------------------------------------------------------
// device code
// ---------------------------------------------------------------------------

// exp(sign 2 pi i m / L)
__device__ __forceinline__ double2 twid(long long m, long long L, double sign)
{
    double s, c;
    sincospi(2.0 * (double)(m % L) / (double)L, &s, &c);
    return make_double2(c, sign * s);
}

// chirp w[j] = exp(-i pi j^2 / M) for 0 <= j < M < 2^28.  j^2 mod 2M from a
// quotient estimate in FP64 (off by at most one, since j^2 / 2M < 2^27) and
// exact 64-bit products; inv2m = 1 / (2M).
__device__ __forceinline__ double2 chirp(long long j, long long M, double inv2m)
{
    const long long m2 = 2 * M;
    const long long sq = j * j;
    long long r = sq - (long long)((double)sq * inv2m) * m2;
    if (r < 0) r += m2;
    if (r >= m2) r -= m2;
    double s, c;
    sincospi((double)r / (double)M, &s, &c);
    return make_double2(c, -s);
}

// Hand global scratch to the other threads of the workgroup.  A barrier's
// workgroup-scope release/acquire covers global memory too: the waves of a
// workgroup share one CU and its write-through L1 (no threadgroup split).  A
// device-scope fence here would write back and invalidate the whole L2.
__device__ __forceinline__ void wg_sync_global() { __syncthreads(); }

// Complex transform of length T.L (sign s) by the workgroup; load(m) gives
// input point m.  Returns true with the result in LDS (natural order), or
// false with the result in `out` in row layout: out[k2 L1 + k1] = X[k2 + L2 k1]
// (four-step, used for the chirp-kernel spectrum table).  `tmp` is global
// scratch of L points.
template <class LOAD>
__device__ bool wg_fft(LOAD load, const RaggedTask &T, double s, double2 *lds, double2 *tmp, double2 *out,
                       int order);

// Column pass of a four-step transform L = L1 L2: columns n1 (length L2,
// stride L1, point m = n1 + L1 n2 from load(m)) are transformed with sign s in
// LDS batches of ca and written times W_L^(s n1 k2) to tmp[n1 + L1 k2].  The
// batches make the strided accesses ca points wide.
template <class LOAD>
__device__ void col_pass(LOAD load, const RaggedTask &T, double s, double2 *lds, double2 *tmp, int order)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long long L = T.L;
    const int L1 = T.L1, L2 = T.L2, ca = T.ca;
    for (int c0 = 0; c0 < L1; c0 += ca) {
        const int cnt = min(ca, L1 - c0);
        for (int e = tid; e < ca * L2; e += nthr) {
            const int t = e % ca, n2 = e / ca;
            lds[lpad(t * L2 + n2)] = t < cnt ? load((long long)(c0 + t) + (long long)L1 * n2) : make_double2(0.0, 0.0);
        }
        __syncthreads();
        lds_fft(lds, L2, ca, s, order, TwSincos{});
        {
            // t = tid % ca is fixed per thread (ca divides nthr); k2 advances by nthr / ca
            const int t = tid % ca, step = nthr / ca;
            const long long n1 = c0 + t;
            double2 w = twid(n1 * (long long)(tid / ca), L, s);
            const double2 ws = twid(n1 * (long long)step, L, s);
            for (int k2 = tid / ca; k2 < L2; k2 += step) {
                if (t < cnt) tmp[n1 + (long long)L1 * k2] = cmul(lds[lpad(t * L2 + k2)], w);
                w = cmul(w, ws);
            }
        }
        __syncthreads();
    }
    wg_sync_global();
}

// Row pass: every row k2 of tmp (L1 contiguous points) is transformed in
// place with sign s, in LDS batches of cb rows.  With `spec` (Bluestein):
// forward transform, times spec[k2 L1 + k1], inverse transform, times
// W_L^(+a k2) -- the middle of the chirp convolution in one pass.
__device__ void row_pass(const RaggedTask &T, double s, double2 *lds, double2 *tmp, int order, const double2 *spec)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long long L = T.L;
    const int L1 = T.L1, L2 = T.L2, cb = T.cb;
    for (int r0 = 0; r0 < L2; r0 += cb) {
        const int cnt = min(cb, L2 - r0);
        double2 *row = tmp + (long long)r0 * L1;
        for (int e = tid; e < cb * L1; e += nthr) lds[lpad(e)] = e < cnt * L1 ? row[e] : make_double2(0.0, 0.0);
        __syncthreads();
        lds_fft(lds, L1, cb, s, order, TwSincos{});
        if (spec) {
            const double2 *sp = spec + (long long)r0 * L1;
            for (int e = tid; e < cnt * L1; e += nthr) lds[lpad(e)] = cmul(lds[lpad(e)], sp[e]);
            __syncthreads();
            lds_fft(lds, L1, cb, -s, order, TwSincos{});
            for (int e = tid; e < cnt * L1; e += nthr) {
                const int t = e / L1, a = e - t * L1;
                row[e] = cmul(lds[lpad(e)], twid((long long)a * (r0 + t), L, -s));
            }
        } else {
            for (int e = tid; e < cnt * L1; e += nthr) row[e] = lds[lpad(e)];
        }
        __syncthreads();
    }
    wg_sync_global();
}

// Column pass producing final values: columns a of tmp (length L2, stride
// L1) transformed with sign s; emit(k, v) receives c[k], k = a + L1 b.
template <class EMIT>
__device__ void col_pass_out(EMIT emit, const RaggedTask &T, double s, double2 *lds, const double2 *tmp, int order)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int L1 = T.L1, L2 = T.L2, ca = T.ca;
    for (int c0 = 0; c0 < L1; c0 += ca) {
        const int cnt = min(ca, L1 - c0);
        for (int e = tid; e < ca * L2; e += nthr) {
            const int t = e % ca, k2 = e / ca;
            lds[lpad(t * L2 + k2)] = t < cnt ? tmp[(long long)(c0 + t) + (long long)L1 * k2] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        lds_fft(lds, L2, ca, s, order, TwSincos{});
        for (int e = tid; e < ca * L2; e += nthr) {
            const int t = e % ca, b = e / ca;
            if (t < cnt) emit((long long)(c0 + t) + (long long)L1 * b, lds[lpad(t * L2 + b)]);
        }
        __syncthreads();
    }
    wg_sync_global();
}

// Output of a four-step transform left in row layout (tmp[k2 L1 + k1] =
// Z[k2 + L2 k1]) in tiles of 16 x 32, transposed through LDS so that both
// the row reads and the natural-order writes are contiguous.  Even n: the
// half-length split X[k] = E + W_n^k O with Z[M-k] read from the mirrored
// tile, and X[M] at the end; odd n: X[k] = Z[k], k <= n/2.
constexpr int kTR = 16, kTC = 32;
static_assert(2 * kTC * (kTR + 1) <= 2048, "two padded output tiles must fit the smallest LDS budget");
__device__ void post_tiles(const RaggedTask &T, double2 *lds, const double2 *tmp, double2 *y, double2 *y2,
                           bool half, double f, double fo)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int L1 = T.L1, L2 = T.L2, n = T.n;
    const long long M = T.L;
    const bool pair = y2 != nullptr;
    const long long limit = half ? M - 1 : n / 2;
    double2 *lds2 = lds + kTC * (kTR + 1);   // second tile (pairs)
    for (int r = 0; r < L2; r += kTR)
        for (int c = 0; c < L1; c += kTC) {
            for (int e = tid; e < kTR * kTC; e += nthr) {
                const int t = e / kTC, u = e % kTC, k2 = r + t, k1 = c + u;
                if (k2 >= L2 || k1 >= L1) continue;
                const long long k = k2 + (long long)L2 * k1;
                if (k > limit) continue;
                const double2 zk = tmp[(long long)k2 * L1 + k1];
                double2 v = zk;
                if (half || pair) {
                    // Z[(M - k) mod M]: row L2 - k2 reversed, or row 0 reversed
                    const long long pm = k2 == 0 ? (long long)((L1 - k1) % L1)
                                                 : (long long)(L2 - k2) * L1 + (L1 - 1 - k1);
                    const double2 zm = tmp[pm];
                    const double2 e1 = cscale(cadd(zk, cconj(zm)), 0.5), d1 = csub(zk, cconj(zm));
                    const double2 o1 = make_double2(0.5 * d1.y, -0.5 * d1.x);
                    if (half) {
                        v = cadd(e1, cmul(o1, twid32((int)k, n, -1.0)));
                    } else {
                        v = e1;
                        lds2[u * (kTR + 1) + t] = make_double2(f * o1.x, fo * o1.y);
                    }
                }
                lds[u * (kTR + 1) + t] = make_double2(f * v.x, fo * v.y);
            }
            __syncthreads();
            for (int e = tid; e < kTR * kTC; e += nthr) {
                const int t = e % kTR, u = e / kTR, k2 = r + t, k1 = c + u;
                if (k2 >= L2 || k1 >= L1) continue;
                const long long k = k2 + (long long)L2 * k1;
                if (k <= limit) {
                    y[k] = lds[u * (kTR + 1) + t];
                    if (pair) y2[k] = lds2[u * (kTR + 1) + t];
                }
            }
            __syncthreads();
        }
    if (half && tid == 0) {
        const double2 z0 = tmp[0];
        y[M] = make_double2(f * (z0.x - z0.y), 0.0);
    }
}

template <class LOAD>
__device__ bool wg_fft(LOAD load, const RaggedTask &T, double s, double2 *lds, double2 *tmp, double2 *out, int order)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long long L = T.L;
    if (T.L1 == 0) {
        for (long long m = tid; m < L; m += nthr) lds[lpad((int)m)] = load(m);
        __syncthreads();
        lds_fft(lds, (int)L, 1, s, order, TwSincos{});
        return true;
    }
    col_pass(load, T, s, lds, tmp, order);
    row_pass(T, s, lds, tmp, order, nullptr);
    for (long long m = tid; m < L; m += nthr) out[m] = tmp[m];
    wg_sync_global();
    return false;
}

__global__ void k_ragged_reset(int *counter)
{
    if (threadIdx.x == 0) counter[0] = 0;
}

// One instance per workgroup size and register budget: WPE = 4 caps the
// registers at 128 per lane (four waves per SIMD, some spills), WPE = 0 lets
// the compiler take what the workgroup size allows (1024 threads: 128).
template <int MAXT, int WPE>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE ? WPE : 1))) k_ragged(const RaggedParams P)
{
    extern __shared__ double2 lds[];
    __shared__ int s_task;
    const int tid = threadIdx.x, nthr = blockDim.x;
    double2 *S0 = P.scratch + (long long)blockIdx.x * P.slot;   // four-step scratch
    for (;;) {
        if (tid == 0) s_task = atomicAdd(P.counter, 1);
        __syncthreads();
        const int task = __builtin_amdgcn_readfirstlane(s_task);   // uniform: the task record loads as scalars
        __syncthreads();
        if (task >= P.ntasks) break;
        const RaggedTask T = P.tasks[task];
        const long long n = T.n, M = T.M;
        const bool half = (n % 2) == 0, pair = T.pair != 0;
        const float *x = P.x + T.in_off;
        const float *x2 = P.x + (pair ? T.in_off2 : T.in_off);
        double2 *y = P.y + T.out_off;
        double2 *y2 = pair ? P.y + T.out_off2 : nullptr;
        auto zin = [&](long long m) -> double2 {
            return half ? make_double2((double)x[2 * m], (double)x[2 * m + 1])
                        : make_double2((double)x[m], pair ? (double)x2[m] : 0.0);
        };
        const double inv2m = 0.5 / (double)M;
        const double f = P.fct, fo = P.fct * P.osign;
        if (P.mode == 1) {
            // chirp kernel b[m] = conj(w[m]) for 0 <= m < K, conj(w[n2 - m]) for n2 - m < M
            // (indices mod n2, zero between); its unnormalized spectrum goes to the table
            // (row layout when the transform is split)
            const long long n2 = T.L, K = (half || pair) ? M : n / 2 + 1;
            double2 *B = P.spec + T.spec_off;
            auto bload = [&](long long m) -> double2 {
                if (m < K) return cconj(chirp(m, M, inv2m));
                if (m > n2 - M) return cconj(chirp(n2 - m, M, inv2m));
                return make_double2(0.0, 0.0);
            };
            if (wg_fft(bload, T, -1.0, lds, S0, B, P.order))
                for (long long m = tid; m < n2; m += nthr) B[m] = lds[lpad((int)m)];
            __syncthreads();
            continue;
        }
        bool in_lds = true;    // Z in LDS (natural order) or in S0 (natural order)
        if (!T.blue) {
            if (T.L1 == 0) {
                wg_fft(zin, T, -1.0, lds, S0, nullptr, P.order);
            } else {
                // columns, rows in place, then the transposing output pass
                col_pass(zin, T, -1.0, lds, S0, P.order);
                row_pass(T, -1.0, lds, S0, P.order, nullptr);
                post_tiles(T, lds, S0, y, y2, half, f, fo);
                __syncthreads();
                continue;
            }
        } else {
            // A = FFT(z w); C = A B; c = IFFT(C) (unnormalized; 1/n2 applied with the output chirp)
            const long long n2 = T.L;
            const double2 *B = P.spec + T.spec_off;
            auto aload = [&](long long m) -> double2 {
                return m < M ? cmul(zin(m), chirp(m, M, inv2m)) : make_double2(0.0, 0.0);
            };
            if (T.L1 == 0) {
                wg_fft(aload, T, -1.0, lds, S0, nullptr, P.order);
                for (long long m = tid; m < n2; m += nthr) lds[lpad((int)m)] = cmul(lds[lpad((int)m)], B[m]);
                __syncthreads();
                lds_fft(lds, (int)n2, 1, 1.0, P.order, TwSincos{});
            } else {
                col_pass(aload, T, -1.0, lds, S0, P.order);
                row_pass(T, -1.0, lds, S0, P.order, B);
                if (half || pair) {
                    // c in natural order back into S0 (each column rewrites the points it read)
                    auto keep = [&](long long k, double2 v) { S0[k] = v; };
                    col_pass_out(keep, T, 1.0, lds, S0, P.order);
                    in_lds = false;
                } else {
                    const long long K = n / 2 + 1;
                    const double sc = 1.0 / (double)n2;
                    auto emit = [&](long long k, double2 v) {
                        if (k < K) {
                            const double2 z = cscale(cmul(v, chirp(k, M, inv2m)), sc);
                            y[k] = make_double2(f * z.x, fo * z.y);
                        }
                    };
                    col_pass_out(emit, T, 1.0, lds, S0, P.order);
                    continue;
                }
            }
        }
        // Z[k] = FFT_M(z)[k], k < K (Bluestein: c[k] w[k] / n2)
        const double bscale = T.blue ? 1.0 / (double)T.L : 1.0;
        auto Z = [&](long long k) -> double2 {
            const double2 v = in_lds ? lds[lpad((int)k)] : S0[k];
            return T.blue ? cscale(cmul(v, chirp(k, M, inv2m)), bscale) : v;
        };
        if (half) {
            // X[k] = E + W_n^k O with E = (Z[k] + conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / 2i;
            // X[M-k] from the same pair with W_n^{M-k} = -conj(W_n^k)
            for (long long k = tid; k <= M / 2; k += nthr) {
                const long long km = (M - k) % M;
                const double2 zk = Z(k), zm = Z(km);
                const double2 w = twid32((int)k, (int)n, -1.0);
                const double2 e1 = cscale(cadd(zk, cconj(zm)), 0.5), d1 = csub(zk, cconj(zm));
                const double2 o1 = make_double2(0.5 * d1.y, -0.5 * d1.x);
                const double2 x1 = cadd(e1, cmul(o1, w));
                y[k] = make_double2(f * x1.x, fo * x1.y);
                if (M - k != k) {
                    const double2 e2 = cscale(cadd(zm, cconj(zk)), 0.5), d2 = csub(zm, cconj(zk));
                    const double2 o2 = make_double2(0.5 * d2.y, -0.5 * d2.x);
                    const double2 x2 = cadd(e2, cmul(o2, make_double2(-w.x, w.y)));
                    y[M - k] = make_double2(f * x2.x, fo * x2.y);
                }
            }
        } else if (pair) {
            // X[k] = (Z[k] + conj Z[n-k]) / 2, X2[k] = (Z[k] - conj Z[n-k]) / 2i
            for (long long k = tid; k <= n / 2; k += nthr) {
                const double2 zk = Z(k), zm = Z((n - k) % n);
                const double2 e1 = cscale(cadd(zk, cconj(zm)), 0.5), d1 = csub(zk, cconj(zm));
                y[k] = make_double2(f * e1.x, fo * e1.y);
                y2[k] = make_double2(f * 0.5 * d1.y, -fo * 0.5 * d1.x);
            }
        } else {
            for (long long k = tid; k <= n / 2; k += nthr) {
                const double2 v = Z(k);
                y[k] = make_double2(f * v.x, fo * v.y);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct Ragged {
    int device = 0;
    std::vector<size_t> n, in_off, out_off;   // as given
    std::vector<std::pair<int, int>> order;   // queue of (signal, paired signal or -1), decreasing cost
    int forward = 1;
    // configuration
    int threads = 0, lds_max = 0, rorder = 0, waves = 0;
    // derived for the configuration
    std::vector<RaggedTask> tasks;
    long long slot = 0;
    int grid = 0;
    DevBuf dtasks, counter, scratch;
    // chirp-kernel spectra, one per distinct Bluestein length
    std::vector<std::pair<long long, long long>> spec_index;   // (2 n + pair, offset), sorted
    DevBuf spec;
    double flops = 0.0;
    size_t in_elems = 0, out_elems = 0;
    // captured graph for (x, y, fct)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *gx = nullptr;
    void *gy = nullptr;
    double gfct = 0.0;
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~Ragged() { drop_graph(); }
};

static const void *ragged_fn(int threads, int waves)
{
    if (threads == 256)
        return waves == 4 ? (const void *)k_ragged<256, 4>
                          : waves == 3 ? (const void *)k_ragged<256, 3> : (const void *)k_ragged<256, 0>;
    if (threads == 512) return waves == 4 ? (const void *)k_ragged<512, 4> : (const void *)k_ragged<512, 0>;
    return (const void *)k_ragged<1024, 0>;
}

static hipError_t launch_ragged(const RaggedParams &P, int threads, int waves, unsigned grid, size_t lds,
                                hipStream_t s)
{
    const dim3 g(grid), b(threads);
    if (threads == 256 && waves == 4) hipLaunchKernelGGL((k_ragged<256, 4>), g, b, lds, s, P);
    else if (threads == 256 && waves == 3) hipLaunchKernelGGL((k_ragged<256, 3>), g, b, lds, s, P);
    else if (threads == 256) hipLaunchKernelGGL((k_ragged<256, 0>), g, b, lds, s, P);
    else if (threads == 512 && waves == 4) hipLaunchKernelGGL((k_ragged<512, 4>), g, b, lds, s, P);
    else if (threads == 512) hipLaunchKernelGGL((k_ragged<512, 0>), g, b, lds, s, P);
    else hipLaunchKernelGGL((k_ragged<1024, 0>), g, b, lds, s, P);
    return hipGetLastError();
}

// transform length of a signal: M, or the Bluestein length n2 >= M + K - 1
static long long transform_len(long long n, bool pair)
{
    const long long M = (n % 2 == 0) ? n / 2 : n;
    if (rg_smooth13(M)) return M;
    const long long K = (n % 2 == 0 || pair) ? M : n / 2 + 1;
    return rg_good7(M + K - 1);
}

// largest divisor of a 13-smooth L that is <= cap
static long long largest_divisor_le(long long L, long long cap)
{
    std::vector<long long> divs = {1};
    long long r = L;
    for (long long p : {2, 3, 5, 7, 11, 13}) {
        int e = 0;
        while (r % p == 0) { r /= p; ++e; }
        const size_t n0 = divs.size();
        long long pw = 1;
        for (int k = 1; k <= e; ++k) {
            pw *= p;
            for (size_t q = 0; q < n0; ++q)
                if (divs[q] * pw <= cap) divs.push_back(divs[q] * pw);
        }
    }
    return *std::max_element(divs.begin(), divs.end());
}

// Rows of a split Bluestein transform are at most this long: the spectrum
// table is stored in the row layout of this split, whatever the configuration.
constexpr int kBlueRow = 2048;

// Plan one signal for an LDS budget of lds_max points (>= kBlueRow); false if
// no split fits.  Split transforms L = L1 L2 take the longest rows L1 that fit
// (L2 = L / L1 as short as possible), so the column passes move batches of
// ca = lds_max / L2 columns and their strided accesses are wide.
static bool plan_task(long long n, bool pair, int lds_max, RaggedTask &t)
{
    const long long M = (n % 2 == 0) ? n / 2 : n;
    t.n = (int)n;
    t.M = (int)M;
    t.blue = rg_smooth13(M) ? 0 : 1;
    t.spec_off = 0;
    t.pair = pair ? 1 : 0;
    t.pad_ = 0;
    t.in_off2 = t.out_off2 = 0;
    const long long L = transform_len(n, pair);
    if (L > (1LL << 30)) return false;
    t.L = (int)L;
    t.L1 = t.L2 = t.ca = t.cb = 0;
    const long long row_cap = t.blue ? kBlueRow : lds_max;
    if (L <= row_cap) return true;
    const long long L1 = largest_divisor_le(L, row_cap), L2 = L / L1;
    if (L2 > lds_max || L1 < 2) return false;
    t.L1 = (int)L1;
    t.L2 = (int)L2;
    t.ca = rg_batch(L2, lds_max);
    t.cb = rg_batch(L1, lds_max);
    return true;
}

constexpr long long kScratchCap = 16LL << 30;

static int build_tasks(Ragged &R, int threads, int lds_max, int waves, Error &err)
{
    std::vector<RaggedTask> tasks(R.order.size());
    long long slot = 1;
    for (size_t q = 0; q < R.order.size(); ++q) {
        const size_t i = (size_t)R.order[q].first;
        const int j = R.order[q].second;
        RaggedTask &t = tasks[q];
        if (!plan_task((long long)R.n[i], j >= 0, lds_max, t)) {
            err = {E_UNSUPPORTED, "ragged signal of length " + std::to_string(R.n[i]) +
                                      " has no four-step split for this configuration"};
            return E_UNSUPPORTED;
        }
        t.in_off = (long long)R.in_off[i];
        t.out_off = (long long)R.out_off[i];
        if (j >= 0) {
            t.in_off2 = (long long)R.in_off[(size_t)j];
            t.out_off2 = (long long)R.out_off[(size_t)j];
        }
        if (t.blue)
            t.spec_off = std::lower_bound(R.spec_index.begin(), R.spec_index.end(),
                                          std::make_pair(2 * (long long)R.n[i] + (j >= 0), -1LL))->second;
        if (t.L1 || t.blue) slot = std::max<long long>(slot, t.L);
    }
    int ncu = 0, per = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, R.device) != hipSuccess || ncu <= 0)
        ncu = 256;
    const size_t lds = (size_t)lds_points_padded(lds_max) * sizeof(double2);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, ragged_fn(threads, waves), threads, lds) != hipSuccess ||
        per <= 0) {
        (void)hipGetLastError();
        per = 1;
    }
    // persistent grid: every resident workgroup, with the per-workgroup scratch
    // held to kScratchCap in total (fewer resident workgroups past that)
    const long long per_wg = slot * (long long)sizeof(double2);
    long long grid_ll = std::min<long long>((long long)ncu * per, (long long)tasks.size());
    grid_ll = std::max<long long>(1, std::min<long long>(grid_ll, std::max<long long>(ncu, kScratchCap / per_wg)));
    const int grid = (int)grid_ll;
    const size_t sbytes = (size_t)grid * (size_t)per_wg;
    if (sbytes > R.scratch.bytes && R.scratch.alloc(sbytes)) {
        err = {E_NOMEM, "ragged scratch allocation failed"};
        return E_NOMEM;
    }
    if (hipMemcpy(R.dtasks.p, tasks.data(), tasks.size() * sizeof(RaggedTask), hipMemcpyHostToDevice) !=
        hipSuccess) {
        (void)hipGetLastError();
        err = {E_HIP, "ragged task upload failed"};
        return E_HIP;
    }
    R.tasks.swap(tasks);
    R.slot = slot;
    R.grid = grid;
    R.threads = threads;
    R.lds_max = lds_max;
    R.waves = waves;
    R.drop_graph();
    return OK;
}

// Spectrum table: one entry per distinct Bluestein length, filled by the
// persistent kernel in spectrum mode.
static int build_spectra(Ragged &R, Error &err)
{
    std::vector<long long> ns;   // keys 2 n + pair
    for (const auto &u : R.order) {
        const long long n = (long long)R.n[(size_t)u.first];
        const long long M = (n % 2 == 0) ? n / 2 : n;
        if (!rg_smooth13(M)) ns.push_back(2 * n + (u.second >= 0 ? 1 : 0));
    }
    std::sort(ns.begin(), ns.end());
    ns.erase(std::unique(ns.begin(), ns.end()), ns.end());
    R.spec_index.clear();
    if (ns.empty()) return OK;
    std::vector<RaggedTask> st(ns.size());
    long long off = 0, slot = 1;
    for (size_t q = 0; q < ns.size(); ++q) {
        if (!plan_task(ns[q] / 2, (ns[q] & 1) != 0, 8192, st[q])) {
            err = {E_UNSUPPORTED, "no four-step split for the Bluestein length of a signal of length " +
                                      std::to_string(ns[q] / 2)};
            return E_UNSUPPORTED;
        }
        st[q].in_off = st[q].out_off = 0;
        st[q].spec_off = off;
        R.spec_index.push_back({ns[q], off});
        off += st[q].L;
        if (st[q].L1) slot = std::max<long long>(slot, st[q].L);
    }
    // longest first
    std::stable_sort(st.begin(), st.end(), [](const RaggedTask &a, const RaggedTask &b) { return a.L > b.L; });
    if (R.spec.alloc((size_t)off * sizeof(double2))) { err = {E_NOMEM, "spectrum table allocation failed"}; return E_NOMEM; }
    DevBuf dt, ws;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, R.device) != hipSuccess || ncu <= 0)
        ncu = 256;
    const int grid = (int)std::min<long long>(ncu, (long long)st.size());
    if (dt.alloc(st.size() * sizeof(RaggedTask)) || ws.alloc((size_t)grid * (size_t)slot * sizeof(double2))) {
        err = {E_NOMEM, "spectrum workspace allocation failed"};
        return E_NOMEM;
    }
    RaggedParams P{};
    P.tasks = (const RaggedTask *)dt.p;
    P.ntasks = (int)st.size();
    P.counter = (int *)R.counter.p;
    P.scratch = (double2 *)ws.p;
    P.slot = slot;
    P.spec = (double2 *)R.spec.p;
    P.mode = 1;
    P.order = 0;
    P.fct = 1.0;
    P.osign = 1.0;
    hipError_t e = hipMemcpy(dt.p, st.data(), st.size() * sizeof(RaggedTask), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ragged_reset, dim3(1), dim3(64), 0, 0, (int *)R.counter.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = launch_ragged(P, 1024, 0, grid, (size_t)lds_points_padded(8192) * sizeof(double2), 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipGetLastError();
        err = {E_HIP, std::string("spectrum table: ") + hipGetErrorString(e)};
        return E_HIP;
    }
    return OK;
}

int ragged_create(Ragged **out, size_t count, const size_t *lengths, const size_t *in_off, const size_t *out_off,
                  int forward, Error &err)
{
    if (!out) { err = {E_INVALID, "null handle pointer"}; return E_INVALID; }
    *out = nullptr;
    if (!count || !lengths || !in_off || !out_off) { err = {E_INVALID, "empty ragged batch"}; return E_INVALID; }
    if (count > (size_t)(1 << 30)) { err = {E_INVALID, "too many signals"}; return E_INVALID; }
    int rc = device_check(err);
    if (rc != OK) return rc;
    auto R = std::make_unique<Ragged>();
    (void)hipGetDevice(&R->device);
    R->forward = forward ? 1 : 0;
    R->n.assign(lengths, lengths + count);
    R->in_off.assign(in_off, in_off + count);
    R->out_off.assign(out_off, out_off + count);
    for (size_t i = 0; i < count; ++i) {
        const long long n = (long long)lengths[i];
        if (n < 1 || n > (1LL << 28)) { err = {E_INVALID, "ragged lengths must be in [1, 2^28]"}; return E_INVALID; }
        if (n > 1) R->flops += 2.5 * (double)n * std::log2((double)n);
        R->in_elems = std::max<size_t>(R->in_elems, in_off[i] + (size_t)n);
        R->out_elems = std::max<size_t>(R->out_elems, out_off[i] + (size_t)n / 2 + 1);
    }
    // Odd lengths > 1 pair up (two real signals, one complex transform) unless
    // the pair's transform would be longer than the longest unpaired one of the
    // batch: each task runs on one workgroup, so the longest task bounds the launch.
    long long longest = 0;
    for (size_t i = 0; i < count; ++i) longest = std::max(longest, transform_len((long long)lengths[i], false));
    std::vector<std::pair<int, int>> units;
    {
        std::vector<std::pair<long long, int>> odd;
        for (size_t i = 0; i < count; ++i)
            if (lengths[i] % 2 == 1 && lengths[i] > 1) odd.push_back({(long long)lengths[i], (int)i});
            else units.push_back({(int)i, -1});
        std::stable_sort(odd.begin(), odd.end());
        for (size_t q = 0; q < odd.size(); ++q) {
            if (q + 1 < odd.size() && odd[q + 1].first == odd[q].first &&
                transform_len(odd[q].first, true) <= longest) {
                units.push_back({odd[q].second, odd[q + 1].second});
                ++q;
            } else {
                units.push_back({odd[q].second, -1});
            }
        }
    }
    std::vector<std::pair<double, size_t>> cost(units.size());
    for (size_t u = 0; u < units.size(); ++u) {
        const long long n = (long long)lengths[(size_t)units[u].first];
        const long long M = (n % 2 == 0) ? n / 2 : n;
        const long long L = transform_len(n, units[u].second >= 0);
        cost[u] = {-(rg_smooth13(M) ? 1.0 : 2.0) * (double)L * std::log2((double)L + 1.0), u};
    }
    std::stable_sort(cost.begin(), cost.end());
    R->order.resize(units.size());
    for (size_t q = 0; q < units.size(); ++q) R->order[q] = units[cost[q].second];
    if (R->dtasks.alloc(units.size() * sizeof(RaggedTask)) || R->counter.alloc(256)) {
        err = {E_NOMEM, "ragged workspace allocation failed"};
        return E_NOMEM;
    }
    const int inst[6][2] = {{256, 0}, {256, 3}, {256, 4}, {512, 0}, {512, 4}, {1024, 0}};
    for (const auto &v : inst)
        if (hipFuncSetAttribute(ragged_fn(v[0], v[1]),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(lds_points_padded(8192) * sizeof(double2))) != hipSuccess) {
            (void)hipGetLastError();
            err = {E_HIP, "cannot raise the LDS limit of the ragged kernel"};
            return E_HIP;
        }
    rc = build_spectra(*R, err);
    if (rc != OK) return rc;
    rc = build_tasks(*R, 1024, 8192, 0, err);
    if (rc != OK) return rc;
    *out = R.release();
    return OK;
}

void ragged_destroy(Ragged *r) { delete r; }

int ragged_configure(Ragged &R, int threads, int lds_max, int order, int waves, Error &err)
{
    const bool pow2 = lds_max > 0 && (lds_max & (lds_max - 1)) == 0;
    if ((threads != 256 && threads != 512 && threads != 1024) || !pow2 || lds_max < kBlueRow || lds_max > 8192 ||
        lds_max > kPointsPerThread * threads || order < 0 || order > 3 ||
        (waves != 0 && waves != 4 && !(waves == 3 && threads == 256))) {
        err = {E_INVALID, "ragged configuration out of range (threads 256/512/1024; lds_points a power of two in "
                          "[2048, 8192] and at most 8 per thread; radix order 0-3; waves per SIMD 0, 4, or 3 "
                          "with 256 threads)"};
        return E_INVALID;
    }
    if (threads == 1024) waves = 0;   // 1024 threads always leave 128 registers
    R.rorder = order;
    if (threads == R.threads && lds_max == R.lds_max && waves == R.waves && !R.tasks.empty()) {
        R.drop_graph();
        return OK;
    }
    return build_tasks(R, threads, lds_max, waves, err);
}

static hipError_t ragged_launch(Ragged &R, const float *x, double2 *y, double fct, hipStream_t s)
{
    RaggedParams P;
    P.x = x;
    P.y = y;
    P.tasks = (const RaggedTask *)R.dtasks.p;
    P.ntasks = (int)R.tasks.size();
    P.counter = (int *)R.counter.p;
    P.scratch = (double2 *)R.scratch.p;
    P.slot = R.slot;
    P.spec = (double2 *)R.spec.p;
    P.mode = 0;
    P.order = R.rorder;
    P.fct = fct;
    P.osign = R.forward ? 1.0 : -1.0;
    hipLaunchKernelGGL(k_ragged_reset, dim3(1), dim3(64), 0, s, (int *)R.counter.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_ragged(P, R.threads, R.waves, (unsigned)R.grid, (size_t)lds_points_padded(R.lds_max) * sizeof(double2),
                         s);
}

static int check_ptrs(const Ragged &R, const float *x, double2 *y, Error &err)
{
    if (!x || !y) { err = {E_INVALID, "null data pointer"}; return E_INVALID; }
    const char *xa = (const char *)x, *ya = (const char *)y;
    const char *xe = xa + R.in_elems * sizeof(float), *ye = ya + R.out_elems * sizeof(double2);
    if (xa < ye && ya < xe) { err = {E_INVALID, "ragged input and output overlap"}; return E_INVALID; }
    return OK;
}

// One hipGraph launch (reset kernel + persistent kernel) per execute.  The
// graph is captured on first use and re-captured when the pointers, the scale
// or the configuration change.
int ragged_execute(Ragged &R, const float *x, double2 *y, double fct, hipStream_t s, bool graph, Error &err)
{
    int rc = check_ptrs(R, x, y, err);
    if (rc != OK) return rc;
    if (!graph) {
        hipError_t e = ragged_launch(R, x, y, fct, s);
        if (e != hipSuccess) { err = {E_HIP, std::string("ragged launch: ") + hipGetErrorString(e)}; return E_HIP; }
        return OK;
    }
    if (!R.exec || R.gx != x || R.gy != y || R.gfct != fct) {
        R.drop_graph();
        hipStream_t cs;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            err = {E_HIP, "stream creation failed"};
            return E_HIP;
        }
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t el = ragged_launch(R, x, y, fct, cs);
            hipError_t ee = hipStreamEndCapture(cs, &R.graph);
            e = el != hipSuccess ? el : ee;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            R.drop_graph();
            err = {E_HIP, std::string("ragged graph capture: ") + hipGetErrorString(e)};
            return E_HIP;
        }
        R.gx = x;
        R.gy = y;
        R.gfct = fct;
    }
    hipError_t e = hipGraphLaunch(R.exec, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        err = {E_HIP, std::string("ragged graph launch: ") + hipGetErrorString(e)};
        return E_HIP;
    }
    return OK;
}

// Autotuner: time every configuration that can plan the whole batch on the
// caller's buffers and keep the fastest.
int ragged_tune(Ragged &R, const float *x, double2 *y, hipStream_t s, int reps, double *best_ms, Error &err)
{
    int rc = check_ptrs(R, x, y, err);
    if (rc != OK) return rc;
    struct Cfg { int t, l, o, w; };
    const Cfg cands[] = {{1024, 8192, 0, 0}, {1024, 8192, 1, 0}, {1024, 4096, 0, 0}, {512, 4096, 0, 0},
                         {512, 4096, 1, 0}, {512, 4096, 1, 4}, {512, 2048, 1, 0}, {512, 2048, 1, 4},
                         {256, 2048, 0, 0}, {256, 2048, 1, 0}, {256, 2048, 0, 4}, {256, 2048, 1, 4},
                         {256, 2048, 0, 3}, {256, 2048, 1, 3}, {256, 2048, 3, 4}, {256, 2048, 2, 4},
                         {256, 2048, 3, 3}, {512, 4096, 3, 4}};
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        (void)hipGetLastError();
        if (a) (void)hipEventDestroy(a);
        err = {E_HIP, "event creation failed"};
        return E_HIP;
    }
    double best = 1e300;
    Cfg bestc = {R.threads, R.lds_max, R.rorder, R.waves};
    reps = std::max(1, reps);
    rc = OK;
    for (const Cfg &c : cands) {
        Error e2{0, ""};
        if (ragged_configure(R, c.t, c.l, c.o, c.w, e2) != OK) {
            if (e2.code == E_UNSUPPORTED) continue;   // this batch has a length the configuration cannot split
            err = e2;
            rc = e2.code;
            break;
        }
        if ((rc = ragged_execute(R, x, y, 1.0, s, false, err)) != OK) break;   // warm-up
        (void)hipEventRecord(a, s);
        for (int r = 0; r < reps && rc == OK; ++r) rc = ragged_execute(R, x, y, 1.0, s, false, err);
        if (rc != OK) break;
        (void)hipEventRecord(b, s);
        float ms = 0.f;
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
            (void)hipGetLastError();
            err = {E_HIP, "autotuner timing failed"};
            rc = E_HIP;
            break;
        }
        const double per = ms / reps;
        if (per < best) { best = per; bestc = c; }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != OK) return rc;
    if (best_ms) *best_ms = best;
    return ragged_configure(R, bestc.t, bestc.l, bestc.o, bestc.w, err);
}

void ragged_info(const Ragged &R, int *threads, int *lds_max, int *order, int *waves, int *grid, double *flops,
                 size_t *scratch_bytes)
{
    if (threads) *threads = R.threads;
    if (waves) *waves = R.waves;
    if (lds_max) *lds_max = R.lds_max;
    if (order) *order = R.rorder;
    if (grid) *grid = R.grid;
    if (flops) *flops = R.flops;
    if (scratch_bytes) *scratch_bytes = R.scratch.bytes;
}

}  // namespace ifft

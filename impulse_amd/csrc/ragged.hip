// Ragged batches: thousands of real transforms of unequal lengths in one
// hipGraph launch of persistent kernels.
//
// For a real signal of length n the engine forms a complex signal z of length
// M (M = n/2, z[m] = x[2m] + i x[2m+1] for even n; for odd n, M = n and two
// signals of the same length share one transform as z = x + i x2, separated as
// X = (Z[k] + conj Z[n-k]) / 2 and X2 = (Z[k] - conj Z[n-k]) / 2i),
// transforms it, and writes the n/2+1 bins.  The host plans every signal for
// the configuration's LDS budget B (points) as tasks of at most B points:
//   small (M 13-smooth and M <= B, or a Bluestein length n2 <= B): one task,
//     the whole transform in LDS (Stockham stages), the bins written from LDS;
//   big, M 13-smooth: four-step M = L1 L2 through a workspace of the signal's
//     own (not of the workgroup): column tiles (FFTs of length L2 over batches
//     of columns, times W_M^(n1 k2)) write it, and row tiles read it back --
//     bands of rows k2 with their mirror rows L2 - k2, which hold Z[k] and
//     Z[M-k] for every bin k of the band, so the real post-process and the
//     transposed store to natural order happen in the same tile;
//   big Bluestein (n2 = L1 L2 7-smooth, n2 >= M + K - 1, K = M outputs for even
//     n, (n+1)/2 for odd): column tiles of z times the chirp, row tiles (forward
//     FFT, times the chirp kernel's spectrum, inverse FFT, times W_n2^(+a k2)),
//     inverse column tiles (the bins of an odd single signal, or c in natural
//     order for the split), and for even n and pairs split tiles (X from Z[k]
//     and Z[Q-k]).
// The tasks of one phase are independent, so each phase is one persistent
// kernel whose workgroups claim tasks with one atomic each (longest first); the
// kernel boundaries order the phases.  One execute is one hipGraph replay of a
// counter reset and the phase kernels.  The chirp kernel's spectrum depends only
// on n: it is computed once per distinct length when the batch is created, in
// natural order.  Twiddles and the chirp exp(-i pi j^2 / M) are generated in
// FP64 with sincospi (the chirp index j^2 mod 2M in 64-bit integers).  Input is
// float32; arithmetic and the output bins (complex128) are FP64.
//
// The autotuner chooses the workgroup size, the LDS budget B (which decides
// which signals are small and how big ones split into tiles) and the radix
// order of the LDS stages.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"
#include "fft_common.h"
#include "lds_fft.h"
#include "mfma_stage.h"

namespace ifft {

// task kinds; the phase of each is given in brackets
enum RgKind : int {
    RG_SMALL = 0,   // [1] whole transform in LDS
    RG_COLS = 1,    // [1] column tile of a smooth split: load z, FFT, twiddle -> workspace
    RG_BCOLS = 2,   // [1] column tile of a Bluestein split: z times the chirp, FFT, twiddle -> workspace
    RG_ROWS = 3,    // [2] band of row pairs (k2, L2 - k2) of a smooth split -> bins
    RG_SELF = 4,    // [2] the self-paired rows 0 and L2/2 of a smooth split -> bins (and X[M])
    RG_BROWS = 5,   // [2] row tile of a Bluestein split: FFT, times the spectrum, inverse FFT, twiddle, in place
    RG_ICOLS = 6,   // [3] inverse column tile of a Bluestein split -> bins (odd single) or c in natural order
    RG_SPLIT = 7    // [4] Bluestein even / pair: bins from c[k] and c[Q - k]
};
constexpr int kPhases = 4;

struct RaggedTask {
    long long in_off;    // first float of the signal
    long long out_off;   // first complex bin of the output
    long long in_off2;   // pair: the second odd signal of the same length (z = x + i x2)
    long long out_off2;
    long long ws_off;    // split transforms: first point of the signal's workspace
    long long spec_off;  // Bluestein: first point of the chirp kernel's spectrum (natural order)
    int kind;
    int n;               // real length
    int M;               // complex length (n/2 or n)
    int L;               // transform length: M, or n2 for Bluestein
    int L1, L2;          // split L = L1 L2: rows of L1 contiguous points, columns of L2 (0: in LDS)
    int first, count;    // tile: first column / row / row pair / bin, and how many
    int width;           // LDS batch: columns (COLS, BCOLS, ICOLS), row slots (ROWS), rows (BROWS)
    int blue;            // Bluestein
    int pair;
    int pad_;
};

struct RaggedParams {
    const float *x;
    double2 *y;
    const RaggedTask *tasks;
    int ntasks;
    int *counter;              // one counter per queue, 64 ints apart
    int qoff[8], qcnt[8];      // the phase's eight queues (one per XCD): first task and task count
    double2 *ws;               // workspaces of the split transforms
    double2 *scratch;          // spectrum mode: per workgroup, `slot` complex points
    long long slot;
    double2 *spec;             // table of chirp-kernel spectra (unnormalized, natural order)
    const u4 *bx;              // non-null: radix-16 stages on the matrix cores (B operand table of both signs)
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

// Stage functions per kernel instance.  A device function is compiled for the
// register budget of every kernel that calls it, so stage functions shared by
// all instances would get the 1,024-thread instance's 128 registers, and the
// larger radices spill on every call.  The tag KI = RgInst<threads, waves,
// precision> makes the FP64 ones distinct per instance (C4 at 256 threads,
// three waves: 27.8 -> 25.9 ms).
template <int MAXT, int WPE, bool MF>
struct RgInst {};
template <class KI>
struct TwK : TwSincos {};

// A radix-16 Stockham stage with the data movement of rstage<16>, whose 16-point DFTs run as
// exact bf16-slice products on the matrix cores (bx_stage).  The NT threads of a transform are
// whole wavefronts; wavefront w takes the butterfly groups [16 grp, 16 grp + 16) for grp = w and
// w + NT / 64 (N <= 8 NT leaves at most two groups a wavefront).  Lane (g, j) loads inputs
// q = 4 g + e of butterfly 16 grp + j times W_16p^(q k), k = b mod p (FP64), and after the
// product stores output j of butterflies 16 grp + 4 g + r.
template <class KI>
__device__ __noinline__ void rstage_bx16(double2 *lds, int base, int N, int p, int i, int NT, double s,
                                         const u4 *bx)
{
    const int nb = N / 16, NW = NT >> 6, w = i >> 6, g = (i >> 4) & 3, j = i & 15;
    const int G = (nb + 15) >> 4;   // butterfly groups of the transform
    double2 v[2][4];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int b = 16 * (w + NW * u) + j;
        if (b < nb) {
            const int k = b % p;
            double2 t = make_double2(1.0, 0.0), w1 = make_double2(1.0, 0.0);
            if (k) {
                t = twid32(4 * g * k, 16 * p, s);
                w1 = twid32(k, 16 * p, s);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[u][e] = cmul(lds[lpad(base + b + (4 * g + e) * nb)], t);
                t = cmul(t, w1);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[u][e] = make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    const BxConsts C = bx_consts_load(bx, s);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int grp = w + NW * u;
        if (grp >= G) break;   // the same for the whole wavefront
        bx_stage(v[u], C);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int b = 16 * grp + 4 * g + r;
            if (b < nb) {
                const int k = b % p;
                lds[lpad(base + (b - k) * 16 + k + j * p)] = v[u][r];
            }
        }
    }
    __syncthreads();
}

// lds_fft with the radix-16 stages on the matrix cores (the other radices in FP64 as there)
template <class KI>
__device__ __noinline__ void lds_fft_bx(double2 *lds, int L, int count, double s, int order, const u4 *bx)
{
    int ns;
    unsigned long long codes = radix_codes(L, order, ns);
    const int NT = blockDim.x / count;
    const int t = threadIdx.x / NT, i = threadIdx.x % NT;
    const int base = t * L;
    const TwK<KI> tw{};
    int p = 1;
    for (int st = 0; st < ns; ++st, codes >>= 4) {
        switch ((int)(codes & 15)) {
            case 1: rstage<2, nb_rg(2)>(lds, base, L, p, i, NT, s, tw); p *= 2; break;
            case 2: rstage<3, nb_rg(3)>(lds, base, L, p, i, NT, s, tw); p *= 3; break;
            case 3: rstage<4, nb_rg(4)>(lds, base, L, p, i, NT, s, tw); p *= 4; break;
            case 4: rstage<5, nb_rg(5)>(lds, base, L, p, i, NT, s, tw); p *= 5; break;
            case 5: rstage<7, nb_rg(7)>(lds, base, L, p, i, NT, s, tw); p *= 7; break;
            case 6: rstage<8, nb_rg(8)>(lds, base, L, p, i, NT, s, tw); p *= 8; break;
            case 7: rstage<11, nb_rg(11)>(lds, base, L, p, i, NT, s, tw); p *= 11; break;
            case 8: rstage<13, nb_rg(13)>(lds, base, L, p, i, NT, s, tw); p *= 13; break;
            default: rstage_bx16<KI>(lds, base, L, p, i, NT, s, bx); p *= 16; break;
        }
    }
    if (ns == 0) __syncthreads();
}

// The LDS transform of the ragged kernel: `count` transforms of length L (blockDim.x / count
// threads each).  With the matrix-core precision (bx given) and transforms that span whole
// wavefronts, the radix-16 stages run on the matrix cores; otherwise every stage is FP64.
template <class KI>
__device__ __forceinline__ void rg_fft(double2 *lds, int L, int count, double s, int order, const u4 *bx)
{
    if (bx && ((blockDim.x / count) & 63) == 0) lds_fft_bx<KI>(lds, L, count, s, order, bx);
    else lds_fft(lds, L, count, s, order, TwK<KI>{});
}

// Hand global scratch to the other threads of the workgroup.  A barrier's
// workgroup-scope release/acquire covers global memory too: the waves of a
// workgroup share one CU and its write-through L1 (no threadgroup split).  A
// device-scope fence here would write back and invalidate the whole L2.
// (Data handed between workgroups goes through a kernel boundary instead.)
__device__ __forceinline__ void wg_sync_global() { __syncthreads(); }

// One LDS batch of columns [c0, c0 + cnt) (cnt <= width, width divides the
// workgroup) of a split transform L = L1 L2: column n1 holds points
// n1 + L1 n2 (load(m) gives point m).  Each column is transformed with sign s,
// multiplied by W_L^(s n1 k2) and stored at out[n1 + L1 k2].  The strided
// accesses are `width` points wide.
template <class KI, class LOAD>
__device__ void col_batch(LOAD load, long long L, int L1, int L2, int c0, int cnt, int width, double s, double2 *lds,
                          double2 *out, int order, const u4 *bx)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int e = tid; e < width * L2; e += nthr) {
        const int t = e % width, n2 = e / width;
        lds[lpad(t * L2 + n2)] = t < cnt ? load((long long)(c0 + t) + (long long)L1 * n2) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    rg_fft<KI>(lds, L2, width, s, order, bx);
    {
        // t = tid % width is fixed per thread (width divides nthr); k2 advances by nthr / width
        const int t = tid % width, step = nthr / width;
        const long long n1 = c0 + t;
        double2 w = twid(n1 * (long long)(tid / width), L, s);
        const double2 ws = twid(n1 * (long long)step, L, s);
        for (int k2 = tid / width; k2 < L2; k2 += step) {
            if (t < cnt) out[n1 + (long long)L1 * k2] = cmul(lds[lpad(t * L2 + k2)], w);
            w = cmul(w, ws);
        }
    }
    __syncthreads();
}

// Rows [r0, r0 + cnt) (cnt <= width) of `buf` (rows of L1 contiguous points)
// transformed in place with sign s.
template <class KI>
__device__ void row_batch(int L1, int r0, int cnt, int width, double s, double2 *lds, double2 *buf, int order, const u4 *bx)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    double2 *row = buf + (long long)r0 * L1;
    for (int e = tid; e < width * L1; e += nthr) lds[lpad(e)] = e < cnt * L1 ? row[e] : make_double2(0.0, 0.0);
    __syncthreads();
    rg_fft<KI>(lds, L1, width, s, order, bx);
    for (int e = tid; e < cnt * L1; e += nthr) row[e] = lds[lpad(e)];
    __syncthreads();
}

// Bluestein row tile: rows [r0, r0 + cnt) of the workspace (row k2, point k1)
// in place: forward FFT, times the chirp kernel's spectrum (natural order: bin
// k2 + L2 k1, read in runs of cnt bins), inverse FFT, times W_L^(+a k2).
template <class KI>
__device__ void brows_batch(long long L, int L1, int L2, int r0, int cnt, int width, double2 *lds, double2 *ws,
                            const double2 *spec, int order, const u4 *bx)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    double2 *row = ws + (long long)r0 * L1;
    for (int e = tid; e < width * L1; e += nthr) lds[lpad(e)] = e < cnt * L1 ? row[e] : make_double2(0.0, 0.0);
    __syncthreads();
    rg_fft<KI>(lds, L1, width, -1.0, order, bx);
    for (int e = tid; e < cnt * L1; e += nthr) {
        const int t = e % cnt, k1 = e / cnt;
        const int q = lpad(t * L1 + k1);
        lds[q] = cmul(lds[q], spec[(long long)(r0 + t) + (long long)L2 * k1]);
    }
    __syncthreads();
    rg_fft<KI>(lds, L1, width, 1.0, order, bx);
    for (int e = tid; e < cnt * L1; e += nthr) {
        const int t = e / L1, a = e - t * L1;
        row[e] = cmul(lds[lpad(e)], twid((long long)a * (r0 + t), L, 1.0));
    }
    __syncthreads();
}

// Inverse column tile: columns [c0, c0 + cnt) of the workspace (length L2,
// stride L1) transformed with sign +1; emit(k, v) receives c[k], k = a + L1 b,
// for the tile's own columns (so an in-place emit touches only its columns).
template <class KI, class EMIT>
__device__ void icols_batch(EMIT emit, int L1, int L2, int c0, int cnt, int width, double2 *lds, const double2 *ws,
                            int order, const u4 *bx)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int e = tid; e < width * L2; e += nthr) {
        const int t = e % width, k2 = e / width;
        lds[lpad(t * L2 + k2)] = t < cnt ? ws[(long long)(c0 + t) + (long long)L1 * k2] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    rg_fft<KI>(lds, L2, width, 1.0, order, bx);
    for (int e = tid; e < width * L2; e += nthr) {
        const int t = e % width, b = e / width;
        if (t < cnt) emit((long long)(c0 + t) + (long long)L1 * b, lds[lpad(t * L2 + b)]);
    }
    __syncthreads();
}

// Bin k of the signal from Z[k] (zk) and Z[Q - k] (zm), Q the complex length M:
// even n: X[k] = E + W_n^k O with E = (Z[k] + conj Z[M-k]) / 2 and
// O = (Z[k] - conj Z[M-k]) / 2i, for every k <= M (k = M from zk = zm = Z[0]
// gives Re Z[0] - Im Z[0]); odd pair: X[k] = E and
// X2[k] = O for k <= n/2; odd single: X[k] = Z[k] for k <= n/2.  f scales the
// real parts, fo the imaginary parts (fo = -f conjugates).
__device__ __forceinline__ void emit_bin(const RaggedTask &T, long long k, double2 zk, double2 zm, double2 *y,
                                         double2 *y2, double f, double fo)
{
    const int n = T.n;
    if (!(n & 1)) {
        const double2 w = twid32((int)k, n, -1.0);
        const double2 e1 = cscale(cadd(zk, cconj(zm)), 0.5), d1 = csub(zk, cconj(zm));
        const double2 o1 = make_double2(0.5 * d1.y, -0.5 * d1.x);
        const double2 v = cadd(e1, cmul(o1, w));
        y[k] = make_double2(f * v.x, fo * v.y);
    } else if (k <= n / 2) {
        if (T.pair) {
            const double2 e1 = cscale(cadd(zk, cconj(zm)), 0.5), d1 = csub(zk, cconj(zm));
            y[k] = make_double2(f * e1.x, fo * e1.y);
            y2[k] = make_double2(f * 0.5 * d1.y, -fo * 0.5 * d1.x);
        } else {
            y[k] = make_double2(f * zk.x, fo * zk.y);
        }
    }
}

// Row band of a smooth split (rows of L1 points, Z[k2 + L2 k1] at row k2, point
// k1): the rows k2 = r0 .. r0 + R - 1 (slots 0 .. R-1) and their mirrors
// L2 - k2 (slots R .. 2R-1; 1 <= k2 < L2 - k2), transformed in LDS.  Z[M - k]
// of a bin k = k2 + L2 k1 is point L1 - 1 - k1 of the mirror row, so every bin
// of the band's 2R rows is finished here; the stores run over R consecutive
// bins (k2 fastest).
template <class KI>
__device__ void rows_split(const RaggedTask &T, double2 *lds, const double2 *ws, double2 *y, double2 *y2, double f,
                           double fo, int order, const u4 *bx)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int L1 = T.L1, L2 = T.L2, r0 = T.first, R = T.count, W = T.width;
    for (int e = tid; e < W * L1; e += nthr) {
        const int t = e / L1, k1 = e - t * L1;
        double2 v = make_double2(0.0, 0.0);
        if (t < 2 * R) {
            const int row = t < R ? r0 + t : L2 - r0 - (t - R);
            v = ws[(long long)row * L1 + k1];
        }
        lds[lpad(e)] = v;
    }
    __syncthreads();
    rg_fft<KI>(lds, L1, W, -1.0, order, bx);
    for (int e = tid; e < 2 * R * L1; e += nthr) {
        const int t = e % (2 * R), k1 = e / (2 * R);
        const int row = t < R ? r0 + t : L2 - r0 - (t - R);
        const int pt = t < R ? t + R : t - R;
        const double2 zk = lds[lpad(t * L1 + k1)], zm = lds[lpad(pt * L1 + L1 - 1 - k1)];
        emit_bin(T, (long long)row + (long long)L2 * k1, zk, zm, y, y2, f, fo);
    }
    __syncthreads();
}

// The self-paired rows of a smooth split: row 0 (Z[M - L2 k1] is its point
// (L1 - k1) mod L1) in slot 0, and for even L2 row L2/2 (mirror point
// L1 - 1 - k1) in slot 1; for even n also X[M] = Re Z[0] - Im Z[0].
template <class KI>
__device__ void self_split(const RaggedTask &T, double2 *lds, const double2 *ws, double2 *y, double2 *y2, double f,
                           double fo, int order, const u4 *bx)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int L1 = T.L1, L2 = T.L2;
    const bool mid = (L2 & 1) == 0;
    for (int e = tid; e < 2 * L1; e += nthr) {
        const int t = e / L1, k1 = e - t * L1;
        lds[lpad(e)] = (t == 0 || mid) ? ws[(long long)(t ? L2 / 2 : 0) * L1 + k1] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    rg_fft<KI>(lds, L1, 2, -1.0, order, bx);
    for (int e = tid; e < 2 * L1; e += nthr) {
        const int t = e & 1, k1 = e >> 1;
        if (t == 1 && !mid) continue;
        const double2 zk = lds[lpad(t * L1 + k1)];
        const int mk = t == 0 ? (L1 - k1) % L1 : L1 - 1 - k1;
        const double2 zm = lds[lpad(t * L1 + mk)];
        emit_bin(T, (long long)(t ? L2 / 2 : 0) + (long long)L2 * k1, zk, zm, y, y2, f, fo);
    }
    if (!(T.n & 1) && tid == 0) {
        const double2 z0 = lds[lpad(0)];
        y[T.M] = make_double2(f * (z0.x - z0.y), 0.0);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// the kernels
// ---------------------------------------------------------------------------

// the B operand table of the matrix-core radix-16 stages (one wavefront)
__global__ void k_bx_table(u4 *tab) { bx_consts_store(tab); }

__global__ void k_ragged_reset(int *counter)
{
    if (threadIdx.x < 8 * kPhases) counter[64 * threadIdx.x] = 0;
}

// the chirp kernel's spectrum of one Bluestein length into the table, natural
// order (create time; the split path goes through the workgroup's scratch `S0`)
template <class KI>
__device__ void spectrum_task(const RaggedTask &T, const RaggedParams &P, double2 *lds, double2 *S0)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long long n = T.n, M = T.M, n2 = T.L;
    const bool half = (n % 2) == 0, pair = T.pair != 0;
    const long long K = (half || pair) ? M : n / 2 + 1;
    const double inv2m = 0.5 / (double)M;
    double2 *B = P.spec + T.spec_off;
    // chirp kernel b[m] = conj(w[m]) for 0 <= m < K, conj(w[n2 - m]) for n2 - m < M (zero between)
    auto bload = [&](long long m) -> double2 {
        if (m < K) return cconj(chirp(m, M, inv2m));
        if (m > n2 - M) return cconj(chirp(n2 - m, M, inv2m));
        return make_double2(0.0, 0.0);
    };
    if (T.L1 == 0) {
        for (long long m = tid; m < n2; m += nthr) lds[lpad((int)m)] = bload(m);
        __syncthreads();
        lds_fft(lds, (int)n2, 1, -1.0, P.order, TwK<KI>{});
        for (long long m = tid; m < n2; m += nthr) B[m] = lds[lpad((int)m)];
        __syncthreads();
        return;
    }
    // width = columns per batch, count = rows per batch
    for (int c0 = 0; c0 < T.L1; c0 += T.width)
        col_batch<KI>(bload, n2, T.L1, T.L2, c0, min(T.width, T.L1 - c0), T.width, -1.0, lds, S0, P.order, nullptr);
    for (int r0 = 0; r0 < T.L2; r0 += T.count)
        row_batch<KI>(T.L1, r0, min(T.count, T.L2 - r0), T.count, -1.0, lds, S0, P.order, nullptr);
    // row layout (S0[k2 L1 + k1] = bin k2 + L2 k1) to natural order
    for (long long m = tid; m < n2; m += nthr) {
        const long long k2 = m / T.L1, k1 = m - k2 * T.L1;
        B[k2 + (long long)T.L2 * k1] = S0[m];
    }
    __syncthreads();
}

// One instance per workgroup size and register budget: WPE = 4 caps the
// registers at 128 per lane (four waves per SIMD, some spills), WPE = 0 lets
// the compiler take what the workgroup size allows (1024 threads: 128).  One
// launch per phase.  A phase's tasks sit in eight queues, one per XCD, each
// holding whole signals: workgroup b runs on XCD b mod 8 and claims from that
// queue first, so the tiles of one signal run on one XCD at about the same
// time and their short strided runs merge in that XCD's L2.  A workgroup whose
// queue is empty takes tasks from the next queues in turn; every workgroup
// leaves once all eight are empty.
template <int MAXT, int WPE, bool MF>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE ? WPE : 1))) k_ragged(const RaggedParams P)
{
    extern __shared__ double2 lds[];
    __shared__ int s_task;
    const u4 *bx = MF ? P.bx : nullptr;   // the FP64 instances carry no matrix-core stage
    // the matrix-core instances share one set (the 128-register budget): with each instance's
    // own budget their stage functions clobber more of the kernel's registers, and the kernel
    // spills around the calls instead (C4: 32.2 ms shared against 38.5-40.9 ms per instance)
    using KI = RgInst<MF ? 1024 : MAXT, MF ? 0 : WPE, MF>;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int home = (int)(blockIdx.x & 7);
    int visited = 0;   // queues found empty so far (thread 0's count is the one that matters)
    for (;;) {
        if (tid == 0) {
            int t = -1;
            while (visited < 8) {
                const int q = (home + visited) & 7;
                const int k = atomicAdd(P.counter + 64 * q, 1);
                if (k < P.qcnt[q]) {
                    t = P.qoff[q] + k;
                    break;
                }
                ++visited;
            }
            s_task = t;
        }
        __syncthreads();
        const int task = __builtin_amdgcn_readfirstlane(s_task);   // uniform: the task record loads as scalars
        __syncthreads();
        if (task < 0) break;
        const RaggedTask T = P.tasks[task];
        if (P.mode == 1) {
            spectrum_task<KI>(T, P, lds, P.scratch + (long long)blockIdx.x * P.slot);
            continue;
        }
        const long long n = T.n, M = T.M;
        const bool half = (n % 2) == 0, pair = T.pair != 0;
        const float *x = P.x + T.in_off;
        const float *x2 = P.x + (pair ? T.in_off2 : T.in_off);
        double2 *y = P.y + T.out_off;
        double2 *y2 = pair ? P.y + T.out_off2 : nullptr;
        double2 *ws = P.ws + T.ws_off;
        auto zin = [&](long long m) -> double2 {
            return half ? make_double2((double)x[2 * m], (double)x[2 * m + 1])
                        : make_double2((double)x[m], pair ? (double)x2[m] : 0.0);
        };
        const double inv2m = 0.5 / (double)M;
        const double f = P.fct, fo = P.fct * P.osign;
        switch (T.kind) {
            case RG_COLS:
                col_batch<KI>(zin, T.L, T.L1, T.L2, T.first, T.count, T.width, -1.0, lds, ws, P.order, bx);
                break;
            case RG_BCOLS: {
                auto aload = [&](long long m) -> double2 {
                    return m < M ? cmul(zin(m), chirp(m, M, inv2m)) : make_double2(0.0, 0.0);
                };
                col_batch<KI>(aload, T.L, T.L1, T.L2, T.first, T.count, T.width, -1.0, lds, ws, P.order, bx);
                break;
            }
            case RG_ROWS: rows_split<KI>(T, lds, ws, y, y2, f, fo, P.order, bx); break;
            case RG_SELF: self_split<KI>(T, lds, ws, y, y2, f, fo, P.order, bx); break;
            case RG_BROWS:
                brows_batch<KI>(T.L, T.L1, T.L2, T.first, T.count, T.width, lds, ws, P.spec + T.spec_off, P.order, bx);
                break;
            case RG_ICOLS:
                if (half || pair) {
                    // c[k] for k < M in natural order, in place: each column rewrites points it read
                    auto keep = [&](long long k, double2 v) {
                        if (k < M) ws[k] = v;
                    };
                    icols_batch<KI>(keep, T.L1, T.L2, T.first, T.count, T.width, lds, ws, P.order, bx);
                } else {
                    const long long K = n / 2 + 1;
                    const double sc = 1.0 / (double)T.L;
                    auto emit = [&](long long k, double2 v) {
                        if (k < K) {
                            const double2 z = cscale(cmul(v, chirp(k, M, inv2m)), sc);
                            y[k] = make_double2(f * z.x, fo * z.y);
                        }
                    };
                    icols_batch<KI>(emit, T.L1, T.L2, T.first, T.count, T.width, lds, ws, P.order, bx);
                }
                break;
            case RG_SPLIT: {
                // Z[k] = c[k] w[k] / n2; bins k in [first, first + count) with their mirrors
                const double sc = 1.0 / (double)T.L;
                auto Z = [&](long long k) -> double2 { return cscale(cmul(ws[k], chirp(k, M, inv2m)), sc); };
                for (long long k = T.first + tid; k < (long long)T.first + T.count; k += nthr) {
                    const long long km = (M - k) % M;
                    const double2 zk = Z(k), zm = Z(km);
                    emit_bin(T, k, zk, zm, y, y2, f, fo);
                    if (half && M - k != k) emit_bin(T, M - k, zm, zk, y, y2, f, fo);
                }
                __syncthreads();
                break;
            }
            default: {   // RG_SMALL: the whole transform in LDS
                const long long L = T.L;
                if (!T.blue) {
                    for (long long m = tid; m < L; m += nthr) lds[lpad((int)m)] = zin(m);
                    __syncthreads();
                    rg_fft<KI>(lds, (int)L, 1, -1.0, P.order, bx);
                } else {
                    // A = FFT(z w); C = A B; c = IFFT(C) (unnormalized; 1/n2 applied with the output chirp)
                    const double2 *B = P.spec + T.spec_off;
                    for (long long m = tid; m < L; m += nthr)
                        lds[lpad((int)m)] = m < M ? cmul(zin(m), chirp(m, M, inv2m)) : make_double2(0.0, 0.0);
                    __syncthreads();
                    rg_fft<KI>(lds, (int)L, 1, -1.0, P.order, bx);
                    for (long long m = tid; m < L; m += nthr) lds[lpad((int)m)] = cmul(lds[lpad((int)m)], B[m]);
                    __syncthreads();
                    rg_fft<KI>(lds, (int)L, 1, 1.0, P.order, bx);
                }
                const double bscale = T.blue ? 1.0 / (double)L : 1.0;
                auto Zl = [&](long long k) -> double2 {
                    const double2 v = lds[lpad((int)k)];
                    return T.blue ? cscale(cmul(v, chirp(k, M, inv2m)), bscale) : v;
                };
                const long long kend = half ? M / 2 : n / 2;
                for (long long k = tid; k <= kend; k += nthr) {
                    const long long km = (M - k) % M;
                    const double2 zk = Zl(k), zm = Zl(km);
                    emit_bin(T, k, zk, zm, y, y2, f, fo);
                    if (half && M - k != k) emit_bin(T, M - k, zm, zk, y, y2, f, fo);
                }
                __syncthreads();
                break;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct Ragged {
    int device = 0;
    std::vector<size_t> n, in_off, out_off;   // as given
    std::vector<std::pair<int, int>> units;   // (signal, paired signal or -1)
    int forward = 1;
    // configuration
    int threads = 0, lds_max = 0, rorder = 0, waves = 0;
    int precision = 0;   // 0: FP64 stages; 1: radix-16 stages as exact bf16-slice MFMA products
    DevBuf bx;           // their B operand table (precision 1)
    // derived for the configuration: the tasks of the phases back to back
    std::vector<RaggedTask> tasks;
    int phase_off[kPhases] = {0, 0, 0, 0}, phase_cnt[kPhases] = {0, 0, 0, 0}, phase_grid[kPhases] = {0, 0, 0, 0};
    int queue_off[kPhases][8] = {}, queue_cnt[kPhases][8] = {};   // per phase: the XCD queues within it
    DevBuf dtasks, counter, ws;
    // chirp-kernel spectra, one per distinct Bluestein length, natural order
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

template <bool MF>
static const void *ragged_fn_t(int threads, int waves)
{
    if (threads == 256)
        return waves == 4 ? (const void *)k_ragged<256, 4, MF>
                          : waves == 3 ? (const void *)k_ragged<256, 3, MF> : (const void *)k_ragged<256, 0, MF>;
    if (threads == 512) return waves == 4 ? (const void *)k_ragged<512, 4, MF> : (const void *)k_ragged<512, 0, MF>;
    return (const void *)k_ragged<1024, 0, MF>;
}

static const void *ragged_fn(int threads, int waves, bool mf)
{
    return mf ? ragged_fn_t<true>(threads, waves) : ragged_fn_t<false>(threads, waves);
}

template <bool MF>
static void launch_ragged_t(const RaggedParams &P, int threads, int waves, dim3 g, size_t lds, hipStream_t s)
{
    const dim3 b(threads);
    if (threads == 256 && waves == 4) hipLaunchKernelGGL((k_ragged<256, 4, MF>), g, b, lds, s, P);
    else if (threads == 256 && waves == 3) hipLaunchKernelGGL((k_ragged<256, 3, MF>), g, b, lds, s, P);
    else if (threads == 256) hipLaunchKernelGGL((k_ragged<256, 0, MF>), g, b, lds, s, P);
    else if (threads == 512 && waves == 4) hipLaunchKernelGGL((k_ragged<512, 4, MF>), g, b, lds, s, P);
    else if (threads == 512) hipLaunchKernelGGL((k_ragged<512, 0, MF>), g, b, lds, s, P);
    else hipLaunchKernelGGL((k_ragged<1024, 0, MF>), g, b, lds, s, P);
}

// the matrix-core instances when P.bx is set
static hipError_t launch_ragged(const RaggedParams &P, int threads, int waves, unsigned grid, size_t lds,
                                hipStream_t s)
{
    if (P.bx) launch_ragged_t<true>(P, threads, waves, dim3(grid), lds, s);
    else launch_ragged_t<false>(P, threads, waves, dim3(grid), lds, s);
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

// divisors of a 13-smooth L
static std::vector<long long> divisors(long long L)
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
            for (size_t q = 0; q < n0; ++q) divs.push_back(divs[q] * pw);
        }
    }
    return divs;
}

static int pow2_floor(long long v)
{
    int c = 1;
    while (2LL * c <= v) c *= 2;
    return c;
}

// Split L = L1 L2 (rows of L1 points, columns of L2) for tiles of at most B
// points.  A column tile holds `ca` columns (its strided reads and writes are
// ca points wide); a row tile holds `cb` rows, or for a smooth split (row_pairs)
// cb / 2 rows and their mirrors (its stores are cb / 2 bins wide).  The split
// with the widest narrow side wins, then the one with the fewest tiles.
static bool split_tiles(long long L, int B, bool row_pairs, int &L1, int &L2, int &ca, int &cb)
{
    int best = 0;
    long long best_tiles = 0;
    for (long long d : divisors(L)) {
        const long long e = L / d;   // d = L1 (row length), e = L2 (column length)
        if (d < 2 || e < 2 || e > B || (row_pairs ? 2 * d : d) > B) continue;
        const int c = std::min(64, pow2_floor(B / e));
        const int r = std::min(64, pow2_floor(B / d));   // row slots
        const int width = std::min(c, r);
        const long long tiles = (d + c - 1) / c + (e + r - 1) / r;
        if (width > best || (width == best && tiles < best_tiles)) {
            best = width;
            best_tiles = tiles;
            L1 = (int)d;
            L2 = (int)e;
            ca = c;
            cb = r;
        }
    }
    return best > 0;
}

// Tasks of one unit (a signal, or an odd pair) for an LDS budget of B points,
// appended to the phase lists; `ws` counts the workspace points handed out.
static bool plan_unit(long long n, bool pair, int B, long long spec_off, RaggedTask base,
                      std::vector<RaggedTask> (&ph)[kPhases], long long &ws)
{
    const long long M = (n % 2 == 0) ? n / 2 : n;
    base.n = (int)n;
    base.M = (int)M;
    base.blue = rg_smooth13(M) ? 0 : 1;
    base.spec_off = base.blue ? spec_off : 0;
    base.pair = pair ? 1 : 0;
    base.pad_ = 0;
    const long long L = transform_len(n, pair);
    if (L > (1LL << 30)) return false;
    base.L = (int)L;
    base.L1 = base.L2 = base.first = base.count = base.width = 0;
    base.ws_off = 0;
    if (L <= B) {
        base.kind = RG_SMALL;
        ph[0].push_back(base);
        return true;
    }
    int L1 = 0, L2 = 0, ca = 0, cb = 0;
    if (!split_tiles(L, B, !base.blue, L1, L2, ca, cb)) return false;
    base.L1 = L1;
    base.L2 = L2;
    base.ws_off = ws;
    ws += L;
    auto add = [&](int phase, int kind, int first, int count, int width) {
        RaggedTask t = base;
        t.kind = kind;
        t.first = first;
        t.count = count;
        t.width = width;
        ph[phase].push_back(t);
    };
    for (int c0 = 0; c0 < L1; c0 += ca) add(0, base.blue ? RG_BCOLS : RG_COLS, c0, std::min(ca, L1 - c0), ca);
    if (!base.blue) {
        // bands of row pairs (k2, L2 - k2), 1 <= k2 <= (L2 - 1) / 2, at most cb / 2 pairs each
        // (power-of-two bands, so that their slots divide the workgroup), then rows 0 and L2/2
        const int P = (L2 - 1) / 2, Rmax = cb / 2;
        for (int r0 = 1; r0 <= P;) {
            const int R = std::min(Rmax, pow2_floor(P - r0 + 1));
            add(1, RG_ROWS, r0, R, 2 * R);
            r0 += R;
        }
        add(1, RG_SELF, 0, (L2 % 2 == 0) ? 2 : 1, 2);
        return true;
    }
    for (int r0 = 0; r0 < L2; r0 += cb) add(1, RG_BROWS, r0, std::min(cb, L2 - r0), cb);
    for (int c0 = 0; c0 < L1; c0 += ca) add(2, RG_ICOLS, c0, std::min(ca, L1 - c0), ca);
    if (n % 2 == 0 || pair) {
        const long long kend = (n % 2 == 0) ? M / 2 : n / 2;   // bins 0 .. kend (with mirrors for even n)
        for (long long k0 = 0; k0 <= kend; k0 += B) add(3, RG_SPLIT, (int)k0, (int)std::min<long long>(B, kend + 1 - k0), 0);
    }
    return true;
}

// work of a task in points times the log of its transform length (queue order: largest first)
static double task_cost(const RaggedTask &t)
{
    switch (t.kind) {
        case RG_SMALL: return (t.blue ? 2.0 : 1.0) * (double)t.L * std::log2((double)t.L + 1.0);
        case RG_COLS: case RG_BCOLS: case RG_ICOLS: return (double)t.width * t.L2 * std::log2((double)t.L2 + 1.0);
        case RG_ROWS: case RG_SELF: return (double)t.width * t.L1 * std::log2((double)t.L1 + 1.0);
        case RG_BROWS: return 2.0 * t.width * t.L1 * std::log2((double)t.L1 + 1.0);
        default: return (double)t.count;
    }
}

static long long spec_offset(const Ragged &R, long long n, bool pair)
{
    return std::lower_bound(R.spec_index.begin(), R.spec_index.end(), std::make_pair(2 * n + (pair ? 1 : 0), -1LL))
        ->second;
}

static int resident_grid(const Ragged &R, int threads, int waves, size_t lds)
{
    int ncu = 0, per = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, R.device) != hipSuccess || ncu <= 0) {
        (void)hipGetLastError();
        ncu = 256;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, ragged_fn(threads, waves, false), threads, lds) != hipSuccess ||
        per <= 0) {
        (void)hipGetLastError();
        per = 1;
    }
    return ncu * per;
}

static int build_tasks(Ragged &R, int threads, int lds_max, int waves, Error &err)
{
    // each unit's tasks per phase, then the units dealt to eight queues per phase (largest unit
    // first, to the queue with the least work), each queue ordered by decreasing task cost
    std::vector<std::vector<RaggedTask>> unit_ph[kPhases];
    for (auto &v : unit_ph) v.resize(R.units.size());
    long long ws = 0;
    for (size_t ui = 0; ui < R.units.size(); ++ui) {
        const auto &u = R.units[ui];
        std::vector<RaggedTask> ph[kPhases];
        const size_t i = (size_t)u.first;
        const bool pair = u.second >= 0;
        const long long n = (long long)R.n[i];
        const long long M = (n % 2 == 0) ? n / 2 : n;
        RaggedTask base;
        std::memset(&base, 0, sizeof(base));
        base.in_off = (long long)R.in_off[i];
        base.out_off = (long long)R.out_off[i];
        if (pair) {
            base.in_off2 = (long long)R.in_off[(size_t)u.second];
            base.out_off2 = (long long)R.out_off[(size_t)u.second];
        }
        const long long so = rg_smooth13(M) ? 0 : spec_offset(R, n, pair);
        if (!plan_unit(n, pair, lds_max, so, base, ph, ws)) {
            err = {E_UNSUPPORTED, "ragged signal of length " + std::to_string(n) +
                                      " has no tile split for this configuration"};
            return E_UNSUPPORTED;
        }
        for (int q = 0; q < kPhases; ++q) unit_ph[q][ui].swap(ph[q]);
    }
    std::vector<RaggedTask> all;
    int off[kPhases], cnt[kPhases];
    int qoff[kPhases][8], qcnt[kPhases][8];
    for (int q = 0; q < kPhases; ++q) {
        std::vector<std::pair<double, size_t>> order;   // (cost of the unit's tasks in this phase, unit)
        for (size_t ui = 0; ui < unit_ph[q].size(); ++ui) {
            double c = 0.0;
            for (const auto &t : unit_ph[q][ui]) c += task_cost(t);
            if (!unit_ph[q][ui].empty()) order.push_back({c, ui});
        }
        std::stable_sort(order.begin(), order.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
        std::vector<RaggedTask> queue[8];
        double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (const auto &o : order) {
            const int k = (int)(std::min_element(load, load + 8) - load);
            load[k] += o.first;
            queue[k].insert(queue[k].end(), unit_ph[q][o.second].begin(), unit_ph[q][o.second].end());
        }
        off[q] = (int)all.size();
        for (int k = 0; k < 8; ++k) {
            std::stable_sort(queue[k].begin(), queue[k].end(),
                             [](const RaggedTask &a, const RaggedTask &b) { return task_cost(a) > task_cost(b); });
            qoff[q][k] = (int)all.size() - off[q];
            qcnt[q][k] = (int)queue[k].size();
            all.insert(all.end(), queue[k].begin(), queue[k].end());
        }
        cnt[q] = (int)all.size() - off[q];
    }
    if (all.size() > (size_t)INT_MAX) { err = {E_UNSUPPORTED, "too many ragged tasks"}; return E_UNSUPPORTED; }
    const size_t wsb = (size_t)std::max<long long>(ws, 1) * sizeof(double2);
    if (wsb > R.ws.bytes && R.ws.alloc(wsb)) { err = {E_NOMEM, "ragged workspace allocation failed"}; return E_NOMEM; }
    if (all.size() * sizeof(RaggedTask) > R.dtasks.bytes && R.dtasks.alloc(all.size() * sizeof(RaggedTask))) {
        err = {E_NOMEM, "ragged task table allocation failed"};
        return E_NOMEM;
    }
    if (hipMemcpy(R.dtasks.p, all.data(), all.size() * sizeof(RaggedTask), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        err = {E_HIP, "ragged task upload failed"};
        return E_HIP;
    }
    const size_t lds = (size_t)lds_points_padded(lds_max) * sizeof(double2);
    const int res = resident_grid(R, threads, waves, lds);
    for (int q = 0; q < kPhases; ++q) {
        R.phase_off[q] = off[q];
        R.phase_cnt[q] = cnt[q];
        R.phase_grid[q] = std::max(1, std::min(res, cnt[q]));
        for (int k = 0; k < 8; ++k) {
            R.queue_off[q][k] = qoff[q][k];
            R.queue_cnt[q][k] = qcnt[q][k];
        }
    }
    R.tasks.swap(all);
    R.threads = threads;
    R.lds_max = lds_max;
    R.waves = waves;
    R.drop_graph();
    return OK;
}

// Spectrum table: one entry per distinct Bluestein length, natural order, filled
// by the persistent kernel in spectrum mode (1024 threads, 8192 points of LDS).
static int build_spectra(Ragged &R, Error &err)
{
    std::vector<long long> ns;   // keys 2 n + pair
    for (const auto &u : R.units) {
        const long long n = (long long)R.n[(size_t)u.first];
        const long long M = (n % 2 == 0) ? n / 2 : n;
        if (!rg_smooth13(M)) ns.push_back(2 * n + (u.second >= 0 ? 1 : 0));
    }
    std::sort(ns.begin(), ns.end());
    ns.erase(std::unique(ns.begin(), ns.end()), ns.end());
    R.spec_index.clear();
    if (ns.empty()) return OK;
    constexpr int kB = 8192;
    std::vector<RaggedTask> st(ns.size());
    long long off = 0, slot = 1;
    for (size_t q = 0; q < ns.size(); ++q) {
        RaggedTask &t = st[q];
        std::memset(&t, 0, sizeof(t));
        const long long n = ns[q] / 2;
        const bool pair = (ns[q] & 1) != 0;
        t.n = (int)n;
        t.M = (int)((n % 2 == 0) ? n / 2 : n);
        t.pair = pair ? 1 : 0;
        t.blue = 1;
        t.kind = RG_SMALL;
        const long long L = transform_len(n, pair);
        t.L = (int)L;
        if (L > kB) {
            int L1 = 0, L2 = 0, ca = 0, cb = 0;
            if (!split_tiles(L, kB, false, L1, L2, ca, cb)) {
                err = {E_UNSUPPORTED, "no split for the Bluestein length of a signal of length " + std::to_string(n)};
                return E_UNSUPPORTED;
            }
            t.L1 = L1;
            t.L2 = L2;
            t.width = ca;   // columns per batch
            t.count = cb;   // rows per batch
            slot = std::max(slot, L);
        }
        t.spec_off = off;
        R.spec_index.push_back({ns[q], off});
        off += L;
    }
    std::stable_sort(st.begin(), st.end(), [](const RaggedTask &a, const RaggedTask &b) { return a.L > b.L; });
    if (R.spec.alloc((size_t)off * sizeof(double2))) { err = {E_NOMEM, "spectrum table allocation failed"}; return E_NOMEM; }
    DevBuf dt, scr;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, R.device) != hipSuccess || ncu <= 0) {
        (void)hipGetLastError();
        ncu = 256;
    }
    const int grid = (int)std::min<long long>(ncu, (long long)st.size());
    if (dt.alloc(st.size() * sizeof(RaggedTask)) || scr.alloc((size_t)grid * (size_t)slot * sizeof(double2))) {
        err = {E_NOMEM, "spectrum workspace allocation failed"};
        return E_NOMEM;
    }
    RaggedParams P{};
    P.tasks = (const RaggedTask *)dt.p;
    P.ntasks = (int)st.size();
    P.qoff[0] = 0;
    P.qcnt[0] = (int)st.size();   // one queue; the other seven are empty
    P.counter = (int *)R.counter.p;
    P.scratch = (double2 *)scr.p;
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
    if (e == hipSuccess) e = launch_ragged(P, 1024, 0, grid, (size_t)lds_points_padded(kB) * sizeof(double2), 0);
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
    // odd lengths > 1 pair up: two real signals, one complex transform
    {
        std::vector<std::pair<long long, int>> odd;
        for (size_t i = 0; i < count; ++i)
            if (lengths[i] % 2 == 1 && lengths[i] > 1) odd.push_back({(long long)lengths[i], (int)i});
            else R->units.push_back({(int)i, -1});
        std::stable_sort(odd.begin(), odd.end());
        for (size_t q = 0; q < odd.size(); ++q) {
            if (q + 1 < odd.size() && odd[q + 1].first == odd[q].first) {
                R->units.push_back({odd[q].second, odd[q + 1].second});
                ++q;
            } else {
                R->units.push_back({odd[q].second, -1});
            }
        }
    }
    if (R->counter.alloc(64 * 8 * kPhases * sizeof(int))) { err = {E_NOMEM, "ragged counter allocation failed"}; return E_NOMEM; }
    const int inst[6][2] = {{256, 0}, {256, 3}, {256, 4}, {512, 0}, {512, 4}, {1024, 0}};
    for (const auto &v : inst)
        for (int mf = 0; mf < 2; ++mf)
            if (hipFuncSetAttribute(ragged_fn(v[0], v[1], mf != 0), hipFuncAttributeMaxDynamicSharedMemorySize,
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
    if ((threads != 256 && threads != 512 && threads != 1024) || !pow2 || lds_max < 2048 || lds_max > 8192 ||
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

int ragged_set_precision(Ragged &R, int precision, Error &err)
{
    if (precision != 0 && precision != 1) {
        err = {E_INVALID, "ragged precision is 0 (FP64) or 1 (radix-16 stages on the matrix cores)"};
        return E_INVALID;
    }
    if (precision == 1 && !R.bx.p) {
        if (R.bx.alloc((size_t)kBxTableWords * sizeof(u4))) { err = {E_NOMEM, "table allocation failed"}; return E_NOMEM; }
        hipLaunchKernelGGL(k_bx_table, dim3(1), dim3(64), 0, nullptr, (u4 *)R.bx.p);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            (void)hipGetLastError();
            err = {E_HIP, std::string("ragged table: ") + hipGetErrorString(e)};
            return E_HIP;
        }
    }
    R.precision = precision;
    R.drop_graph();
    return OK;
}

// the reset and one persistent launch per non-empty phase
static hipError_t ragged_launch(Ragged &R, const float *x, double2 *y, double fct, hipStream_t s)
{
    hipLaunchKernelGGL(k_ragged_reset, dim3(1), dim3(64), 0, s, (int *)R.counter.p);
    hipError_t e = hipGetLastError();
    const size_t lds = (size_t)lds_points_padded(R.lds_max) * sizeof(double2);
    for (int q = 0; q < kPhases && e == hipSuccess; ++q) {
        if (!R.phase_cnt[q]) continue;
        RaggedParams P;
        std::memset(&P, 0, sizeof(P));
        P.x = x;
        P.y = y;
        P.tasks = (const RaggedTask *)R.dtasks.p + R.phase_off[q];
        P.ntasks = R.phase_cnt[q];
        P.counter = (int *)R.counter.p + 64 * 8 * q;
        for (int k = 0; k < 8; ++k) {
            P.qoff[k] = R.queue_off[q][k];
            P.qcnt[k] = R.queue_cnt[q][k];
        }
        P.ws = (double2 *)R.ws.p;
        P.spec = (double2 *)R.spec.p;
        P.mode = 0;
        // the matrix-core precision plans powers of two as radix-16 stages (order bit 1 clear)
        P.order = R.precision ? (R.rorder & ~2) : R.rorder;
        P.bx = R.precision ? (const u4 *)R.bx.p : nullptr;
        P.fct = fct;
        P.osign = R.forward ? 1.0 : -1.0;
        e = launch_ragged(P, R.threads, R.waves, (unsigned)R.phase_grid[q], lds, s);
    }
    return e;
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
    if (grid) *grid = *std::max_element(R.phase_grid, R.phase_grid + kPhases);
    if (flops) *flops = R.flops;
    if (scratch_bytes) *scratch_bytes = R.ws.bytes;
}

}  // namespace ifft

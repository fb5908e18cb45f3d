// Matrix-core (MFMA) DFT stages for complex FFTs of length 256 and 65536.
//
// A 256-point transform is two radix-16 stages, and each stage is one
// complex 16x16x16 matrix product on a wavefront.  With the input viewed as
// M[q][b] = x[b + 16 q]:
//   stage 1   Y1^T = M^T F16                  (DFT over q for each b)
//   twiddle   T[k1][b] = Y1[k1][b] W256^(b k1)
//   stage 2   Y2 = T F16                      X[k1 + 16 k2] = Y2[k1][k2]
// For v_mfma_f32_16x16x* lane l holds row (l mod 16) of the A operand and
// column (l mod 16) of B, and the result D[4 (l/16) + r][l mod 16] in its
// accumulator r.  The summation index k may be assigned to (l/16, element)
// at will as long as A and B agree, so the stage-1 result a lane holds,
// Y1[l mod 16][4 (l/16) + r], is already the lane's A operand for stage 2:
// no data moves between lanes.  A complex product is a real one of depth 32
// over [Re A, Im A] x [Re B; -Im B] (real part) and [Re A, Im A] x [Im B; Re B]
// (imaginary part).
//
// Two precisions:
//   F32    FP32 data and v_mfma_f32_16x16x4_f32 (FP32 products and
//          accumulation); twiddles are FP64-generated and applied as split
//          FP32 (hi + lo) compensated products.  complex64 in and out.
//   BF16X  FP64 data.  Before each stage the 256 values of a transform are
//          cut, at one power-of-two scale 2^E >= max |v|, into S = 5 slices of
//          8-bit integers (bf16-exact); the DFT matrix is cut the same way.
//          Every slice product is an integer below 2^14 and the 32-term MFMA
//          sums of all slice pairs with the same weight 2^-8(a+f) stay below
//          2^22, so v_mfma_f32_16x16x32_bf16 accumulates them exactly in FP32.
//          Pairs with a + f <= S - 1 are kept (relative truncation ~2^-40 per
//          stage); the weighted sums are combined in FP64.  complex64 in,
//          complex128 out.
//
// Lengths: 256 (one pass) and 256 n2 for n2 = 2 .. 8192 (four-step: 256-point
// column FFTs with the W_L^(t k) twiddle, then the n2-point row FFTs with a
// transposed store -- in the same launch for n2 <= 16 (and n2 = 32 with FP32
// data), on the matrix cores for n2 = 256, by an FP64 LDS row kernel
// otherwise).  A
// workgroup of four wavefronts owns a tile of 16 sub-transforms in LDS; each
// wavefront computes four of them.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "butterflies.h"
#include "engine.h"
#include "fft_common.h"
#include "lds_fft.h"
#include "mfma_stage.h"

namespace ifft {


enum { MF_F32 = 0, MF_BF16X = 1 };
// MP_FUSED: the column pass of L = 256 n2 for n2 <= 16 (and n2 = 32 with FP32 data, whose
// 32-column tile is one whole signal), with the n2-point row DFTs of the tile's whole signals
// done in LDS before the store (one pass, no workspace)
enum { MP_COL = 0, MP_ROW = 1, MP_SINGLE = 2, MP_FUSED = 3 };
constexpr int kFusedMaxN2 = 16;
__host__ __device__ constexpr bool fused_n2(int n2, int var) { return (n2 > 1 && n2 <= kFusedMaxN2) || (n2 == 32 && var == 0); }
constexpr int kRow = 273;           // LDS row stride (elements): 256 + 256/16 padding + 1
// sub-transforms per workgroup (four per wavefront): 32 for FP32 data (256-byte
// segments in the strided passes), 16 for FP64 data (LDS budget)
template <int VAR>
constexpr int tile_of() { return VAR == 0 ? 32 : 16; }
// threads per workgroup: eight wavefronts; a wavefront computes TILE / 8 transforms
// (four for FP32 data, two for FP64 data, so that the 78 KB FP64 tile still lets
// two workgroups, four wavefronts per SIMD, share a CU)
#ifndef IMPULSE_MFMA_BX_THREADS
#define IMPULSE_MFMA_BX_THREADS 512
#endif
#ifndef IMPULSE_MFMA_BX_WPE
#define IMPULSE_MFMA_BX_WPE 4
#endif
template <int VAR>
constexpr int threads_of() { return VAR == 0 ? 512 : IMPULSE_MFMA_BX_THREADS; }
template <int VAR>
constexpr int wpe_of() { return VAR == 0 ? 4 : IMPULSE_MFMA_BX_WPE; }

struct MfmaParams {
    const void *src;
    void *dst;
    long long batch;                // transforms of length L
    int L;                          // 256 n2
    int n2, ln2;                    // columns of the column pass (L / 256) and log2 of it
    double sign;                    // -1 forward, +1 backward
    double fct;                     // applied by the last pass
    const double2 *tw;              // 512 FP64 twiddles (forward sign): W_256^m, then W_L^m, m < 256
    const double2 *tw2;             // W_n2^q, q < n2 (forward sign), for n2 != 256
};

__device__ __forceinline__ int lidx(int n) { return n + (n >> 4); }

template <int VAR>
struct Elem;
template <>
struct Elem<MF_F32> {
    using T = float2;
};
template <>
struct Elem<MF_BF16X> {
    using T = double2;
};

// ---------------------------------------------------------------------------
// FP32 MFMA stage: in/out are the lane's four complex values
// ---------------------------------------------------------------------------

struct F32Consts {
    float br[8], bi[8];   // B operand for the real / imaginary accumulators, k-chunk kb = 0..7
};

__device__ __forceinline__ F32Consts f32_consts(double sign)
{
    F32Consts C;
    const int l = threadIdx.x & 63, g = l >> 4, j = l & 15;
    for (int e = 0; e < 4; ++e) {
        const double2 w = w16((4 * g + e) * j);
        const float fr = (float)w.x, fi = (float)(-sign * w.y);   // F = cos + i sign sin
        C.br[e] = fr;
        C.bi[e] = fi;
        C.br[4 + e] = -fi;
        C.bi[4 + e] = fr;
    }
    return C;
}

__device__ __forceinline__ void f32_stage(float2 (&a)[4], const F32Consts &C)
{
    f4 dr = {0.f, 0.f, 0.f, 0.f}, di = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        const float av = kb < 4 ? a[kb].x : a[kb - 4].y;
        dr = __builtin_amdgcn_mfma_f32_16x16x4f32(av, C.br[kb], dr, 0, 0, 0);
        di = __builtin_amdgcn_mfma_f32_16x16x4f32(av, C.bi[kb], di, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) a[r] = make_float2(dr[r], di[r]);
}

// y * (w_hi + w_lo): w (FP64, sign applied) split into FP32 high and low parts
__device__ __forceinline__ float2 f32_twiddle(float2 y, double2 w)
{
    const float wr = (float)w.x, wi = (float)w.y;
    const float wrl = (float)(w.x - (double)wr), wil = (float)(w.y - (double)wi);
    float re = fmaf(y.x, wr, -y.y * wi);
    float im = fmaf(y.x, wi, y.y * wr);
    re += fmaf(y.x, wrl, -y.y * wil);
    im += fmaf(y.x, wil, y.y * wrl);
    return make_float2(re, im);
}

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------

__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ void put_out(float2 *y, long long i, double2 v) { y[i] = make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ void put_out(double2 *y, long long i, double2 v) { y[i] = v; }

// The n2-point row DFTs of a fused tile: its columns are the n2 columns t of ns whole signals
// (column g n2 + t of the tile holds the twiddled 256-point transform of column t of signal
// s0 + g).  X[k1 + 256 k2] = sum_t T[t][k1] W_n2^(t k2) in FP64, scaled by fct; one job per
// (signal, k1), k1 fastest, so the stores run over 256 consecutive outputs.
// n2 = 32 with FP32 data (the tile is one signal): the 32-point row DFTs as radix 8 x 4 in LDS.
// Step 1, per (k1, t1): the 8-point DFT over t2 of T[t1 + 4 t2][k1], times W_32^(t1 q), back into
// the same columns (column t1 + 4 q; each job reads and writes only its own eight); step 2, per
// (k1, q): X[k1 + 256 (q + 8 r)] = sum_t1 W_4^(t1 r) Y[t1 + 4 q][k1].  W_32^j = W_256^(8 j) comes
// from the LDS table tws (forward sign).
template <typename O>
__device__ __forceinline__ void fused_row_dft32(float2 *lds, const double2 *tws, O *y, long long sg, long long L,
                                                double sign, double fct, int tid, int NT)
{
    for (int q = tid; q < 4 * 256; q += NT) {
        const int k1 = q & 255, t1 = q >> 8;
        double2 v[8];
#pragma unroll
        for (int t2 = 0; t2 < 8; ++t2) v[t2] = to_d2(lds[(t1 + 4 * t2) * kRow + lidx(k1)]);
        dft<8>(v, sign);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double2 w = v[k];
            const double2 r = t1 * k ? cmul(w, tw_sign(tws[8 * t1 * k], sign)) : w;
            lds[(t1 + 4 * k) * kRow + lidx(k1)] = make_float2((float)r.x, (float)r.y);
        }
    }
    __syncthreads();
    O *ys = y + sg * L;
    for (int q = tid; q < 8 * 256; q += NT) {
        const int k1 = q & 255, k2a = q >> 8;
        double2 v[4];
#pragma unroll
        for (int t1 = 0; t1 < 4; ++t1) v[t1] = to_d2(lds[(t1 + 4 * k2a) * kRow + lidx(k1)]);
        dft<4>(v, sign);
#pragma unroll
        for (int r = 0; r < 4; ++r) put_out(ys, k1 + 256LL * (k2a + 8 * r), cscale(v[r], fct));
    }
}

template <int N2, typename E, typename O>
__device__ __forceinline__ void fused_row_dfts(const E *lds, O *y, long long s0, long long batch, long long L,
                                               int ns, double sign, double fct, int tid, int NT)
{
    for (int q = tid; q < ns * 256; q += NT) {
        const int g = q >> 8, k1 = q & 255;
        const long long sg = s0 + g;
        if (sg >= batch) break;   // q only grows, so every later job is past the batch too
        double2 v[N2];
#pragma unroll
        for (int t = 0; t < N2; ++t) v[t] = to_d2(lds[(g * N2 + t) * kRow + lidx(k1)]);
        dft<N2>(v, sign);
        O *ys = y + sg * L + k1;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) put_out(ys, 256LL * k2, cscale(v[k2], fct));
    }
}

template <int VAR, int PASS>
__global__ void __launch_bounds__(threads_of<VAR>()) __attribute__((amdgpu_waves_per_eu(wpe_of<VAR>())))
k_mfma256(const MfmaParams P)
{
    constexpr int TILE = tile_of<VAR>(), NT = threads_of<VAR>(), TPS = 256 / TILE;   // threads, tiles per signal
    constexpr int PER = TILE / (NT / 64);   // transforms per wavefront
    static_assert(PER * (NT / 64) == TILE, "the tile splits evenly over the wavefronts");
    using E = typename Elem<VAR>::T;
    __shared__ E lds[TILE * kRow];
    __shared__ double2 tws[512];   // W_256^m and W_L^m (m < 256), forward sign
    const int tid = threadIdx.x;
    // the tiles of one signal run on one XCD, so its 2 KB rows are fetched into one L2
    const long long tile = xcd_span_block(blockIdx.x, gridDim.x);
    const long long L = P.L;
    long long base = 0;   // first element of the signal (ROW) or of transform 0 of the tile (SINGLE)
    int t0 = 0, ntr = TILE;
    if (PASS == MP_SINGLE) {
        base = tile * TILE * 256;
        ntr = (int)min((long long)TILE, P.batch - tile * TILE);
    } else if (PASS == MP_ROW) {
        base = (tile / TPS) * L;
        t0 = (int)(tile % TPS) * TILE;
    }
    // the column pass: the tile's transforms are the batch's columns gc = tile TILE + c (batch n2 of
    // them; column t of signal s holds points s L + t + n2 n, n < 256), so a tile may span signals
    const long long ncols = P.batch * (long long)P.n2;
    auto col_at = [&](int c, long long &off, int &t) {   // first point of column c of the tile, its t
        const long long gc = tile * TILE + c;
        t = (int)(gc & (P.n2 - 1));
        off = (gc >> P.ln2) * L + t;
        return gc < ncols;
    };
    constexpr bool COLS = PASS == MP_COL || PASS == MP_FUSED;   // the column pass of L = 256 n2
    for (int m = tid; m < (COLS ? 512 : 256); m += NT) tws[m] = P.tw[m];
    auto put = [&](int c, int n, float2 v) {
        if constexpr (VAR == MF_F32) lds[c * kRow + lidx(n)] = v;
        else lds[c * kRow + lidx(n)] = make_double2((double)v.x, (double)v.y);
    };
    // ---- load (16-byte accesses: complex64 pairs, or one complex128) ----
    if (COLS) {
        // row n of the tile: columns 2p, 2p + 1 are neighbours in one signal (n2 even), one
        // 16-byte load of two complex64.  Consecutive e walk a signal's memory in order: the
        // PW pairs of a row, then the rows, then (n2 < TILE) the tile's next signal.
        const float2 *x = (const float2 *)P.src;
        constexpr int LT = TILE == 32 ? 5 : 4;
        static_assert(1 << LT == TILE, "the tile is 16 or 32 columns");
        const int lpw = min(P.ln2, LT) - 1;   // log2 of the PW = min(n2, TILE) / 2 pairs of a row
        for (int e = tid; e < TILE * 128; e += NT) {
            const int p = ((e >> (lpw + 8)) << lpw) + (e & ((1 << lpw) - 1)), n = (e >> lpw) & 255;
            long long off;
            int t;
            float4 v = float4{0.f, 0.f, 0.f, 0.f};   // columns past the batch compute zeros
            if (col_at(2 * p, off, t)) v = *(const float4 *)(x + off + (long long)P.n2 * n);
            put(2 * p, n, make_float2(v.x, v.y));
            put(2 * p + 1, n, make_float2(v.z, v.w));
        }
    } else if (PASS == MP_ROW) {
        if constexpr (VAR == MF_F32) {
            const float4 *x = (const float4 *)((const float2 *)P.src + base + 256LL * t0);
            for (int e = tid; e < TILE * 128; e += NT) {
                const int c = e >> 7, pp = e & 127;
                const float4 v = x[e];
                put(c, 2 * pp, make_float2(v.x, v.y));
                put(c, 2 * pp + 1, make_float2(v.z, v.w));
            }
        } else {
            const double2 *x = (const double2 *)P.src + base + 256LL * t0;
            for (int e = tid; e < TILE * 256; e += NT) lds[(e >> 8) * kRow + lidx(e & 255)] = x[e];
        }
    } else {
        const float4 *x = (const float4 *)((const float2 *)P.src + base);
        for (int e = tid; e < TILE * 128; e += NT) {
            const int c = e >> 7, pp = e & 127;
            const float4 v = c < ntr ? x[e] : float4{0.f, 0.f, 0.f, 0.f};   // a partial last tile computes zeros
            put(c, 2 * pp, make_float2(v.x, v.y));
            put(c, 2 * pp + 1, make_float2(v.z, v.w));
        }
    }
    __syncthreads();
    // ---- compute: wave w owns transforms PER w .. PER w + PER - 1 ----
    const int l = tid & 63, w = tid >> 6, g = l >> 4, j = l & 15;
    const double sign = P.sign;
    const bool last = PASS == MP_ROW || PASS == MP_SINGLE;
    // W_256^(b k1) for this lane's four (b, k1) pairs
    double2 w1[4];
    if constexpr (VAR == MF_F32) {
#pragma unroll
        for (int r = 0; r < 4; ++r) w1[r] = tw_sign(tws[(4 * g + r) * j], sign);
    }
    // W_L^(t k) = W_L^(m mod 256) W_n2^(m / 256), m = t k < L: the first factor from LDS; the second
    // for n2 <= 256 is W_256^((m / 256) (256 / n2)), from LDS too, and from the W_n2 table otherwise;
    // product in FP64
    const int hstep = P.n2 <= 256 ? 256 >> P.ln2 : 0;
    auto step_tw = [&](int t, int k) -> double2 {
        const int m = t * k;
        const double2 hi = hstep ? tws[(m >> 8) * hstep] : P.tw2[m >> 8];
        return tw_sign(cmul(tws[256 + (m & 255)], hi), sign);
    };
    auto col_t = [&](int c) { return (int)((tile * TILE + c) & (P.n2 - 1)); };
    if constexpr (VAR == MF_F32) {
        const F32Consts C = f32_consts(sign);
        for (int u = 0; u < PER; ++u) {
            const int c = PER * w + u;
            float2 a[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = lds[c * kRow + lidx(j + 16 * (4 * g + r))];
            f32_stage(a, C);
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = f32_twiddle(a[r], w1[r]);
            f32_stage(a, C);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * g + r + 16 * j;
                float2 v = a[r];
                if (COLS) v = f32_twiddle(v, step_tw(col_t(c), k));
                if (last && P.fct != 1.0) v = make_float2((float)(v.x * P.fct), (float)(v.y * P.fct));
                lds[c * kRow + lidx(k)] = v;
            }
        }
    } else {
        const BxConsts C = bx_consts(sign);
        for (int u = 0; u < PER; ++u) {
            const int c = PER * w + u;
            double2 a[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = lds[c * kRow + lidx(j + 16 * (4 * g + r))];
            bx_stage(a, C);
            // W_256^(b k1) from the LDS table at use (keeps 16 registers free)
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = cmul(a[r], tw_sign(tws[(4 * g + r) * j], sign));
            bx_stage(a, C);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * g + r + 16 * j;
                double2 v = a[r];
                if (COLS) v = cmul(v, step_tw(col_t(c), k));
                if (last) v = cscale(v, P.fct);
                lds[c * kRow + lidx(k)] = v;
            }
        }
    }
    __syncthreads();
    // ---- store ----
    if constexpr (PASS == MP_FUSED) {
        using O = typename std::conditional<VAR == MF_F32, float2, double2>::type;
        O *y = (O *)P.dst;
        const int ns = TILE / P.n2;
        const long long s0 = tile * ns;
        if constexpr (VAR == MF_F32) {
            if (P.n2 == 32) {   // one signal per tile
                if (s0 < P.batch) fused_row_dft32(lds, tws, y, s0, L, sign, P.fct, tid, NT);
                return;
            }
        }
        switch (P.n2) {
            case 2: fused_row_dfts<2>(lds, y, s0, P.batch, L, ns, sign, P.fct, tid, NT); break;
            case 4: fused_row_dfts<4>(lds, y, s0, P.batch, L, ns, sign, P.fct, tid, NT); break;
            case 8: fused_row_dfts<8>(lds, y, s0, P.batch, L, ns, sign, P.fct, tid, NT); break;
            default: fused_row_dfts<16>(lds, y, s0, P.batch, L, ns, sign, P.fct, tid, NT); break;
        }
    } else if constexpr (VAR == MF_F32) {
        float2 *y = (float2 *)P.dst + base;
        if (PASS == MP_SINGLE) {
            float4 *y4 = (float4 *)y;
            for (int e = tid; e < ntr * 128; e += NT) {
                const int c = e >> 7, kk = 2 * (e & 127);
                const float2 a = lds[c * kRow + lidx(kk)], b = lds[c * kRow + lidx(kk + 1)];
                y4[e] = make_float4(a.x, a.y, b.x, b.y);
            }
        } else if (PASS == MP_COL) {
            // element (c, k) of column t of signal s goes to s L + t + n2 k (pairs of neighbours)
            float2 *yc = (float2 *)P.dst;
            for (int e = tid; e < TILE * 128; e += NT) {
                const int p = e % (TILE / 2), k = e / (TILE / 2);
                long long off;
                int t;
                if (!col_at(2 * p, off, t)) continue;
                const float2 a = lds[2 * p * kRow + lidx(k)], b = lds[(2 * p + 1) * kRow + lidx(k)];
                *(float4 *)(yc + off + (long long)P.n2 * k) = make_float4(a.x, a.y, b.x, b.y);
            }
        } else {
            // element (c, k) goes to t0 + c + 256 k: TILE consecutive complex64 per k
            float4 *y4 = (float4 *)(y + t0);
            for (int e = tid; e < TILE * 128; e += NT) {
                const int p = e % (TILE / 2), k = e / (TILE / 2);
                const float2 a = lds[2 * p * kRow + lidx(k)], b = lds[(2 * p + 1) * kRow + lidx(k)];
                y4[128LL * k + p] = make_float4(a.x, a.y, b.x, b.y);
            }
        }
    } else {
        double2 *y = (double2 *)P.dst + base;
        if (PASS == MP_SINGLE) {
            for (int e = tid; e < ntr * 256; e += NT) y[e] = lds[(e >> 8) * kRow + lidx(e & 255)];
        } else if (PASS == MP_COL) {
            double2 *yc = (double2 *)P.dst;
            for (int e = tid; e < TILE * 256; e += NT) {
                const int c = e % TILE, k = e / TILE;
                long long off;
                int t;
                if (col_at(c, off, t)) yc[off + (long long)P.n2 * k] = lds[c * kRow + lidx(k)];
            }
        } else {
            for (int e = tid; e < TILE * 256; e += NT) {
                const int c = e % TILE, k = e / TILE;
                y[(long long)(t0 + c) + 256LL * k] = lds[c * kRow + lidx(k)];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// the row pass of L = 256 n2 for n2 = 32 .. 8192, n2 != 256 (FP64 LDS stages)
// ---------------------------------------------------------------------------

// The column pass leaves row k1 of signal s as n2 contiguous points t at s L + n2 k1.  A
// workgroup takes R = NT kPointsPerThread / n2 consecutive rows of one signal -- one contiguous
// run of R n2 points -- transforms them in LDS (radix-8 powers of two, W_n2 from the table)
// and stores X[k1 + 256 k2] k1 fastest, in runs of R consecutive outputs.
// Four waves per SIMD (128 registers, 13 of them spilled) measured faster than the
// compiler's 142 registers at three, and 512 threads faster than 256 (longer store runs).
template <int VAR, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) k_mfma_rows(const MfmaParams P)
{
    constexpr int PTS = NT * kPointsPerThread;
    __shared__ double2 lds[lds_points_padded(PTS)];
    using I = typename Elem<VAR>::T;   // the workspace: complex64 (F32) or complex128 (BF16X)
    using O = typename std::conditional<VAR == MF_F32, float2, double2>::type;
    const int n2 = P.n2, lr = __builtin_ctz((unsigned)PTS) - P.ln2, R = 1 << lr;
    const long long blk = xcd_span_block(blockIdx.x, gridDim.x);
    const long long sg = blk >> (8 - lr);   // 256 / R workgroups per signal
    const int r0 = (int)(blk & ((256 >> lr) - 1)) << lr;
    const I *x = (const I *)P.src + sg * (long long)P.L + (long long)r0 * n2;
    for (int e = threadIdx.x; e < PTS; e += NT) lds[lpad(e)] = to_d2(x[e]);
    __syncthreads();
    lds_fft_inline<kPointsPerThread, true>(lds, n2, R, P.sign, 2, TwTable{P.tw2, n2});
    O *y = (O *)P.dst + sg * (long long)P.L + r0;
    for (int e = threadIdx.x; e < PTS; e += NT) {
        const int r = e & (R - 1), k2 = e >> lr;
        put_out(y, r + 256LL * k2, cscale(lds[lpad(r * n2 + k2)], P.fct));
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct Mfma {
    int precision = 0;
    long long L = 0, batch = 0;
    int n2 = 1, ln2 = 0;                      // L = 256 n2
    DevBuf tw, tw2, tmp;
};

static int upload(DevBuf &b, const void *p, size_t n)
{
    if (b.alloc(n)) return E_NOMEM;
    if (hipMemcpy(b.p, p, n, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        return E_HIP;
    }
    return OK;
}

// L = 256 n2 as a four-step transform: the 256-point column transforms (stride n2) on the
// matrix cores with the twiddle W_L^(t k1), then the n2-point row transforms with the
// transposed store X[k1 + 256 k2] -- in the same launch for n2 <= 16 (and n2 = 32 with FP32
// data), on the matrix cores for n2 = 256, otherwise by the FP64 LDS row kernel.
int mfma_create(Mfma **out, size_t length, size_t batch, int precision, Error &err)
{
    if (!out) { err = {E_INVALID, "null handle pointer"}; return E_INVALID; }
    *out = nullptr;
    if (length < 256 || length > ((size_t)1 << 21) || (length & (length - 1))) {
        err = {E_UNSUPPORTED, "matrix-core FFT lengths are the powers of two from 256 to 2^21"};
        return E_UNSUPPORTED;
    }
    if (batch == 0 || batch > (size_t)1 << 40) { err = {E_INVALID, "batch out of range"}; return E_INVALID; }
    if (precision != MF_F32 && precision != MF_BF16X) { err = {E_INVALID, "unknown precision"}; return E_INVALID; }
    int rc = device_check(err);
    if (rc != OK) return rc;
    auto M = std::make_unique<Mfma>();
    M->precision = precision;
    M->L = (long long)length;
    M->batch = (long long)batch;
    M->n2 = (int)(length / 256);
    while ((1 << M->ln2) < M->n2) ++M->ln2;
    // W_256^m then W_L^m for m < 256, and W_n2^q for q < n2, from the long-double host generator
    std::vector<double2> t256, tL;
    twiddle_table(t256, 256, 256, 1);
    twiddle_table(tL, (uint64_t)length, 256, 1);
    t256.insert(t256.end(), tL.begin(), tL.end());
    rc = upload(M->tw, t256.data(), t256.size() * sizeof(double2));
    if (rc == OK && M->n2 > 1 && M->n2 != 256) {
        std::vector<double2> t2;
        twiddle_table(t2, (uint64_t)M->n2, (uint64_t)M->n2, 1);
        rc = upload(M->tw2, t2.data(), t2.size() * sizeof(double2));
    }
    const size_t es = precision == MF_F32 ? sizeof(float2) : sizeof(double2);
    const bool fused = fused_n2(M->n2, precision);
    if (rc == OK && length > 256 && !fused && M->tmp.alloc((size_t)length * batch * es)) rc = E_NOMEM;
    if (rc != OK) {
        err = {rc, "matrix-core FFT table or workspace allocation failed"};
        return rc;
    }
    *out = M.release();
    return OK;
}

void mfma_destroy(Mfma *m) { delete m; }

size_t mfma_workspace_bytes(const Mfma &m) { return m.tmp.bytes; }

template <int VAR>
static hipError_t launch_var(const MfmaParams &P, int pass, long long batch, hipStream_t s)
{
    constexpr int TILE = tile_of<VAR>();
    const long long grid = pass == MP_SINGLE ? (batch + TILE - 1) / TILE
                         : pass == MP_COL || pass == MP_FUSED ? (batch * (long long)P.n2 + TILE - 1) / TILE
                                            : batch * (256 / TILE);
    if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 g((unsigned)grid), b(threads_of<VAR>());
    if (pass == MP_FUSED) hipLaunchKernelGGL((k_mfma256<VAR, MP_FUSED>), g, b, 0, s, P);
    else if (pass == MP_COL) hipLaunchKernelGGL((k_mfma256<VAR, MP_COL>), g, b, 0, s, P);
    else if (pass == MP_ROW) hipLaunchKernelGGL((k_mfma256<VAR, MP_ROW>), g, b, 0, s, P);
    else hipLaunchKernelGGL((k_mfma256<VAR, MP_SINGLE>), g, b, 0, s, P);
    return hipGetLastError();
}

int mfma_execute(Mfma &m, const void *in, void *out, int forward, double fct, hipStream_t s, Error &err)
{
    if (!in || !out) { err = {E_INVALID, "null data pointer"}; return E_INVALID; }
    if (((uintptr_t)in | (uintptr_t)out) & 15) {
        err = {E_INVALID, "matrix-core FFT buffers must be 16-byte aligned"};
        return E_INVALID;
    }
    MfmaParams P{};
    P.batch = m.batch;
    P.L = (int)m.L;
    P.n2 = m.n2;
    P.ln2 = m.ln2;
    P.sign = forward ? -1.0 : 1.0;
    P.fct = fct;
    P.tw = (const double2 *)m.tw.p;
    P.tw2 = (const double2 *)m.tw2.p;
    auto go = [&](int pass, const void *src, void *dst) {
        P.src = src;
        P.dst = dst;
        return m.precision == MF_F32 ? launch_var<MF_F32>(P, pass, m.batch, s)
                                     : launch_var<MF_BF16X>(P, pass, m.batch, s);
    };
    hipError_t e;
    const bool fused = fused_n2(m.n2, m.precision);
    if (m.L == 256) {
        e = go(MP_SINGLE, in, out);
    } else if (fused) {
        e = go(MP_FUSED, in, out);
    } else {
        e = go(MP_COL, in, m.tmp.p);
        if (e == hipSuccess && m.n2 == 256) {
            e = go(MP_ROW, m.tmp.p, out);
        } else if (e == hipSuccess) {
            // R n2 = 4096 points per workgroup (8192 for n2 = 8192); 256 / R workgroups a signal
            P.src = m.tmp.p;
            P.dst = out;
            const int NT = std::max(512, m.n2 / kPointsPerThread), R = NT * kPointsPerThread / m.n2;
            const long long grid = m.batch * (256 / R);
            if (grid > 0x7fffffffLL) {
                e = hipErrorInvalidValue;
            } else {
                const dim3 g((unsigned)grid);
                if (m.precision == MF_F32) {
                    if (NT == 1024) hipLaunchKernelGGL((k_mfma_rows<MF_F32, 1024>), g, dim3(1024), 0, s, P);
                    else hipLaunchKernelGGL((k_mfma_rows<MF_F32, 512>), g, dim3(512), 0, s, P);
                } else {
                    if (NT == 1024) hipLaunchKernelGGL((k_mfma_rows<MF_BF16X, 1024>), g, dim3(1024), 0, s, P);
                    else hipLaunchKernelGGL((k_mfma_rows<MF_BF16X, 512>), g, dim3(512), 0, s, P);
                }
                e = hipGetLastError();
            }
        }
    }
    if (e != hipSuccess) {
        err = {E_HIP, std::string("matrix-core FFT launch: ") + hipGetErrorString(e)};
        return E_HIP;
    }
    return OK;
}

}  // namespace ifft

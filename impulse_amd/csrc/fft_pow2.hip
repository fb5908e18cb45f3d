// Power-of-two FP64 FFT kernel with compile-time length N (64 .. 8192).
//
// Radix-16 Stockham stages (one radix-2/4/8 stage last when log2 N is not a
// multiple of 4).  Each of the N/16 threads of a transform holds 16 points in
// registers for every stage: the first stage reads them straight from global
// memory (x[i + q N/16], coalesced), the middle stages exchange through LDS,
// and the last stage writes straight to global memory (y[b + q N/R]).
//
// LDS holds ONE component per exchange: real parts are written, read back
// into the next stage's positions, then the imaginary parts follow.  A
// 4096-point FP64 transform thus needs 34 KiB of LDS instead of 68 KiB, which
// doubles the workgroups a CU can keep in flight; rows are padded by one word
// every 16 (index + index/16) so the radix-16 scatter is free of bank
// conflicts.  Twiddles W^{qm} are built from four table loads (W^m, W^2m,
// W^4m, W^8m) and at most three products, which keeps them within a few ulp.
//
// Per-point load modes (complex, real-half, real-pair, c2r pre-processing,
// Bluestein chirp) and store modes (plain, four-step twiddle, Bluestein
// spectrum product or chirp, c2r outputs) are fused into the global accesses;
// FAST instantiations cover plain complex loads and plain / twiddled complex
// stores (complex double, or complex float widened on load and rounded on
// store) with direct pointer arithmetic.
//
// MAP selects the lane mapping of the global accesses:
//   MAP_ROW    : lanes walk along a transform (contiguous axis)
//   MAP_COL    : lanes walk across the TB transforms of a workgroup first
//                (transforms adjacent in memory, strided axis); LDS rows are
//                interleaved so the exchange stays conflict-free
//   MAP_ROWCOL : loads along the transform, stores across transforms (the
//                second pass of a four-step transform, whose output stride is
//                the first factor): the last stage is transposed through LDS.
#include "fft_common.h"

namespace ifft {

template <int N> struct Pow2Radices;
template <> struct Pow2Radices<64> { static constexpr int n = 2; static constexpr int r[2] = {16, 4}; };
template <> struct Pow2Radices<128> { static constexpr int n = 2; static constexpr int r[2] = {16, 8}; };
template <> struct Pow2Radices<256> { static constexpr int n = 2; static constexpr int r[2] = {16, 16}; };
template <> struct Pow2Radices<512> { static constexpr int n = 3; static constexpr int r[3] = {16, 16, 2}; };
template <> struct Pow2Radices<1024> { static constexpr int n = 3; static constexpr int r[3] = {16, 16, 4}; };
template <> struct Pow2Radices<2048> { static constexpr int n = 3; static constexpr int r[3] = {16, 16, 8}; };
template <> struct Pow2Radices<4096> { static constexpr int n = 3; static constexpr int r[3] = {16, 16, 16}; };
template <> struct Pow2Radices<8192> { static constexpr int n = 4; static constexpr int r[4] = {16, 16, 16, 2}; };

__device__ __forceinline__ int padx(int idx) { return idx + (idx >> 4); }

enum Pow2Map { MAP_ROW = 0, MAP_COL = 1, MAP_ROWCOL = 2 };

// transforms per workgroup, fixed per length and mapping: 256 threads along
// rows, 512 threads across transforms (at most 64 transforms)
__host__ __device__ constexpr int pow2_tb(int N, int map)
{
    return map == MAP_ROW ? (N >= 4096 ? 1 : 4096 / N) : (N >= 8192 ? 1 : (8192 / N > 64 ? 64 : 8192 / N));
}

// doubles of LDS per transform: one padded component row plus one word so
// that rows of neighbouring transforms start on different banks; the
// row-load / column-store variant reuses the row as N/2 complex words with
// an odd stride for its transposed store, hence two words
template <int N, int MAP> constexpr int pow2_ld() { return N + N / 16 + (MAP == MAP_ROWCOL ? 2 : 1); }

// v[q] *= W^{q m} (forward table, conjugated for the backward sign), q < R.
// W^m, W^2m, W^4m, W^8m come from the table, every other power is one to three
// products of those, and each twiddle is used as soon as it is formed so that
// at most five are live at a time.
// IO_BLUE_IN / IO_BLUE_OUT: the Bluestein column passes with direct buffer
// accesses -- chirp-in loads (complex or odd real pairs) with a twiddled store,
// and plain loads with a chirp-out complex store (rows and columns mappings)
// IO_W4 (added to IO_CONV or a Bluestein mode) selects the instance held to four waves per
// SIMD, for grids of a few rounds.  IO_S32 / IO_D32 (added to IO_PLAIN, or IO_S32 to
// IO_TWIDDLE) select the instances whose source / destination array holds complex float: widened
// to FP64 on load, rounded to nearest on store.  They are instances of their own so that the
// FP64 instances carry no float code (a runtime branch in the shared instances cost the FP64
// passes 1-2 %).
// IO_R2C_HALF: a whole even-length r2c signal per transform (its half-length complex sequence,
// per-point loads), the real post-process fused into the store.
enum Pow2Io { IO_GENERIC = 0, IO_PLAIN = 1, IO_TWIDDLE = 2, IO_R2C_ROWPAIR = 3, IO_CONV = 4, IO_BLUE_IN = 5,
              IO_BLUE_OUT = 6, IO_R2C_HALF = 7, IO_W4 = 8, IO_S32 = 16, IO_D32 = 32 };

// Row assignment of the fused real post-process (ST_R2C_ROWPAIR, rows of L1 = N
// points, L2 rows per signal, Z[j + L2 n] in row j).  Z[M - k] for k = j + L2 n
// sits in row L2 - j at point L1 - 1 - n, so a workgroup of TB rows takes TB/2
// pairs (j, L2 - j), j = 1 .. L2/2 - 1: pair slot p holds row j in slot p and its
// partner in slot p + TB/2.  The self-paired rows 0 and L2/2 take the first free
// pair slot of the signal's last group (slots cnt and cnt + TB/2), or a group of
// their own when the pairs fill every group; then each signal has as many groups
// as its rows fill, and the grid no round with a few groups left over.
struct RowpairGroup {
    long long u;   // signal
    long long w;   // group within the signal
    int cnt;       // pair slots holding pairs
    bool self;     // the self-paired rows sit in pair slot cnt
};

__host__ __device__ __forceinline__ long long rowpair_groups(long long L2, int TB)
{
    const long long pp = TB / 2, npairs = L2 / 2 - 1;
    const long long W = (npairs + pp - 1) / pp;
    return W * pp > npairs ? W : W + 1;
}

__device__ __forceinline__ RowpairGroup rowpair_group(long long blk, int TB, long long L2)
{
    const int pp = TB / 2;
    const long long npairs = L2 / 2 - 1;
    const long long G = rowpair_groups(L2, TB);
    RowpairGroup R;
    R.u = blk / G;
    R.w = blk - R.u * G;
    const long long left = npairs - R.w * pp;
    R.cnt = (int)(left < 0 ? 0 : (left < pp ? left : pp));
    R.self = R.w == G - 1;
    return R;
}

// logical transform of slot t, or -1 for an idle slot
__device__ __forceinline__ long long rowpair_transform(const RowpairGroup &R, int t, int TB, long long L2)
{
    const int pp = TB / 2, p = t % pp;
    const bool partner = t >= pp;
    long long row;
    if (p < R.cnt) {
        const long long j = R.w * pp + 1 + p;
        row = partner ? L2 - j : j;
    } else if (R.self && p == R.cnt) {
        row = partner ? L2 / 2 : 0;
    } else {
        return -1;
    }
    return R.u * L2 + row;
}

// X[k] and X[M-k] of an even real transform from Z[k], Z[M-k] (M complex
// points, W = W_{2M}); adj_w doubles the interior bins, fwd_flag 0 conjugates
__device__ __forceinline__ void real_split_store(const FftParams &P, const TfInfo &U, long long k, double2 zk,
                                                 double2 zm)
{
    const long long M = P.M;
    const double2 w = real_tw(P, k);   // W_{2M}^k: the direct table or the two-level one
    const double2 e1 = cscale(cadd(zk, cconj(zm)), 0.5), d1 = csub(zk, cconj(zm));
    const double2 o1 = make_double2(0.5 * d1.y, -0.5 * d1.x);
    const double2 e2 = cscale(cadd(zm, cconj(zk)), 0.5), d2 = csub(zm, cconj(zk));
    const double2 o2 = make_double2(0.5 * d2.y, -0.5 * d2.x);
    const long long km = M - k;
    const double fk = (P.adj_w && k != 0 && k != M) ? 2.0 * P.fct : P.fct;
    const double fm = (P.adj_w && km != 0 && km != M) ? 2.0 * P.fct : P.fct;
    double2 Xk = cscale(cadd(e1, cmul(o1, w)), fk);
    double2 Xm = cscale(cadd(e2, cmul(o2, make_double2(-w.x, w.y))), fm);
    if (!P.fwd_flag) {
        Xk = cconj(Xk);
        Xm = cconj(Xm);
    }
    const long long sax = P.out.sax;
    gst_c(P.dst, P.dst32, U.oa + k * sax, Xk);
    if (km != k) gst_c(P.dst, P.dst32, U.oa + km * sax, Xm);
}

template <int R>
__device__ __forceinline__ void apply_twiddles(double2 *v, __amdgpu_buffer_rsrc_t tw, int m, double s)
{
    const double2 w1 = bload2(tw, 16 * m);
    v[1] = cmul(v[1], tw_sign(w1, s));
    if constexpr (R >= 4) {
        const double2 w2 = bload2(tw, 32 * m);
        v[2] = cmul(v[2], tw_sign(w2, s));
        const double2 w3 = cmul(w1, w2);
        v[3] = cmul(v[3], tw_sign(w3, s));
        if constexpr (R >= 8) {
            const double2 w4 = bload2(tw, 64 * m);
            v[4] = cmul(v[4], tw_sign(w4, s));
            v[5] = cmul(v[5], tw_sign(cmul(w4, w1), s));
            v[6] = cmul(v[6], tw_sign(cmul(w4, w2), s));
            v[7] = cmul(v[7], tw_sign(cmul(w4, w3), s));
            if constexpr (R >= 16) {
                const double2 w8 = bload2(tw, 128 * m);
                v[8] = cmul(v[8], tw_sign(w8, s));
                v[9] = cmul(v[9], tw_sign(cmul(w8, w1), s));
                v[10] = cmul(v[10], tw_sign(cmul(w8, w2), s));
                v[11] = cmul(v[11], tw_sign(cmul(w8, w3), s));
                const double2 w12 = cmul(w8, w4);
                v[12] = cmul(v[12], tw_sign(w12, s));
                v[13] = cmul(v[13], tw_sign(cmul(w12, w1), s));
                v[14] = cmul(v[14], tw_sign(cmul(w12, w2), s));
                v[15] = cmul(v[15], tw_sign(cmul(w12, w3), s));
            }
        }
    }
}

template <int R>
__device__ __forceinline__ void dft_groups(double2 *v, double s)
{
#pragma unroll
    for (int c = 0; c < 16 / R; ++c) dft<R>(v + c * R, s);
}

// waves per SIMD the register allocation must allow: along rows three
// (no spills, three 256-thread workgroups per CU); across transforms four, so
// that two 512-thread workgroups share a CU
#ifndef IFFT_ROWCOL_WPE
#define IFFT_ROWCOL_WPE 3
#endif
#ifndef IFFT_ROW_WPE
#define IFFT_ROW_WPE 3
#endif
#ifndef IFFT_COL_WPE
#define IFFT_COL_WPE 3
#endif
template <int MAP, int IOW> constexpr int pow2_wpe()
{
    if (IOW & IO_W4) return 4;
    return MAP == MAP_ROW ? IFFT_ROW_WPE : (MAP == MAP_COL ? IFFT_COL_WPE : IFFT_ROWCOL_WPE);
}

template <int N, int MAP, int IOW>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(pow2_wpe<MAP, IOW>()))) k_fft_pow2(const FftParams P)
{
    constexpr int IO = IOW & (IO_W4 - 1);
    constexpr bool S32 = IOW & IO_S32, D32 = IOW & IO_D32;
    extern __shared__ double lds[];
    __shared__ TfInfo info[MAP == MAP_ROWCOL ? kMaxTB : 1];
    constexpr bool COL = MAP == MAP_COL;
    constexpr int NT = N / 16;
    constexpr int LDP = pow2_ld<N, MAP>();
    using RS = Pow2Radices<N>;
    constexpr int NST = RS::n;
    constexpr int TB = pow2_tb(N, MAP);
    const int tid = threadIdx.x;
    const int t = COL ? tid % TB : tid / NT;
    const int i = COL ? tid / TB : tid % NT;
    const double s = (double)P.sign;
    long long gidx;
    RowpairGroup RG{};
    if constexpr (IO == IO_R2C_ROWPAIR) {
        RG = rowpair_group(xcd_block(blockIdx.x, gridDim.x), TB, P.ni);
        const long long g = rowpair_transform(RG, t, TB, P.ni);
        gidx = g < 0 ? P.ntf : g;
    } else {
        gidx = (long long)xcd_block(blockIdx.x, gridDim.x) * TB + t;
    }
    const TfInfo T = make_tfinfo(P, gidx);
    const bool valid = T.flags & 1;
    constexpr bool FAST = IO != IO_GENERIC && IO != IO_R2C_HALF;
    // twiddle table through a buffer resource: 32-bit offsets, scalar base
    const __amdgpu_buffer_rsrc_t rtw = make_rsrc(P.tw, (long long)N * 16);

    auto at = [&](int idx) -> double & { return COL ? lds[padx(idx) * TB + t] : lds[t * LDP + padx(idx)]; };

    double2 v[16];
    // ---- stage 0 (radix 16, p = 1): inputs x[i + q N/16] from global memory
    if constexpr (IO == IO_BLUE_IN) {
        // z[pos] = (x[pos] or conj x[pos]) w[pos] for pos < ilim, zero beyond; x complex or the
        // odd real pair xa + i xb.  Every load is issued (position 0 stands in for padding and
        // idle slots) and the zero is selected afterwards, so no load waits on a branch.
        // Positions and byte offsets fit 31 bits (checked at plan time).
        const __amdgpu_buffer_rsrc_t rin = make_rsrc((const char *)P.src + P.src_lo, P.src_span);
        const __amdgpu_buffer_rsrc_t rch = make_rsrc(P.chirp, P.ilim * 16);
        const bool pairs = P.load_mode == LD_R2C_PAIR;
        const int es = (pairs ? 8 : 16) / (S32 ? 2 : 1);   // float / complex float sources: half
        const int ba = (int)(T.ia * es - P.src_lo), bb = (int)(T.ib * es - P.src_lo);
        const int estep = (int)P.in.sax * es;
        const int ilim = (int)P.ilim;
        const double cj = P.conj_in ? -1.0 : 1.0;
        const bool hasb = T.flags & 2;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int pos = (int)T.ip0 + (i + q * (N / 16)) * (int)P.in.pst;
            const bool ok = valid && pos < ilim;
            const int pc = ok ? pos : 0;
            double2 z;
            if constexpr (S32) {
                if (pairs) {
                    const float a = bloadf1(rin, ba + pc * estep), b = bloadf1(rin, bb + pc * estep);
                    z = make_double2((double)a, hasb ? (double)b : 0.0);
                } else {
                    const float2 f = bloadf2(rin, ba + pc * estep);
                    z = make_double2((double)f.x, (double)f.y);
                }
            } else if (pairs) {
                const double a = bload1(rin, ba + pc * estep), b = bload1(rin, bb + pc * estep);
                z = make_double2(a, hasb ? b : 0.0);
            } else {
                z = bload2(rin, ba + pc * estep);
            }
            z.y *= cj;
            const double2 w = bload2(rch, pc * 16);
            v[q] = ok ? cmul(z, w) : make_double2(0.0, 0.0);
        }
    } else if constexpr (FAST) {
        // The r2c row-pair pass reads whole rows that no other workgroup touches: its loads are
        // non-temporal, which keeps the L2 for merging its short strided stores (pass time
        // 0.337 -> 0.330 ms at r2c 2^22 x16).  The column passes read 32- to 128-byte pieces of
        // lines their neighbours read too, and lose 4-17 % with the same hint.  The r2c pass A,
        // whose 128-byte pieces are whole lines, gains 5 % with it (0.243 -> 0.232 ms) but leaves
        // pass B 5 % slower, so it keeps cached loads too.
        constexpr bool NTLOAD = IO == IO_R2C_ROWPAIR;
        // byte offsets inside the source window fit 31 bits (checked at plan time)
        const __amdgpu_buffer_rsrc_t rin = make_rsrc((const char *)P.src + P.src_lo, P.src_span);
        if constexpr (S32) {
            // complex float: 8-byte loads, widened to FP64
            const int step = (int)(P.in.pst * P.in.sax * 8);
            const int base = (int)((T.ia + T.ip0 * P.in.sax) * 8 - P.src_lo) + i * step;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float2 x = bloadf2(rin, base + q * (N / 16) * step);
                v[q] = valid ? make_double2((double)x.x, (double)x.y) : make_double2(0.0, 0.0);
            }
        } else {
            const int step = (int)(P.in.pst * P.in.sax * 16);
            const int base = (int)((T.ia + T.ip0 * P.in.sax) * 16 - P.src_lo) + i * step;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const double2 x = bload2_data<NTLOAD>(rin, base + q * (N / 16) * step);
                v[q] = valid ? x : make_double2(0.0, 0.0);
            }
        }
    } else if (IO == IO_R2C_HALF && P.load_mode == LD_C2C && P.src_span > 0 && P.src_span <= 0x7fffffffLL) {
        // the real-half load of an aligned contiguous row seen as complex words (the plan's
        // alternative parameters): direct 16-byte loads, or 8-byte loads of float pairs
        const __amdgpu_buffer_rsrc_t rin = make_rsrc((const char *)P.src + P.src_lo, P.src_span);
        const int es = P.src32 ? 8 : 16;
        const int step = (int)(P.in.pst * P.in.sax * es);
        const int base = (int)((T.ia + T.ip0 * P.in.sax) * es - P.src_lo) + i * step;
        if (P.src32) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float2 x = bloadf2(rin, base + q * (N / 16) * step);
                v[q] = valid ? make_double2((double)x.x, (double)x.y) : make_double2(0.0, 0.0);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const double2 x = bload2(rin, base + q * (N / 16) * step);
                v[q] = valid ? x : make_double2(0.0, 0.0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int n = i + q * (N / 16);
            v[q] = valid ? load_full(P, T, T.ip0 + (long long)n * P.in.pst) : make_double2(0.0, 0.0);
        }
    }
    dft<16>(v, s);

    // ---- exchanges and the remaining stages (sign sg)
    auto stages = [&](double sg) {
    int p = 16;   // product of the radices applied so far
#pragma unroll
    for (int st = 1; st < NST; ++st) {
        const int R = RS::r[st];
        const int nb = N / R;
        // positions the previous stage wrote (radix 16 for all but the last stage)
        const int kp = (st == 1) ? 0 : i % (p / 16);
        const int jp = (st == 1) ? 16 * i : (i - kp) * 16 + kp;
        const int pp = p / 16;
        // real parts, then imaginary parts
#pragma unroll
        for (int half = 0; half < 2; ++half) {
#pragma unroll
            for (int q = 0; q < 16; ++q) at(jp + q * pp) = half ? v[q].y : v[q].x;
            __syncthreads();
#pragma unroll
            for (int c = 0; c < 16 / R; ++c)
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const double x = at(i + c * NT + q * nb);
                    if (half) v[c * R + q].y = x;
                    else v[c * R + q].x = x;
                }
            __syncthreads();
        }
        // twiddles and butterflies
#pragma unroll
        for (int c = 0; c < 16 / R; ++c) {
            const int b = i + c * NT;
            const int k = b % p;
            switch (R) {
                case 16: apply_twiddles<16>(v + c * R, rtw, k * (N / (p * 16)), sg); break;
                case 8: apply_twiddles<8>(v + c * R, rtw, k * (N / (p * 8)), sg); break;
                case 4: apply_twiddles<4>(v + c * R, rtw, k * (N / (p * 4)), sg); break;
                default: apply_twiddles<2>(v + c * R, rtw, k * (N / (p * 2)), sg); break;
            }
        }
        switch (R) {
            case 16: dft_groups<16>(v, sg); break;
            case 8: dft_groups<8>(v, sg); break;
            case 4: dft_groups<4>(v, sg); break;
            default: dft_groups<2>(v, sg); break;
        }
        p *= R;
    }
    };
    stages(s);
    // the last stage had p_last = N / R_L, so butterfly b = i + c NT (k = b)
    // holds outputs y[b + q N / R_L]
    constexpr int RL = RS::r[NST - 1];
    constexpr int NB = 16 / RL;

    if constexpr (IO == IO_CONV) {
        // Bluestein row pass: y[n] times row j of the transposed kernel spectrum
        // (one table of L = P.tw_L points, 32-bit byte offsets), then back to the
        // stage-0 layout through LDS (real parts, then imaginary parts; the last
        // exchange ended with a barrier) and the transform with the opposite sign
        const __amdgpu_buffer_rsrc_t rsp = make_rsrc(P.mtab, P.tw_L * 16);
        const int sbase = (int)T.j * N;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int b = i + c * NT;
#pragma unroll
            for (int q = 0; q < RL; ++q)
                v[c * RL + q] = cmul(v[c * RL + q], bload2(rsp, (sbase + b + q * (N / RL)) * 16));
        }
#pragma unroll
        for (int half = 0; half < 2; ++half) {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const int b = i + c * NT;
#pragma unroll
                for (int q = 0; q < RL; ++q) at(b + q * (N / RL)) = half ? v[c * RL + q].y : v[c * RL + q].x;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const double x = at(i + q * (N / 16));
                if (half) v[q].y = x;
                else v[q].x = x;
            }
            __syncthreads();
        }
        dft<16>(v, -s);
        stages(-s);
    }

    if constexpr (IO == IO_R2C_HALF) {
        // X[k] and X[M - k] from Z[k] and Z[M - k] (M = N): the lower half of Z goes to LDS
        // (the last stage ended with a barrier); the holder of Z[n], n > N/2, reads Z[N - n] and
        // writes X[N - n] and X[n]; the holders of Z[0] and Z[N/2] write X[0], X[N] and X[N/2].
        double2 *ldc = (double2 *)lds;
        constexpr int H = N / 2;
        constexpr int SH = LDP / 2;
        static_assert(SH >= H + H / 16, "half a transform must fit its LDS row");
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int b = i + c * NT;
#pragma unroll
            for (int q = 0; q < RL; ++q) {
                const int n = b + q * (N / RL);
                if (n < H) ldc[t * SH + padx(n)] = v[c * RL + q];
            }
        }
        __syncthreads();
        if (!valid) return;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int b = i + c * NT;
#pragma unroll
            for (int q = 0; q < RL; ++q) {
                const int n = b + q * (N / RL);
                const double2 z = v[c * RL + q];
                if (n > H) real_split_store(P, T, N - n, ldc[t * SH + padx(N - n)], z);
                else if (n == 0 || n == H) real_split_store(P, T, n, z, z);
            }
        }
        return;
    }
    if constexpr (MAP != MAP_ROWCOL) {
        if (!valid) return;
        if constexpr (FAST) {
            const __amdgpu_buffer_rsrc_t rout = make_rsrc((char *)P.dst + P.dst_lo, P.dst_span);
            // complex float outputs: rounded to nearest, 8-byte stores
            constexpr int oes = D32 ? 8 : 16;
            const int ostep = (int)(P.out.pst * P.out.sax * oes);
            const int obase = (int)((T.oa + T.op0 * P.out.sax) * oes - P.dst_lo);
            auto put = [&](int off, double2 z) {
                if constexpr (D32) bstoref2(rout, off, z);
                else bstore2(rout, off, z);
            };
            constexpr bool twid = IO == IO_TWIDDLE || IO == IO_CONV || IO == IO_BLUE_IN;
            const double so = IO == IO_CONV ? -s : s;   // the row pass stores conj(W^{j n})
            if constexpr (twid) {
                // W_L^{j n}, n = b + q N/RL: W^{j b} times powers of W^{j N/RL}
                // (binary powers, at most four products per twiddle)
                const long long tS = 1LL << P.tw_logS;
                const __amdgpu_buffer_rsrc_t rlo = make_rsrc(P.tw_lo, tS * 16);
                const __amdgpu_buffer_rsrc_t rhi = make_rsrc(P.tw_hi, (P.tw_L / tS + 2) * 16);
                const double2 step1 = twolevel_b(rlo, rhi, P.tw_logS, T.j * (long long)(N / RL));
                double2 pw[4];
                pw[0] = step1;
#pragma unroll
                for (int e = 1; e < 4; ++e) pw[e] = cmul(pw[e - 1], pw[e - 1]);
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const int b = i + c * NT;
                    const double2 base = twolevel_b(rlo, rhi, P.tw_logS, T.j * (long long)b);
#pragma unroll
                    for (int q = 0; q < RL; ++q) {
                        const int n = b + q * (N / RL);
                        double2 w = base;
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (q & (1 << e)) w = cmul(w, pw[e]);
                        put(obase + n * ostep, cmul(v[c * RL + q], tw_sign(w, so)));
                    }
                }
            } else if constexpr (IO == IO_BLUE_OUT) {
                // X[pos] = (c[pos] w[pos], conjugated on request) fct for pos < olim; the chirp
                // loads are issued for every point (position 0 for those past the limit) so that
                // only the stores sit under the branch
                const __amdgpu_buffer_rsrc_t rch = make_rsrc(P.chirp, P.olim * 16);
                const int olim = (int)P.olim, op0 = (int)T.op0, opst = (int)P.out.pst;
                const double cj = P.conj_out ? -1.0 : 1.0;
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const int b = i + c * NT;
#pragma unroll
                    for (int q = 0; q < RL; ++q) {
                        const int n = b + q * (N / RL);
                        const int pos = op0 + n * opst;
                        const double2 w = bload2(rch, (pos < olim ? pos : 0) * 16);
                        double2 z = cmul(v[c * RL + q], w);
                        z.y *= cj;
                        if (pos < olim) put(obase + n * ostep, cscale(z, P.fct));
                    }
                }
            } else {
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const int b = i + c * NT;
#pragma unroll
                    for (int q = 0; q < RL; ++q) {
                        const int n = b + q * (N / RL);
                        put(obase + n * ostep, cscale(v[c * RL + q], P.fct));
                    }
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const int b = i + c * NT;
#pragma unroll
                for (int q = 0; q < RL; ++q) store_value(P, T, b + q * (N / RL), v[c * RL + q]);
            }
        }
    } else if constexpr (IO == IO_R2C_ROWPAIR) {
        // rows j (slots p < cnt) and L2 - j (slots p + TB/2): point n of row j
        // pairs with point L1 - 1 - n of row L2 - j.  Round r keeps half r of the
        // j rows and half 1 - r of their partners in LDS, so every pair of a
        // round is resident.  Then the self-paired rows 0 (slot cnt) and L2/2
        // (slot cnt + TB/2), where the group holds them, keep both halves in LDS
        // slots 0, 1 and 2, 3.
        if (i == 0) info[t] = T;
        double2 *ldc = (double2 *)lds;
        constexpr int H = N / 2;
        constexpr int SH = LDP / 2;
        constexpr int PP = TB / 2;
        constexpr int nthr = TB * NT;
        static_assert(TB >= 4 && TB % 2 == 0, "row pairs and the two self-paired rows need four slots");
        const long long L2 = P.ni;
        const int cnt = RG.cnt;
        __syncthreads();
        if (cnt > 0) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                __syncthreads();
                const int hsel = t < PP ? r : 1 - r;
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const int b = i + c * NT;
#pragma unroll
                    for (int q = 0; q < RL; ++q) {
                        const int n = b + q * (N / RL) - hsel * H;
                        if (n >= 0 && n < H) ldc[t * SH + padx(n)] = v[c * RL + q];
                    }
                }
                __syncthreads();
                // a fixed number of items per thread, unrolled so that the twiddle and LDS
                // loads of all items issue before the first product
                constexpr int ITEMS = (PP * H + nthr - 1) / nthr;
#pragma unroll
                for (int it = 0; it < ITEMS; ++it) {
                    const int e = tid + it * nthr;
                    const int pp = e % PP, nl = e / PP;
                    const TfInfo &U = info[pp];
                    if (e >= PP * H || pp >= cnt || !(U.flags & 1)) continue;
                    const long long k = U.j + L2 * (long long)(r * H + nl);
                    real_split_store(P, U, k, ldc[pp * SH + padx(nl)], ldc[(pp + PP) * SH + padx(H - 1 - nl)]);
                }
            }
        }
        if (RG.self) {
            // row 0 (slot cnt) into LDS slots 0 and 1, row L2/2 (slot cnt + PP) into 2 and 3
            __syncthreads();
            if ((t == cnt || t == cnt + PP) && valid) {
                const int s0 = t == cnt ? 0 : 2;
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const int b = i + c * NT;
#pragma unroll
                    for (int q = 0; q < RL; ++q) {
                        const int n = b + q * (N / RL);
                        ldc[(s0 + (n < H ? 0 : 1)) * SH + padx(n < H ? n : n - H)] = v[c * RL + q];
                    }
                }
            }
            __syncthreads();
            auto zat = [&](int s0, int n) { return n < H ? ldc[s0 * SH + padx(n)] : ldc[(s0 + 1) * SH + padx(n - H)]; };
            // row 0: pairs (n, L1 - n), n = 0 .. L1/2; row L2/2: pairs (n, L1 - 1 - n), n < L1/2
            for (int e = tid; e < N + 1; e += nthr) {
                if (e <= N / 2) {
                    const int n = e;
                    const TfInfo &U = info[cnt];
                    if (U.flags & 1) real_split_store(P, U, L2 * (long long)n, zat(0, n), zat(0, (N - n) % N));
                } else {
                    const int n = e - (N / 2 + 1);
                    const TfInfo &U = info[cnt + PP];
                    if (U.flags & 1)
                        real_split_store(P, U, L2 / 2 + L2 * (long long)n, zat(2, n), zat(2, N - 1 - n));
                }
            }
        }
    } else {
        // transpose through LDS in two halves of the point range (one component
        // row holds half the points as complex values)
        if (i == 0) info[t] = T;
        double2 *ldc = (double2 *)lds;
        constexpr int H = N / 2;
        constexpr int SH = LDP / 2;   // complex words per transform: H + H/16 + 1, odd
        static_assert(SH >= H + H / 16, "half row must fit the component row");
        constexpr int nthr = TB * NT;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const int b = i + c * NT;
#pragma unroll
                for (int q = 0; q < RL; ++q) {
                    const int n = b + q * (N / RL) - half * H;
                    if (n >= 0 && n < H) ldc[t * SH + padx(n)] = v[c * RL + q];
                }
            }
            __syncthreads();
            for (int e = tid; e < TB * H; e += nthr) {
                const int tt = e % TB, n = e / TB;
                const TfInfo &U = info[tt];
                if (!(U.flags & 1)) continue;
                if constexpr (IO == IO_PLAIN) {
                    const __amdgpu_buffer_rsrc_t rout = make_rsrc((char *)P.dst + P.dst_lo, P.dst_span);
                    const long long pos = U.op0 + (long long)(n + half * H) * P.out.pst;
                    const double2 z = cscale(ldc[tt * SH + padx(n)], P.fct);
                    if constexpr (D32) bstoref2(rout, (int)((U.oa + pos * P.out.sax) * 8 - P.dst_lo), z);
                    else bstore2(rout, (int)((U.oa + pos * P.out.sax) * 16 - P.dst_lo), z);
                } else {
                    store_value(P, U, n + half * H, ldc[tt * SH + padx(n)]);
                }
            }
        }
    }
}

template <int N, int IO>
static hipError_t launch_io(const FftParams &P, int map, unsigned grid, unsigned block, size_t lds, hipStream_t s)
{
    // the Bluestein and real-half modes are built for the row and column mappings only (pow2_io)
    constexpr bool ROWCOL = (IO & (IO_W4 - 1)) != IO_BLUE_IN && (IO & (IO_W4 - 1)) != IO_BLUE_OUT &&
                            (IO & (IO_W4 - 1)) != IO_R2C_HALF;
    static bool attr_done = false;
    if (!attr_done) {
        for (const void *f : {(const void *)k_fft_pow2<N, MAP_ROW, IO>, (const void *)k_fft_pow2<N, MAP_COL, IO>})
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192) != hipSuccess)
                (void)hipGetLastError();
        if constexpr (ROWCOL)
            if (hipFuncSetAttribute((const void *)k_fft_pow2<N, MAP_ROWCOL, IO>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192) != hipSuccess)
                (void)hipGetLastError();
        attr_done = true;
    }
    switch (map) {
        case MAP_COL: hipLaunchKernelGGL((k_fft_pow2<N, MAP_COL, IO>), dim3(grid), dim3(block), lds, s, P); break;
        case MAP_ROWCOL:
            if constexpr (ROWCOL) {
                hipLaunchKernelGGL((k_fft_pow2<N, MAP_ROWCOL, IO>), dim3(grid), dim3(block), lds, s, P);
                break;
            } else {
                return hipErrorInvalidValue;
            }
        default: hipLaunchKernelGGL((k_fft_pow2<N, MAP_ROW, IO>), dim3(grid), dim3(block), lds, s, P); break;
    }
    return hipGetLastError();
}

// the fused real post-process runs on the row-load / column-store mapping
// with two (N = 2048) or four (N = 1024) row pairs per workgroup
template <int N>
static hipError_t launch_rowpair(const FftParams &P, unsigned grid, unsigned block, size_t lds, hipStream_t s)
{
    if constexpr (N == 1024 || N == 2048) {
        static bool attr_done = false;
        if (!attr_done) {
            if (hipFuncSetAttribute((const void *)k_fft_pow2<N, MAP_ROWCOL, IO_R2C_ROWPAIR>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192) != hipSuccess)
                (void)hipGetLastError();
            attr_done = true;
        }
        hipLaunchKernelGGL((k_fft_pow2<N, MAP_ROWCOL, IO_R2C_ROWPAIR>), dim3(grid), dim3(block), lds, s, P);
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
}

// the four-step Bluestein row pass (transform, spectrum product, inverse transform)
// on the row mapping
template <int N, int IOW>
static hipError_t launch_conv(const FftParams &P, unsigned grid, unsigned block, size_t lds, hipStream_t s)
{
    static bool attr_done = false;
    if (!attr_done) {
        if (hipFuncSetAttribute((const void *)k_fft_pow2<N, MAP_ROW, IOW>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192) != hipSuccess)
            (void)hipGetLastError();
        attr_done = true;
    }
    hipLaunchKernelGGL((k_fft_pow2<N, MAP_ROW, IOW>), dim3(grid), dim3(block), lds, s, P);
    return hipGetLastError();
}

template <int N>
static hipError_t launch_n(const FftParams &P, int map, int io, unsigned grid, unsigned block, size_t lds,
                           hipStream_t s)
{
    if (io == IO_R2C_ROWPAIR) return map == MAP_ROWCOL ? launch_rowpair<N>(P, grid, block, lds, s) : hipErrorInvalidValue;
    if ((io & (IO_W4 - 1)) == IO_CONV) {
        if (map != MAP_ROW) return hipErrorInvalidValue;
        return (io & IO_W4) ? launch_conv<N, IO_CONV | IO_W4>(P, grid, block, lds, s)
                            : launch_conv<N, IO_CONV>(P, grid, block, lds, s);
    }
    switch (io) {
        case IO_PLAIN: return launch_io<N, IO_PLAIN>(P, map, grid, block, lds, s);
        case IO_TWIDDLE: return launch_io<N, IO_TWIDDLE>(P, map, grid, block, lds, s);
        case IO_BLUE_IN: return launch_io<N, IO_BLUE_IN>(P, map, grid, block, lds, s);
        case IO_R2C_HALF: return launch_io<N, IO_R2C_HALF>(P, map, grid, block, lds, s);
        case IO_BLUE_OUT: return launch_io<N, IO_BLUE_OUT>(P, map, grid, block, lds, s);
        case IO_BLUE_IN | IO_W4: return launch_io<N, IO_BLUE_IN | IO_W4>(P, map, grid, block, lds, s);
        case IO_BLUE_IN | IO_S32: return launch_io<N, IO_BLUE_IN | IO_S32>(P, map, grid, block, lds, s);
        case IO_BLUE_IN | IO_S32 | IO_W4: return launch_io<N, IO_BLUE_IN | IO_S32 | IO_W4>(P, map, grid, block, lds, s);
        case IO_BLUE_OUT | IO_D32: return launch_io<N, IO_BLUE_OUT | IO_D32>(P, map, grid, block, lds, s);
        case IO_BLUE_OUT | IO_D32 | IO_W4: return launch_io<N, IO_BLUE_OUT | IO_D32 | IO_W4>(P, map, grid, block, lds, s);
        case IO_PLAIN | IO_S32: return launch_io<N, IO_PLAIN | IO_S32>(P, map, grid, block, lds, s);
        case IO_PLAIN | IO_D32: return launch_io<N, IO_PLAIN | IO_D32>(P, map, grid, block, lds, s);
        case IO_PLAIN | IO_S32 | IO_D32: return launch_io<N, IO_PLAIN | IO_S32 | IO_D32>(P, map, grid, block, lds, s);
        case IO_TWIDDLE | IO_S32: return launch_io<N, IO_TWIDDLE | IO_S32>(P, map, grid, block, lds, s);
        case IO_BLUE_OUT | IO_W4: return launch_io<N, IO_BLUE_OUT | IO_W4>(P, map, grid, block, lds, s);
        default: return launch_io<N, IO_GENERIC>(P, map, grid, block, lds, s);
    }
}

bool pow2_supported(int N) { return N >= 64 && N <= 8192 && (N & (N - 1)) == 0; }
int pow2_tb_host(int N, int map) { return pow2_tb(N, map); }
// LDS bytes per transform (one padded component row of doubles)
int pow2_lds_bytes(int N, int map) { return (N + N / 16 + (map == MAP_ROWCOL ? 2 : 1)) * 8; }

// plain complex loads and plain / twiddled complex stores (complex double or complex float
// arrays) take the FAST path when no position can reach the zero-padding or store limits
bool pow2_fast(const FftParams &P)
{
    const long long imax = (P.ni - 1) * P.in.ipos + (long long)(P.N - 1) * P.in.pst;
    const long long omax = (P.ni - 1) * P.out.ipos + (long long)(P.N - 1) * P.out.pst;
    const long long win = 0x7fffffffLL;   // buffer windows are addressed with 31-bit byte offsets
    return P.load_mode == LD_C2C && !P.conj_in && !P.chirp_in && imax < P.ilim &&
           (P.store_mode == ST_C2C || (P.store_mode == ST_TWIDDLE && P.tw_mul == 1)) && !P.chirp_out &&
           !P.conj_out && omax < P.olim && P.src_span > 0 && P.src_span <= win && P.dst_span > 0 &&
           P.dst_span <= win;
}

// plain complex loads inside a 31-bit byte window (the fused row-pair store
// takes only this load form)
bool pow2_fast_load(const FftParams &P)
{
    const long long imax = (P.ni - 1) * P.in.ipos + (long long)(P.N - 1) * P.in.pst;
    return P.load_mode == LD_C2C && !P.src32 && !P.conj_in && !P.chirp_in && imax < P.ilim && P.src_span > 0 &&
           P.src_span <= 0x7fffffffLL;
}

// the Bluestein column passes: f64 data, every position and byte offset within 31 bits
static bool blue_windows(const FftParams &P)
{
    const long long win = 0x7fffffffLL;
    const long long imax = (P.ni - 1) * P.in.ipos + (long long)(P.N - 1) * P.in.pst;
    const long long omax = (P.ni - 1) * P.out.ipos + (long long)(P.N - 1) * P.out.pst;
    return P.chirp && P.src_span > 0 && P.src_span <= win && P.dst_span > 0 &&
           P.dst_span <= win && imax >= 0 && omax >= 0 && (imax + 1) * 16 <= win && (omax + 1) * 16 <= win &&
           P.in.ipos >= 0 && P.in.pst > 0 && P.out.ipos >= 0 && P.out.pst > 0;
}

// chirp-in loads (complex or odd real pairs) with the plain four-step twiddle store
static bool pow2_blue_in(const FftParams &P)
{
    const long long omax = (P.ni - 1) * P.out.ipos + (long long)(P.N - 1) * P.out.pst;
    return blue_windows(P) && !P.dst32 && P.chirp_in && (P.load_mode == LD_C2C || P.load_mode == LD_R2C_PAIR) &&
           P.ilim <= 0x7fffffffLL / 16 && P.store_mode == ST_TWIDDLE && P.tw_mul == 1 && !P.chirp_out &&
           !P.conj_out && omax < P.olim;
}

// plain loads with the chirp-out complex store
static bool pow2_blue_out(const FftParams &P)
{
    return blue_windows(P) && pow2_fast_load(P) && P.chirp_out && P.olim <= 0x7fffffffLL / 16 &&
           P.store_mode == ST_C2C;
}

// compute units of the current device (read once per device)
static int device_cus()
{
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return 256;
    }
    if (!cus[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) {
            (void)hipGetLastError();
            v = 256;
        }
        cus[dev] = v;
    }
    return cus[dev];
}

// A grid of at most four rounds of resident workgroups (three waves per SIMD) runs its
// Bluestein passes on the four-wave instances: a second workgroup per CU overlaps one's loads
// with the other's stages, which the register spills of those instances do not outweigh.  On
// r2c 1,000,003 x4 the three passes took 0.164 -> 0.155 ms; on larger batches (c2c 512 x
// 60,013: 1.50 -> 1.78 ms) the spills cost more than the overlap gains.
static bool few_rounds(unsigned grid, unsigned block, size_t lds)
{
    const int waves = (int)((block + 63) / 64);
    int per_cu = 12 / std::max(1, waves);
    const size_t lds_wg = lds + 2048;   // the kernel's static shared memory
    per_cu = std::min<int>(per_cu, (int)((size_t)160 * 1024 / lds_wg));
    return per_cu >= 1 && (long long)grid <= 4LL * per_cu * device_cus();
}

int pow2_io(const FftParams &P, int map, unsigned grid, unsigned block, size_t lds)
{
    if (P.store_mode == ST_R2C_ROWPAIR) return IO_R2C_ROWPAIR;
    if (P.store_mode == ST_R2C_HALF) return map == MAP_ROWCOL ? IO_GENERIC : IO_R2C_HALF;   // plan routes rows only
    int io = IO_GENERIC;
    if (P.store_mode == ST_CONV_TWIDDLE) io = IO_CONV;
    else if (pow2_fast(P)) {
        io = (P.store_mode == ST_TWIDDLE ? IO_TWIDDLE : IO_PLAIN) | (P.src32 ? IO_S32 : 0) | (P.dst32 ? IO_D32 : 0);
        // the twiddled store writes a workspace: no complex float instance for it
        return (io & IO_D32) && (io & (IO_W4 - 1)) == IO_TWIDDLE ? IO_GENERIC : io;
    }
    else if (map != MAP_ROWCOL && pow2_blue_in(P)) io = IO_BLUE_IN | (P.src32 ? IO_S32 : 0);
    else if (map != MAP_ROWCOL && pow2_blue_out(P)) io = IO_BLUE_OUT | (P.dst32 ? IO_D32 : 0);
    else return IO_GENERIC;
    return few_rounds(grid, block, lds) ? (io | IO_W4) : io;
}

// workgroups of the fused row-pair pass: the pair groups of every signal, the self-paired rows
// in the last one's free slot or in one more group
long long pow2_rowpair_grid(long long signals, long long L2, int TB)
{
    return signals * rowpair_groups(L2, TB);
}

hipError_t launch_fft_pow2(const FftParams &P, int map, unsigned grid, unsigned block, size_t lds, hipStream_t s)
{
    const int fast = pow2_io(P, map, grid, block, lds);
    switch (P.N) {
        case 64: return launch_n<64>(P, map, fast, grid, block, lds, s);
        case 128: return launch_n<128>(P, map, fast, grid, block, lds, s);
        case 256: return launch_n<256>(P, map, fast, grid, block, lds, s);
        case 512: return launch_n<512>(P, map, fast, grid, block, lds, s);
        case 1024: return launch_n<1024>(P, map, fast, grid, block, lds, s);
        case 2048: return launch_n<2048>(P, map, fast, grid, block, lds, s);
        case 4096: return launch_n<4096>(P, map, fast, grid, block, lds, s);
        case 8192: return launch_n<8192>(P, map, fast, grid, block, lds, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ifft

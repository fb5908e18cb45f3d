// A mixed-radix Stockham FFT of runtime length on data in LDS, run by the
// threads of a workgroup: radices 2, 3, 4, 5, 7, 8, 11, 13, 16, FP64.  The
// twiddles come from sincospi (TwSincos: the ragged-batch kernel, whose
// signals all have different lengths) or from a table W_L^j, j < L, of the
// transform length (TwTable: the fused DCT and r2r_fftpack kernels, whose
// lines share one length).
#pragma once

#include "butterflies.h"

namespace ifft {

// Every configuration keeps lds_points <= 8 threads, so a thread owns at most
// 8 points of any LDS batch and ceil(8 / R) butterflies of a radix-R stage.
constexpr int kPointsPerThread = 8;
__host__ __device__ constexpr int nb_rg(int R) { return (kPointsPerThread + R - 1) / R; }

// exp(sign 2 pi i k / L), 0 <= k < L < 2^31
__device__ __forceinline__ double2 twid32(int k, int L, double sign)
{
    double s, c;
    sincospi(2.0 * (double)k / (double)L, &s, &c);
    return make_double2(c, sign * s);
}

struct TwSincos {
    // exp(sign 2 pi i k / n)
    __device__ __forceinline__ double2 operator()(int k, int n, double sign) const { return twid32(k, n, sign); }
};

struct TwTable {
    const double2 *tab;   // W_L^j = exp(-2 pi i j / L), j < L (forward sign)
    int L;
    // exp(sign 2 pi i k / n) = W_L^(k L / n), conjugated for the backward sign; n divides L
    __device__ __forceinline__ double2 operator()(int k, int n, double sign) const
    {
        const double2 w = tab[k * (L / n)];
        return make_double2(w.x, -sign * w.y);
    }
};

// LDS holds transform data at padded positions: one word skipped every 16, so
// that the power-of-two strides of the stages (16 b, 16 p) spread over the banks.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
constexpr int lds_points_padded(int n) { return n + n / 16 + 1; }

// Stockham stage on one transform in LDS: butterflies b = i, i + NT, ...
// The twiddle roots are generated (or their table loads issued) before the
// values are loaded, so that work does not hold the R NB values live.
// SCALAR: read N, p and NT (the same for every thread of the workgroup: the transforms of one
// LDS batch share their length) as scalars.  Inlined into a kernel that keeps the stage loop
// in scalar registers this took the fused DCT rows from 0.118 to 0.083 ms and the convolution
// from 1.96 to 1.80 ms; in the stage calls of the ragged kernel it cost 1.4 %, so they skip it.
template <int R, int NB, typename TW, bool SCALAR = false>
__device__ __forceinline__ void rstage_body(double2 *lds, int base, int N, int p, int i, int NT, double s,
                                            const TW &tw)
{
    if constexpr (SCALAR) {
        N = __builtin_amdgcn_readfirstlane(N);
        p = __builtin_amdgcn_readfirstlane(p);
        NT = __builtin_amdgcn_readfirstlane(NT);
    }
    const int nb = N / R;
    const float rp = __builtin_amdgcn_rcpf((float)p);
    int kk[NB];
    double2 w1[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int k = mod_small(i + c * NT, p, rp);
        kk[c] = k;
        w1[c] = k ? tw(k, p * R, s) : make_double2(1.0, 0.0);
    }
    double2 v[NB][R];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int b = i + c * NT;
        if (b < nb)
#pragma unroll
            for (int q = 0; q < R; ++q) v[c][q] = lds[lpad(base + b + q * nb)];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int b = i + c * NT;
        if (b < nb) {
            const int k = kk[c];
            if (k) {
                double2 w = w1[c];
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    v[c][q] = cmul(v[c][q], w);
                    if (q + 1 < R) w = cmul(w, w1[c]);
                }
            }
            dft<R>(v[c], s);
            const int j = (b - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lpad(base + j + q * p)] = v[c][q];
        }
    }
    __syncthreads();
}

// the stage as a call of its own (one copy per radix in the ragged kernel, which keeps its
// register budget to the largest stage)
template <int R, int NB, typename TW>
__device__ __noinline__ void rstage(double2 *lds, int base, int N, int p, int i, int NT, double s, TW tw)
{
    rstage_body<R, NB>(lds, base, N, p, i, NT, s, tw);
}

// Radix sequence of a 13-smooth L as 4-bit codes (stage 0 in the low bits),
// in registers.  Order bit 0: odd radices ascending after the powers of two
// (0) or descending before them (1).  Order bit 1: powers of two as 16s and one
// 2/4/8 (0), or as 8s and one 2/4 (1), which keeps every thread of a
// kPointsPerThread = 8 layout busy in those stages at the cost of more stages.
__device__ __forceinline__ unsigned long long radix_codes(int L, int order, int &ns)
{
    // code -> radix: 1:2 2:3 3:4 4:5 5:7 6:8 7:11 8:13 9:16
    unsigned long long pw = 0, od = 0;
    int np = 0, no = 0;
    int n = L;
    int twos = __builtin_ctz((unsigned)n);
    n >>= twos;
    if (order & 2) {
        for (; twos >= 3; twos -= 3) pw |= 6ull << (4 * np++);
        if (twos) pw |= (unsigned long long)(twos == 2 ? 3 : 1) << (4 * np++);
    } else {
        for (; twos >= 4; twos -= 4) pw |= 9ull << (4 * np++);
        if (twos) pw |= (unsigned long long)(twos == 3 ? 6 : twos == 2 ? 3 : 1) << (4 * np++);
    }
    const int ps[5] = {3, 5, 7, 11, 13}, cs[5] = {2, 4, 5, 7, 8};
#pragma unroll
    for (int q = 0; q < 5; ++q)
        while (n % ps[q] == 0) {
            n /= ps[q];
            if (order & 1) od = (od << 4) | (unsigned long long)cs[q];   // descending: latest first
            else od |= (unsigned long long)cs[q] << (4 * no);
            ++no;
        }
    ns = np + no;
    return (order & 1) ? (od | (pw << (4 * no))) : (pw | (od << (4 * np)));
}

// `count` transforms of length L at lds + t L, blockDim.x / count threads each
// (count divides blockDim.x).  Ends with a barrier.
template <typename TW>
__device__ __noinline__ void lds_fft(double2 *lds, int L, int count, double s, int order, TW tw)
{
    int ns;
    unsigned long long codes = radix_codes(L, order, ns);
    const int NT = blockDim.x / count;
    const int t = threadIdx.x / NT, i = threadIdx.x % NT;
    const int base = t * L;
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
            default: rstage<16, nb_rg(16)>(lds, base, L, p, i, NT, s, tw); p *= 16; break;
        }
    }
    if (ns == 0) __syncthreads();
}

// The same transform with every stage inlined into the caller: no call frames,
// so no per-lane stack in scratch memory.  A call's frame (return address and
// the callee-saved registers, 150-250 bytes a lane) lives in scratch, and with
// thousands of resident waves those frames outgrow the L2 and are written back
// on every stage -- more traffic than the transform's data in a streaming
// kernel.  PPT bounds the points a thread owns (count L <= PPT blockDim.x), so a
// radix-R stage takes ceil(PPT / R) butterflies per thread.
template <int PPT>
__host__ __device__ constexpr int nb_ppt(int R) { return (PPT + R - 1) / R; }

// SMALL: only radices 2, 3, 4, 5, 7 and 8 are compiled in, for 7-smooth lengths
// with the radix-8 order (order bit 1); the largest stage then holds at most 14
// values a thread, which keeps a 1,024-thread workgroup within 128 registers.
// The caller guarantees the length (no other radix may occur).
template <int PPT, bool SMALL, typename TW>
__device__ __forceinline__ void lds_fft_inline(double2 *lds, int L, int count, double s, int order, TW tw)
{
    int ns;
    unsigned long long codes = radix_codes(L, order, ns);
    const int NT = blockDim.x / count;
    const int t = threadIdx.x / NT, i = threadIdx.x % NT;
    const int base = t * L;
    int p = 1;
    for (int st = 0; st < ns; ++st, codes >>= 4) {
        const int code = (int)(codes & 15);
        if constexpr (SMALL) {
            switch (code) {
                case 1: rstage_body<2, nb_ppt<PPT>(2), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 2; break;
                case 2: rstage_body<3, nb_ppt<PPT>(3), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 3; break;
                case 3: rstage_body<4, nb_ppt<PPT>(4), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 4; break;
                case 4: rstage_body<5, nb_ppt<PPT>(5), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 5; break;
                case 5: rstage_body<7, nb_ppt<PPT>(7), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 7; break;
                default: rstage_body<8, nb_ppt<PPT>(8), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 8; break;
            }
        } else {
            switch (code) {
                case 1: rstage_body<2, nb_ppt<PPT>(2), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 2; break;
                case 2: rstage_body<3, nb_ppt<PPT>(3), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 3; break;
                case 3: rstage_body<4, nb_ppt<PPT>(4), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 4; break;
                case 4: rstage_body<5, nb_ppt<PPT>(5), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 5; break;
                case 5: rstage_body<7, nb_ppt<PPT>(7), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 7; break;
                case 6: rstage_body<8, nb_ppt<PPT>(8), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 8; break;
                case 7: rstage_body<11, nb_ppt<PPT>(11), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 11; break;
                case 8: rstage_body<13, nb_ppt<PPT>(13), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 13; break;
                default: rstage_body<16, nb_ppt<PPT>(16), TW, true>(lds, base, L, p, i, NT, s, tw); p *= 16; break;
            }
        }
    }
    if (ns == 0) __syncthreads();
}

}  // namespace ifft

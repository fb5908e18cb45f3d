// 16-point DFTs of FP64 data as exact bf16-slice products on the matrix cores
// (v_mfma_f32_16x16x32_bf16), shared by the matrix-core FFT kernels and the
// ragged-batch kernel.
//
// A wavefront computes 16 DFTs of 16 points at once.  Lane l = 16 g + j holds
// inputs q = 4 g + e (e = 0..3) of DFT j in a[e]; after bx_stage it holds
// output j of DFTs 4 g + r in a[r].  In matrix terms the DFTs are the rows of a
// 16 x 32 real A ([Re x, Im x] per row) times a 32 x 16 real B ([Re F; -Im F]
// for the real part, [Im F; Re F] for the imaginary part); the MFMA lane
// layout puts row l mod 16 of A, column l mod 16 of B and D[4 (l / 16) + r][l
// mod 16] in lane l.
//
// Exactness: before the product the wavefront's values are scaled by one power
// of two, 2^(7 - E) with 2^E > max |v|, and cut into kSlices balanced 8-bit
// integer digits (each exact in bf16); F 2^7 is cut the same way.  A slice
// product is an integer below 2^14, the slice pairs of one weight 2^(-8 (a + f))
// accumulate into one FP32 accumulator whose 32-term sums stay below 2^24, so
// the MFMA accumulation is exact in any order; pairs with a + f <= kSlices - 1
// are kept (truncation ~2^-40 of the maximum) and combined in FP64.
#pragma once

#include "butterflies.h"

namespace ifft {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

constexpr int kSlices = 5;

// exp(-2 pi i m / 16) for m < 16 (forward sign), exact to FP64 rounding
__device__ __forceinline__ double2 w16(int m)
{
    constexpr double c1 = 0.92387953251128674, s1 = 0.38268343236508978, h = 0.70710678118654752;
    const double cs[5] = {1.0, c1, h, s1, 0.0};
    const int o = m & 15;
    // cos and sin of 2 pi o / 16 from the first quadrant
    const int q = o >> 2, r = o & 3;
    double c = cs[r], s = cs[4 - r];
    double cr, sr;
    switch (q) {
        case 0: cr = c; sr = s; break;
        case 1: cr = -s; sr = c; break;
        case 2: cr = -c; sr = -s; break;
        default: cr = s; sr = -c; break;
    }
    return make_double2(cr, -sr);
}

struct BxConsts {
    bf8 br[kSlices], bi[kSlices];   // slices of [Re F; -Im F] and [Im F; Re F], scaled by 2^7
};

// bf16 of two floats that are integers of magnitude <= 256 (exact: their low
// 16 bits are zero), packed lo | hi << 16
__device__ __forceinline__ unsigned pk_bf16(float lo, float hi)
{
    return __builtin_amdgcn_perm(__float_as_uint(hi), __float_as_uint(lo), 0x07060302u);
}

// Next slice of v in [-128, 128]: t = rint(v) (an integer in [-128, 128]) and
// v <- 256 (v - t), exact in FP64.
__device__ __forceinline__ float take(double &v)
{
    const double t = rint(v);
    v = fma(v, 256.0, -256.0 * t);
    return (float)t;
}

// the slices of four values as bf16 in elements e0 .. e0+3 of each A[s] / B[s]
__device__ __forceinline__ void cut4(double v0, double v1, double v2, double v3, unsigned (&lo)[kSlices],
                                     unsigned (&hi)[kSlices])
{
#pragma unroll
    for (int a = 0; a < kSlices; ++a) {
        const float f0 = take(v0), f1 = take(v1), f2 = take(v2), f3 = take(v3);
        lo[a] = pk_bf16(f0, f1);
        hi[a] = pk_bf16(f2, f3);
    }
}

__device__ __forceinline__ BxConsts bx_consts(double sign)
{
    BxConsts C;
    const int l = threadIdx.x & 63, g = l >> 4, j = l & 15;
    double fr[4], fi[4];
    for (int e = 0; e < 4; ++e) {
        const double2 w = w16((4 * g + e) * j);
        fr[e] = w.x * 128.0;
        fi[e] = -sign * w.y * 128.0;
    }
    unsigned rl[kSlices], rh[kSlices], il[kSlices], ih[kSlices], nl[kSlices], nh[kSlices];
    cut4(fr[0], fr[1], fr[2], fr[3], rl, rh);
    cut4(fi[0], fi[1], fi[2], fi[3], il, ih);
    cut4(-fi[0], -fi[1], -fi[2], -fi[3], nl, nh);
    for (int f = 0; f < kSlices; ++f) {
        C.br[f] = __builtin_bit_cast(bf8, (u4){rl[f], rh[f], nl[f], nh[f]});
        C.bi[f] = __builtin_bit_cast(bf8, (u4){il[f], ih[f], rl[f], rh[f]});
    }
    return C;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ void bx_stage(double2 (&a)[4], const BxConsts &C)
{
    double mx = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) mx = fmax(mx, fmax(fabs(a[r].x), fabs(a[r].y)));
    mx = wave_max(mx);
    if (mx == 0.0) return;   // an all-zero transform stays zero
    int E;
    (void)frexp(mx, &E);     // mx < 2^E
    const double up = ldexp(1.0, 7 - E);
    bf8 A[kSlices];
    {
        unsigned rl[kSlices], rh[kSlices], il[kSlices], ih[kSlices];
        cut4(a[0].x * up, a[1].x * up, a[2].x * up, a[3].x * up, rl, rh);
        cut4(a[0].y * up, a[1].y * up, a[2].y * up, a[3].y * up, il, ih);
#pragma unroll
        for (int s = 0; s < kSlices; ++s) A[s] = __builtin_bit_cast(bf8, (u4){rl[s], rh[s], il[s], ih[s]});
    }
    f4 accr[kSlices], acci[kSlices];
#pragma unroll
    for (int d = 0; d < kSlices; ++d) {
        accr[d] = f4{0.f, 0.f, 0.f, 0.f};
        acci[d] = f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int s = 0; s < kSlices; ++s)
#pragma unroll
        for (int f = 0; f + s < kSlices; ++f) {
            accr[s + f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[s], C.br[f], accr[s + f], 0, 0, 0);
            acci[s + f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[s], C.bi[f], acci[s + f], 0, 0, 0);
        }
    const double down = ldexp(1.0, E - 14);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int d = kSlices - 1; d >= 0; --d) {
            re = re * (1.0 / 256.0) + (double)accr[d][r];
            im = im * (1.0 / 256.0) + (double)acci[d][r];
        }
        a[r] = make_double2(re * down, im * down);
    }
}

// bx_consts of both signs for the 64 lanes in a table, so that kernels that run
// the stage now and then load them instead of cutting them: entry
// (si 64 + lane) kBxWords + f, si = 0 for the forward sign (-1) and 1 for the
// backward sign; words f < kSlices hold br, the next kSlices bi.
constexpr int kBxWords = 2 * kSlices;
constexpr int kBxTableWords = 2 * 64 * kBxWords;

__device__ __forceinline__ void bx_consts_store(u4 *tab)   // one wavefront
{
    const int l = threadIdx.x & 63;
    for (int si = 0; si < 2; ++si) {
        const BxConsts C = bx_consts(si ? 1.0 : -1.0);
        u4 *t = tab + (si * 64 + l) * kBxWords;
#pragma unroll
        for (int f = 0; f < kSlices; ++f) {
            t[f] = __builtin_bit_cast(u4, C.br[f]);
            t[kSlices + f] = __builtin_bit_cast(u4, C.bi[f]);
        }
    }
}

__device__ __forceinline__ BxConsts bx_consts_load(const u4 *tab, double sign)
{
    const u4 *t = tab + ((sign > 0.0 ? 64 : 0) + (int)(threadIdx.x & 63)) * kBxWords;
    BxConsts C;
#pragma unroll
    for (int f = 0; f < kSlices; ++f) {
        C.br[f] = __builtin_bit_cast(bf8, t[f]);
        C.bi[f] = __builtin_bit_cast(bf8, t[kSlices + f]);
    }
    return C;
}

}  // namespace ifft

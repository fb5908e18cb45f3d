// Register-resident DFT butterflies for the Stockham stages (FP64, gfx950).
//
// `s` is the transform sign as a double: -1 for exp(-2 pi i jk/R) (forward),
// +1 for the backward direction.  rot(t, s) multiplies t by s*i.  Radix 4, 8
// and 16 are split into radix-4 / radix-2 steps with exact rotations; odd
// primes use the symmetric-pair form (cosine sums on v_q + v_{R-q}, sine sums
// on v_q - v_{R-q}), whose constants are cos/sin(2 pi k / R) rounded from
// extended precision.
#pragma once

#include <hip/hip_runtime.h>

namespace ifft {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double f) { return make_double2(a.x * f, a.y * f); }
// t * (s i)
__device__ __forceinline__ double2 rot(double2 t, double s) { return make_double2(-s * t.y, s * t.x); }
// twiddle stored with forward sign; conjugated for s = +1
__device__ __forceinline__ double2 tw_sign(double2 w, double s) { return make_double2(w.x, -s * w.y); }

__device__ __forceinline__ void dft2(double2 &a, double2 &b)
{
    double2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

__device__ __forceinline__ void dft4(double2 &v0, double2 &v1, double2 &v2, double2 &v3, double s)
{
    double2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
    double2 t2 = cadd(v1, v3), t3 = rot(csub(v1, v3), s);
    v0 = cadd(t0, t2);
    v2 = csub(t0, t2);
    v1 = cadd(t1, t3);
    v3 = csub(t1, t3);
}

constexpr double kSqrtHalf = 0.7071067811865476;
constexpr double kCos8 = 0.9238795325112867;   // cos(pi/8)
constexpr double kSin8 = 0.3826834323650898;   // sin(pi/8)

__device__ __forceinline__ void dft8(double2 *v, double s)
{
    dft4(v[0], v[2], v[4], v[6], s);
    dft4(v[1], v[3], v[5], v[7], s);
    // odd half times W8^k, k = 1..3
    double2 o1 = cmul(v[3], make_double2(kSqrtHalf, s * kSqrtHalf));
    double2 o2 = rot(v[5], s);
    double2 o3 = cmul(v[7], make_double2(-kSqrtHalf, s * kSqrtHalf));
    double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

__device__ __forceinline__ void dft16(double2 *v, double s)
{
    // columns j = 0..3: DFT4 over v[j + 4m]
    dft4(v[0], v[4], v[8], v[12], s);
    dft4(v[1], v[5], v[9], v[13], s);
    dft4(v[2], v[6], v[10], v[14], s);
    dft4(v[3], v[7], v[11], v[15], s);
    // twiddles W16^{j k1}: element (j, k1) lives at v[j + 4 k1]
    v[5] = cmul(v[5], make_double2(kCos8, s * kSin8));             // e=1
    v[9] = cmul(v[9], make_double2(kSqrtHalf, s * kSqrtHalf));     // e=2
    v[13] = cmul(v[13], make_double2(kSin8, s * kCos8));           // e=3
    v[6] = cmul(v[6], make_double2(kSqrtHalf, s * kSqrtHalf));     // e=2
    v[10] = rot(v[10], s);                                         // e=4
    v[14] = cmul(v[14], make_double2(-kSqrtHalf, s * kSqrtHalf));  // e=6
    v[7] = cmul(v[7], make_double2(kSin8, s * kCos8));             // e=3
    v[11] = cmul(v[11], make_double2(-kSqrtHalf, s * kSqrtHalf));  // e=6
    v[15] = cmul(v[15], make_double2(-kCos8, -s * kSin8));         // e=9
    // rows k1: DFT4 over j, output y[k1 + 4 k2]
    double2 r[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        double2 a = v[4 * k1], b = v[4 * k1 + 1], c = v[4 * k1 + 2], d = v[4 * k1 + 3];
        dft4(a, b, c, d, s);
        r[k1] = a; r[k1 + 4] = b; r[k1 + 8] = c; r[k1 + 12] = d;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = r[q];
}

template <int R> struct OddConst;
template <> struct OddConst<3> {
    static constexpr double c[3] = {1.0, -0.5, -0.5};
    static constexpr double s[3] = {0.0, 0.8660254037844386, -0.8660254037844386};
};
template <> struct OddConst<5> {
    static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749475, -0.8090169943749475,
                                    0.30901699437494745};
    static constexpr double s[5] = {0.0, 0.9510565162951535, 0.5877852522924731, -0.5877852522924731,
                                    -0.9510565162951535};
};
template <> struct OddConst<7> {
    static constexpr double c[7] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191,
                                    -0.9009688679024191, -0.2225209339563144, 0.6234898018587335};
    static constexpr double s[7] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581,
                                    -0.4338837391175581, -0.9749279121818236, -0.7818314824680298};
};
template <> struct OddConst<11> {
    static constexpr double c[11] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                                     -0.6548607339452851, -0.9594929736144974, -0.9594929736144974,
                                     -0.6548607339452851, -0.14231483827328514, 0.41541501300188644,
                                     0.8412535328311812};
    static constexpr double s[11] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                                     0.7557495743542583, 0.28173255684142967, -0.28173255684142967,
                                     -0.7557495743542583, -0.9898214418809327, -0.9096319953545183,
                                     -0.5406408174555976};
};
template <> struct OddConst<13> {
    static constexpr double c[13] = {1.0, 0.8854560256532099, 0.5680647467311558, 0.12053668025532305,
                                     -0.3546048870425356, -0.7485107481711011, -0.970941817426052,
                                     -0.970941817426052, -0.7485107481711011, -0.3546048870425356,
                                     0.12053668025532305, 0.5680647467311558, 0.8854560256532099};
    static constexpr double s[13] = {0.0, 0.46472317204376856, 0.8229838658936564, 0.992708874098054,
                                     0.9350162426854148, 0.6631226582407952, 0.23931566428755777,
                                     -0.23931566428755777, -0.6631226582407952, -0.9350162426854148,
                                     -0.992708874098054, -0.8229838658936564, -0.46472317204376856};
};

template <int R>
__device__ __forceinline__ void dft_odd(double2 *v, double s)
{
    constexpr int H = (R - 1) / 2;
    double2 a[H + 1], b[H + 1];
    double2 y0 = v[0];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        a[q] = cadd(v[q], v[R - q]);
        b[q] = csub(v[q], v[R - q]);
        y0 = cadd(y0, a[q]);
    }
    double2 out[R];
    out[0] = y0;
#pragma unroll
    for (int m = 1; m <= H; ++m) {
        double2 A = v[0], B = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            const int idx = (q * m) % R;
            A.x = fma(a[q].x, OddConst<R>::c[idx], A.x);
            A.y = fma(a[q].y, OddConst<R>::c[idx], A.y);
            B.x = fma(b[q].x, OddConst<R>::s[idx], B.x);
            B.y = fma(b[q].y, OddConst<R>::s[idx], B.y);
        }
        double2 rb = rot(B, s);
        out[m] = cadd(A, rb);
        out[R - m] = csub(A, rb);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
}

template <int R> __device__ __forceinline__ void dft(double2 *v, double s);
template <> __device__ __forceinline__ void dft<2>(double2 *v, double) { dft2(v[0], v[1]); }
template <> __device__ __forceinline__ void dft<3>(double2 *v, double s) { dft_odd<3>(v, s); }
template <> __device__ __forceinline__ void dft<4>(double2 *v, double s) { dft4(v[0], v[1], v[2], v[3], s); }
template <> __device__ __forceinline__ void dft<5>(double2 *v, double s) { dft_odd<5>(v, s); }
template <> __device__ __forceinline__ void dft<7>(double2 *v, double s) { dft_odd<7>(v, s); }
template <> __device__ __forceinline__ void dft<8>(double2 *v, double s) { dft8(v, s); }
template <> __device__ __forceinline__ void dft<11>(double2 *v, double s) { dft_odd<11>(v, s); }
template <> __device__ __forceinline__ void dft<13>(double2 *v, double s) { dft_odd<13>(v, s); }
template <> __device__ __forceinline__ void dft<16>(double2 *v, double s) { dft16(v, s); }

// b mod p for 0 <= b < 2^13 and p >= 1, from rp ~ 1 / p in FP32 (the hardware
// reciprocal, within 1 ulp): the quotient estimate b rp is within 2^-9 of b / p,
// so one correction step gives the exact remainder.  About six VALU
// instructions against the seventeen of the integer division.
__device__ __forceinline__ int mod_small(int b, int p, float rp)
{
    const int q = (int)((float)b * rp);
    int r = b - q * p;
    r = r < 0 ? r + p : r;
    return r >= p ? r - p : r;
}

// Butterflies a thread may hold per stage of radix R (register budget: 16 points).
__host__ __device__ constexpr int nb_max(int R) { return R >= 16 ? 1 : 16 / R; }

}  // namespace ifft

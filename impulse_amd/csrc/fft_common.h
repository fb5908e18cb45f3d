// Device helpers shared by the FFT kernels: batch geometry and the per-point
// load modes (see fft_params.h for the geometry model).
#pragma once

#include "butterflies.h"
#include "fft_params.h"

namespace ifft {

// ---------------------------------------------------------------------------
// batch geometry helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ long long dims_off(const Side &s, unsigned p)
{
    long long off = 0;
    for (int d = 0; d < s.nbd; ++d) {
        const unsigned c = (unsigned)s.cnt[d];
        const unsigned q = p / c;
        off += (long long)(p - q * c) * s.str[d];
        p = q;
    }
    return off;
}

// base offsets of logical outer index u: row a (and row b for pair sides)
__device__ __forceinline__ void side_base(const Side &s, unsigned u, long long &a, long long &b, int &hasb)
{
    if (s.kind == SIDE_LINEAR) {
        a = (long long)u * s.lin;
        b = 0;
        hasb = 0;
    } else if (s.kind == SIDE_DIMS_PAIR) {
        const unsigned p = 2 * (u + (unsigned)s.u0);
        a = dims_off(s, p);
        hasb = (long long)(p + 1) < s.nphys;
        b = hasb ? dims_off(s, p + 1) : 0;
    } else {
        a = dims_off(s, u + (unsigned)s.u0);
        b = 0;
        hasb = 0;
    }
}

__device__ __forceinline__ double2 twolevel(const double2 *lo, const double2 *hi, int logS, long long m)
{
    return cmul(lo[m & ((1LL << logS) - 1)], hi[m >> logS]);
}

__device__ __forceinline__ double2 real_tw(const FftParams &P, long long k)
{
    // W_{2M}^k for the real-half modes
    return P.rtw ? P.rtw[k] : twolevel(P.rtw_lo, P.rtw_hi, P.rtw_logS, k);
}

// ---------------------------------------------------------------------------
// element access of user arrays: double / double2, or float / float2 widened
// to FP64 on load and rounded to nearest on store (the flag is uniform per
// launch)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double2 gld_c(const void *p, int f32, long long i)
{
    if (f32) {
        const float2 v = ((const float2 *)p)[i];
        return make_double2((double)v.x, (double)v.y);
    }
    return ((const double2 *)p)[i];
}

__device__ __forceinline__ double gld_r(const void *p, int f32, long long i)
{
    return f32 ? (double)((const float *)p)[i] : ((const double *)p)[i];
}

__device__ __forceinline__ void gst_c(void *p, int f32, long long i, double2 v)
{
    if (f32) ((float2 *)p)[i] = make_float2((float)v.x, (float)v.y);
    else ((double2 *)p)[i] = v;
}

__device__ __forceinline__ void gst_r(void *p, int f32, long long i, double v)
{
    if (f32) ((float *)p)[i] = (float)v;
    else ((double *)p)[i] = v;
}

// ---------------------------------------------------------------------------
// per-transform info and the load of one point
// ---------------------------------------------------------------------------

struct TfInfo {
    long long ia, ib, oa, ob;   // signal bases (input rows a/b, output rows a/b)
    long long ip0, op0;         // position of element 0 (input / output side)
    long long j;                // inner index
    int flags;                  // 1 valid, 2 input row b exists, 4 output row b exists
};

__device__ __forceinline__ double2 load_point(const FftParams &P, const TfInfo &T, long long pos)
{
    const Side &S = P.in;
    const int f32 = P.src32;
    switch (P.load_mode) {
        case LD_C2C:
            return gld_c(P.src, f32, T.ia + pos * S.sax);
        case LD_REAL_HALF:
            return make_double2(gld_r(P.src, f32, T.ia + 2 * pos * S.sax), gld_r(P.src, f32, T.ia + (2 * pos + 1) * S.sax));
        case LD_R2C_PAIR: {
            double a = gld_r(P.src, f32, T.ia + pos * S.sax);
            double b = (T.flags & 2) ? gld_r(P.src, f32, T.ib + pos * S.sax) : 0.0;
            return make_double2(a, b);
        }
        case LD_C2R_HALF: {
            const long long M = P.M;
            double2 xp = gld_c(P.src, f32, T.ia + pos * S.sax);
            double2 xq = gld_c(P.src, f32, T.ia + (M - pos) * S.sax);
            if (P.fwd_flag) { xp = cconj(xp); xq = cconj(xq); }
            if (pos == 0) { xp.y = 0.0; xq.y = 0.0; }
            else if (P.adj_w) { xp = cscale(xp, 0.5); xq = cscale(xq, 0.5); }
            double2 e = cadd(xp, cconj(xq));
            double2 o = cmul(csub(xp, cconj(xq)), cconj(real_tw(P, pos)));
            return make_double2(e.x - o.y, e.y + o.x);
        }
        default: {  // LD_C2R_PAIR
            const long long L = P.M;
            const bool mirror = 2 * pos > L;
            const long long k = mirror ? L - pos : pos;
            double2 a = gld_c(P.src, f32, T.ia + k * S.sax);
            double2 b = (T.flags & 2) ? gld_c(P.src, f32, T.ib + k * S.sax) : make_double2(0.0, 0.0);
            if (P.fwd_flag != (int)mirror) { a = cconj(a); b = cconj(b); }
            if (pos == 0) { a.y = 0.0; b.y = 0.0; }
            else if (P.adj_w) { a = cscale(a, 0.5); b = cscale(b, 0.5); }
            return make_double2(a.x - b.y, a.y + b.x);
        }
    }
}

// raw buffer access: a scalar 48-bit base and 32-bit byte offsets (loads past
// the window return zero, stores past it are dropped)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, long long span)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                             (int)(span < 0x7fffffffLL ? span : 0x7fffffffLL), 0x00020000);
}

__device__ __forceinline__ double2 bload2(__amdgpu_buffer_rsrc_t r, int off)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

__device__ __forceinline__ double bload1(__amdgpu_buffer_rsrc_t r, int off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

__device__ __forceinline__ float bloadf1(__amdgpu_buffer_rsrc_t r, int off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

__device__ __forceinline__ float2 bloadf2(__amdgpu_buffer_rsrc_t r, int off)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// a transform's input; NT = true marks the load non-temporal (streamed past the L2), for
// passes whose lines no other workgroup reads
template <bool NT>
__device__ __forceinline__ double2 bload2_data(__amdgpu_buffer_rsrc_t r, int off)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, NT ? 2 : 0));
}

__device__ __forceinline__ void bstore2(__amdgpu_buffer_rsrc_t r, int off, double2 v)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned int, v),
                                           r, off, 0, 0);
}

// v rounded to complex float
__device__ __forceinline__ void bstoref2(__amdgpu_buffer_rsrc_t r, int off, double2 v)
{
    const float2 f = make_float2((float)v.x, (float)v.y);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned int, f),
                                          r, off, 0, 0);
}

__device__ __forceinline__ double2 twolevel_b(__amdgpu_buffer_rsrc_t lo, __amdgpu_buffer_rsrc_t hi, int logS,
                                              long long m)
{
    return cmul(bload2(lo, (int)(m & ((1LL << logS) - 1)) * 16), bload2(hi, (int)(m >> logS) * 16));
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Renumber them so that XCD x takes logical blocks
// j*8K + x*K + [0, K) in turn (K = IMPULSE_XCD_CHUNK, one block per CU of an
// XCD): the blocks an XCD runs at once are neighbours, so the short runs
// they read or write in the same lines merge in its L2, while the 8 XCDs
// together stay on one contiguous stretch of 8K blocks and so spread over
// all HBM channels.  (Giving each XCD one contiguous eighth of the grid
// instead keeps the L2 merging but sends the 8 XCDs to addresses 1/8 of the
// array apart at the same time; with power-of-two strides those fall on the
// same channels, and transposing passes ran 10-30% slower.)  K = 0 selects
// that contiguous-eighth mapping.
#ifndef IMPULSE_XCD_CHUNK
#define IMPULSE_XCD_CHUNK 32
#endif
__device__ __forceinline__ unsigned xcd_span_block(unsigned b, unsigned grid)
{
    const unsigned g8 = grid & ~7u;
    return b < g8 ? (b & 7u) * (g8 >> 3) + (b >> 3) : b;
}

__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned grid)
{
    constexpr unsigned K = IMPULSE_XCD_CHUNK;
    if constexpr (K == 0) {
        return xcd_span_block(b, grid);
    } else {
        const unsigned g = grid / (8 * K) * (8 * K);
        if (b >= g) return b;
        const unsigned j = b >> 3;
        return (j / K) * (8 * K) + (b & 7u) * K + (j % K);
    }
}

__device__ __forceinline__ TfInfo make_tfinfo(const FftParams &P, long long g)
{
    TfInfo T;
    T.flags = 0;
    T.ia = T.ib = T.oa = T.ob = T.ip0 = T.op0 = T.j = 0;
    if (g < P.ntf) {
        const unsigned gg = (unsigned)g, ni = (unsigned)P.ni;
        const unsigned u = gg / ni, j = gg - u * ni;
        int hb_in, hb_out;
        side_base(P.in, u, T.ia, T.ib, hb_in);
        side_base(P.out, u, T.oa, T.ob, hb_out);
        T.ia += j * P.in.isig;
        T.ib += j * P.in.isig;
        T.oa += j * P.out.isig;
        T.ob += j * P.out.isig;
        T.ip0 = (long long)j * P.in.ipos;
        T.op0 = (long long)j * P.out.ipos;
        T.j = j;
        T.flags = 1 | (hb_in ? 2 : 0) | (hb_out ? 4 : 0);
    }
    return T;
}

// load of input point `pos` with the Bluestein pre-processing (conjugate,
// chirp, zero padding beyond ilim)
__device__ __forceinline__ double2 load_full(const FftParams &P, const TfInfo &T, long long pos)
{
    double2 z = make_double2(0.0, 0.0);
    if (pos < P.ilim) {
        z = load_point(P, T, pos);
        if (P.conj_in) z = cconj(z);
        if (P.chirp_in) z = cmul(z, P.chirp[pos]);
    }
    return z;
}

// store of local point n (value z) for the per-point store modes; the
// Bluestein post-processing (chirp, conjugate) is applied here
__device__ __forceinline__ void store_value(const FftParams &P, const TfInfo &T, long long n, double2 z)
{
    const Side &S = P.out;
    const long long pos = T.op0 + n * S.pst;
    if (P.chirp_out && pos < P.olim) z = cmul(z, P.chirp[pos]);
    if (P.conj_out) z = cconj(z);
    switch (P.store_mode) {
        case ST_TWIDDLE: {
            double2 w = tw_sign(twolevel(P.tw_lo, P.tw_hi, P.tw_logS, (long long)P.tw_mul * T.j * n), (double)P.sign);
            gst_c(P.dst, P.dst32, T.oa + pos * S.sax, cmul(z, w));
            return;
        }
        case ST_MUL_TABLE:
            gst_c(P.dst, P.dst32, T.oa + pos * S.sax, cmul(z, P.mtab[pos]));
            return;
        case ST_C2R_HALF: {
            if (pos >= P.olim) return;
            gst_r(P.dst, P.dst32, T.oa + 2 * pos * S.sax, z.x * P.fct);
            gst_r(P.dst, P.dst32, T.oa + (2 * pos + 1) * S.sax, z.y * P.fct);
            return;
        }
        case ST_C2R_PAIR: {
            if (pos >= P.olim) return;
            gst_r(P.dst, P.dst32, T.oa + pos * S.sax, z.x * P.fct);
            if (T.flags & 4) gst_r(P.dst, P.dst32, T.ob + pos * S.sax, z.y * P.fct);
            return;
        }
        default:  // ST_C2C
            if (pos >= P.olim) return;
            gst_c(P.dst, P.dst32, T.oa + pos * S.sax, cscale(z, P.fct));
            return;
    }
}

}  // namespace ifft

// Kernel-side parameter blocks of the impulse FFT engine (host and device).
//
// Geometry model.  (Transform counts, batch counts and positions fit 31
// bits; the planner rejects larger launches, so device code indexes batches
// with 32-bit integer arithmetic and only offsets are 64-bit.)
//  A launch runs a batch of "logical transforms" of length N
// (complex points) through LDS.  Logical transform g is split into an inner
// index j = g % ni and an outer index u = g / ni.  Element n of transform g
// sits at signal position   pos = j*ipos + n*pst   of a signal whose element 0
// is at   base(u) + j*isig ;  its address is   base + pos*sax   in units of the
// array's scalar (double or float for real arrays, double2 or float2 for
// complex arrays).
// base(u) comes from one of three "side" layouts:
//   SIDE_DIMS      : u is a physical batch index decomposed over nbd dims
//   SIDE_DIMS_PAIR : u covers physical rows 2u and 2u+1 (row 2u+1 may not exist)
//   SIDE_LINEAR    : base = u * lin (a contiguous workspace)
// One-pass transforms use ni = the batch dim kept for coalescing (isig = its
// stride, ipos = 0); four-step passes use ni = the other factor (isig = 0,
// ipos = 1 or N1) so that loads and stores can address the full signal.
#pragma once

#include <hip/hip_runtime.h>

namespace ifft {

constexpr int kMaxStages = 24;
constexpr int kMaxBatchDims = 8;
constexpr int kMaxTB = 64;          // transforms per workgroup
constexpr int kMaxLdsLen = 8192;    // longest single-pass FP64 transform (128 KiB of LDS)
constexpr int kMaxThreads = 512;       // workgroup size cap (256 VGPRs per lane available)

enum LoadMode : int {
    LD_C2C = 0,       // z = x[pos]                                  (complex input)
    LD_REAL_HALF = 1, // z = x[2 pos] + i x[2 pos + 1]               (even-length real input)
    LD_R2C_PAIR = 2,  // z = xa[pos] + i xb[pos]                     (odd-length real rows 2u, 2u+1)
    LD_C2R_HALF = 3,  // z = (X[p] + X*[M-p]) + i (X[p] - X*[M-p]) W_{2M}^{-p}  (even c2r)
    LD_C2R_PAIR = 4   // z = A[pos] + i B[pos], A, B Hermitian-extended half spectra (odd c2r)
};

enum StoreMode : int {
    ST_C2C = 0,       // y[pos] = fct z[n]
    ST_TWIDDLE = 1,   // y[pos] = z[n] W_L^{j n}                     (four-step pass A)
    ST_R2C_HALF = 2,  // X[k], k = 0..M, from Z[k], Z[M-k]           (even r2c)
    ST_R2C_PAIR = 3,  // Xa[k], Xb[k], k = 0..L/2, from Z[k], Z[L-k] (odd r2c pair)
    ST_C2R_HALF = 4,  // x[2 pos] = Re z, x[2 pos + 1] = Im z
    ST_C2R_PAIR = 5,  // xa[pos] = Re z, xb[pos] = Im z
    ST_MUL_TABLE = 6, // y[pos] = z[n] tab[pos]                      (Bluestein spectrum product)
    ST_R2C_ROWPAIR = 7, // even r2c, last four-step pass: rows j and L2-j of one workgroup hold
                        // Z[k] and Z[M-k], stored as the half spectrum X[k], X[M-k]
    ST_CONV_TWIDDLE = 8 // four-step Bluestein row pass: row j is transformed, multiplied by
                        // row j of the transposed kernel spectrum (mtab), transformed back with
                        // the opposite sign and stored as y[pos] = z[n] conj(W_L^{j n})
};

enum SideKind : int { SIDE_DIMS = 0, SIDE_DIMS_PAIR = 1, SIDE_LINEAR = 2 };

struct Side {
    int kind;
    int nbd;                        // SIDE_DIMS*: number of batch dims (dim 0 fastest)
    long long cnt[kMaxBatchDims];   // physical counts
    long long str[kMaxBatchDims];   // strides (scalar units of this side's array)
    long long nphys;                // physical batch count (pair bound check)
    long long lin;                  // SIDE_LINEAR stride per logical transform
    long long isig;                 // added to the base per inner index j
    long long ipos;                 // added to the position per inner index j
    long long pst;                  // position step per element n
    long long sax;                  // scalar stride per position
    long long u0;                   // SIDE_DIMS*: first logical outer index of this launch (batch chunks)
};

struct FftParams {
    // transform in LDS
    int N;                     // transform length in LDS (points)
    int nst;                   // number of Stockham stages
    int radix[kMaxStages];
    int NT;                    // threads per transform
    int TB;                    // transforms per workgroup
    int LD;                    // LDS words (double2) per transform
    // modes
    int load_mode, store_mode;
    int col_load, col_store;   // lane mapping: 1 = lanes walk transforms first
    int chirp_in, conv, chirp_out;  // Bluestein pieces
    int conj_in, conj_out;     // backward Bluestein runs the forward tables on conjugates
    int sign;                  // -1: forward stages, +1: backward stages
    int fwd_flag;              // r2c: 0 => conjugate output; c2r: 1 => conjugate input
    int adj_w;                 // adjoint weights: c2r loads take interior bins x1/2, r2c stores
                               // give interior bins x2 (bins 0 and n/2 keep weight 1)
    long long ilim;            // input positions >= ilim load as zero (Bluestein zero padding)
    long long olim;            // output positions >= olim are not stored
    int M;                     // half length (real-half modes) or odd length (pair modes)
    int src32, dst32;          // the source / destination array holds float (complex: float2)
                               // rather than double; arithmetic stays FP64 (workspaces are FP64)
    int nload;                 // local points loaded per transform (the rest of the row is zero)
    int nstore;                // points stored per transform
    double fct;
    // batch
    long long ntf;             // logical transforms in this launch
    long long ni;              // inner count
    Side in, out;
    // arrays
    const void *src;
    void *dst;
    long long src_lo, src_span;   // byte window of the source array around src (lo <= 0)
    long long dst_lo, dst_span;   // byte window of the destination array around dst
    const double2 *tw;         // W_N^j (forward sign), j < N
    const double2 *chirp;      // w[n] = exp(-i pi n^2 / nsig), n < nsig
    const double2 *spec;       // Bluestein kernel spectrum / N, N entries
    const double2 *rtw;        // W_{2M}^k, k <= M (real-half modes)
    const double2 *tw_lo;      // two-level W_L^m = lo[m & (S-1)] * hi[m >> log2 S], S a power of two
    const double2 *tw_hi;
    long long tw_S, tw_L;
    int tw_logS;
    int tw_mul;                // ST_TWIDDLE uses W_L^{tw_mul j n}
    const double2 *rtw_lo;     // two-level W_{2M} used by the real-half modes when rtw is null
    const double2 *rtw_hi;
    long long rtw_S;
    int rtw_logS;
    const double2 *mtab;       // ST_MUL_TABLE
};

// Post-process of a four-step real transform: Z (contiguous per logical
// signal) -> half spectra in the output layout.
struct RealPostParams {
    int half;                  // 1: even length, Z has M points; 0: odd pair, Z has L points
    int adj_w;                 // interior bins x2 (adjoint of c2r)
    int M;                     // M (half) or L (pair)
    int fwd_flag;
    int dst32;                 // the destination holds complex float rather than complex double
    double fct;
    long long ntf;             // logical signals
    long long zlen;            // Z stride per logical signal
    Side out;
    const double2 *Z;
    void *dst;
    const double2 *rtw;        // W_{2M}^k, k <= M (half)
    const double2 *tw_lo, *tw_hi;
    long long tw_S;
    int tw_logS;
    int two_level;             // rtw replaced by the two-level table
};

}  // namespace ifft

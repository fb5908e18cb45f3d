// Power-of-two FP64 FFT kernel with compile-time length N (64 .. 8192).
//
// Radix-16 Stockham stages (one radix-2/4/8 stage last when log2 N is not a
// multiple of 4).  Each of the N/16 threads of a transform holds 16 points in
// registers for every stage: the first stage reads them straight from global
// memory (x[i + q N/16], coalesced), the middle stages exchange through a
// padded LDS row (index + index/16, which keeps the radix-16 scatter free of
// bank conflicts), and the last stage writes straight to global memory
// (y[b + q N/R], coalesced).  Per-point load modes (complex, real-half,
// real-pair, c2r pre-processing, Bluestein chirp) and store modes (plain,
// four-step twiddle, Bluestein spectrum product or chirp, c2r outputs) are
// fused into those global accesses.
//
// MAP selects the lane mapping of the global accesses:
//   MAP_ROW    : lanes walk along a transform (contiguous axis), loads and stores
//   MAP_COL    : lanes walk across the TB transforms of a workgroup first
//                (transforms adjacent in memory, strided axis); the LDS rows
//                are interleaved so the exchange stays conflict-free
//   MAP_ROWCOL : loads along the transform, stores across transforms (the
//                second pass of a four-step transform, which writes its output
//                with the stride of the first factor): the last stage goes
//                through LDS once more to transpose the store.
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

// LDS words per transform: padded row plus one word so that rows of
// neighbouring transforms start on different banks
template <int N> constexpr int pow2_ld() { return N + N / 16 + 1; }

enum Pow2Map { MAP_ROW = 0, MAP_COL = 1, MAP_ROWCOL = 2 };

// radix-R butterflies held by one thread: R-point groups v[c*R .. c*R+R)
template <int R>
__device__ __forceinline__ void dft_groups(double2 *v, double s)
{
#pragma unroll
    for (int c = 0; c < 16 / R; ++c) dft<R>(v + c * R, s);
}

template <int N, int MAP>
__global__ void __launch_bounds__(512) k_fft_pow2(const FftParams P)
{
    extern __shared__ double2 lds[];
    __shared__ TfInfo info[MAP == MAP_ROWCOL ? kMaxTB : 1];
    constexpr bool COL = MAP == MAP_COL;
    constexpr int NT = N / 16;
    constexpr int LDP = pow2_ld<N>();
    using RS = Pow2Radices<N>;
    const int TB = P.TB;
    const int tid = threadIdx.x;
    const int t = COL ? tid % TB : tid / NT;
    const int i = COL ? tid / TB : tid % NT;
    const double s = (double)P.sign;
    const TfInfo T = make_tfinfo(P, (long long)blockIdx.x * TB + t);
    const bool valid = T.flags & 1;
    const double2 *__restrict__ tw = P.tw;

    auto at = [&](int idx) -> double2 & {
        return COL ? lds[padx(idx) * TB + t] : lds[t * LDP + padx(idx)];
    };

    double2 v[16];
    // ---- stage 0 (radix 16, p = 1): inputs x[i + q N/16] from global memory
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int n = i + q * (N / 16);
        v[q] = valid ? load_full(P, T, T.ip0 + (long long)n * P.in.pst) : make_double2(0.0, 0.0);
    }
    dft<16>(v, s);
    // outputs y[16 i + q]
#pragma unroll
    for (int q = 0; q < 16; ++q) at(16 * i + q) = v[q];
    __syncthreads();

    int p = 16;
    // ---- middle stages (all radix 16)
#pragma unroll
    for (int st = 1; st + 1 < RS::n; ++st) {
        constexpr int R = 16;
        const int nb = N / R;
        const int k = i % p;
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = at(i + q * nb);
        __syncthreads();
        const int twstep = N / (p * R);
#pragma unroll
        for (int q = 1; q < 16; ++q) v[q] = cmul(v[q], tw_sign(tw[q * k * twstep], s));
        dft<16>(v, s);
        const int j = (i - k) * R + k;
#pragma unroll
        for (int q = 0; q < 16; ++q) at(j + q * p) = v[q];
        __syncthreads();
        p *= R;
    }

    // ---- last stage (radix RL, p = N / RL): 16 / RL butterflies b = i + c NT
    constexpr int RL = RS::r[RS::n - 1];
    constexpr int NB = 16 / RL;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int b = i + c * NT;
#pragma unroll
        for (int q = 0; q < RL; ++q) v[c * RL + q] = at(b + q * (N / RL));
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int b = i + c * NT;   // k = b since p = N / RL > b
#pragma unroll
        for (int q = 1; q < RL; ++q) v[c * RL + q] = cmul(v[c * RL + q], tw_sign(tw[q * b], s));
    }
    dft_groups<RL>(v, s);
    if constexpr (MAP != MAP_ROWCOL) {
        if (valid) {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const int b = i + c * NT;
#pragma unroll
                for (int q = 0; q < RL; ++q) store_value(P, T, b + q * (N / RL), v[c * RL + q]);
            }
        }
    } else {
        // transpose through LDS: natural order rows, then lanes across transforms
        if (i == 0) info[t] = T;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int b = i + c * NT;
#pragma unroll
            for (int q = 0; q < RL; ++q) at(b + q * (N / RL)) = v[c * RL + q];
        }
        __syncthreads();
        const int nthr = blockDim.x;
        for (int e = tid; e < TB * N; e += nthr) {
            const int tt = e % TB, n = e / TB;
            const TfInfo &U = info[tt];
            if (U.flags & 1) store_value(P, U, n, lds[tt * LDP + padx(n)]);
        }
    }
}

template <int N>
static hipError_t launch_n(const FftParams &P, int map, unsigned grid, unsigned block, size_t lds, hipStream_t s)
{
    static bool attr_done = false;
    if (!attr_done) {
        for (const void *f : {(const void *)k_fft_pow2<N, MAP_ROW>, (const void *)k_fft_pow2<N, MAP_COL>,
                              (const void *)k_fft_pow2<N, MAP_ROWCOL>})
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192) != hipSuccess)
                (void)hipGetLastError();
        attr_done = true;
    }
    switch (map) {
        case MAP_COL: hipLaunchKernelGGL((k_fft_pow2<N, MAP_COL>), dim3(grid), dim3(block), lds, s, P); break;
        case MAP_ROWCOL: hipLaunchKernelGGL((k_fft_pow2<N, MAP_ROWCOL>), dim3(grid), dim3(block), lds, s, P); break;
        default: hipLaunchKernelGGL((k_fft_pow2<N, MAP_ROW>), dim3(grid), dim3(block), lds, s, P); break;
    }
    return hipGetLastError();
}

bool pow2_supported(int N) { return N >= 64 && N <= 8192 && (N & (N - 1)) == 0; }
int pow2_lds_words(int N) { return N + N / 16 + 1; }

hipError_t launch_fft_pow2(const FftParams &P, int map, unsigned grid, unsigned block, size_t lds, hipStream_t s)
{
    switch (P.N) {
        case 64: return launch_n<64>(P, map, grid, block, lds, s);
        case 128: return launch_n<128>(P, map, grid, block, lds, s);
        case 256: return launch_n<256>(P, map, grid, block, lds, s);
        case 512: return launch_n<512>(P, map, grid, block, lds, s);
        case 1024: return launch_n<1024>(P, map, grid, block, lds, s);
        case 2048: return launch_n<2048>(P, map, grid, block, lds, s);
        case 4096: return launch_n<4096>(P, map, grid, block, lds, s);
        case 8192: return launch_n<8192>(P, map, grid, block, lds, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ifft

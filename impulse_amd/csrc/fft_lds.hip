// Generic single-pass FP64 FFT kernel: a batch of length-N transforms staged
// through LDS, mixed-radix Stockham autosort stages with runtime radices
// (2, 3, 4, 5, 7, 8, 11, 13, 16), and fused load/store modes for real
// transforms, four-step passes and Bluestein.  See fft_params.h for the
// geometry model.
//
// Stockham stage (radix R, p = product of the radices already applied):
//   butterfly b in [0, N/R): k = b mod p,
//   inputs  x[b + q N/R] * W_{pR}^{q k},  q < R
//   outputs y[(b - k) R + k + q p] = DFT_R(inputs)[q]
// which leaves the result in natural order after the last stage.  Each stage
// reads its inputs into registers, synchronises, and writes back in place.
#include "fft_common.h"

namespace ifft {

// ---------------------------------------------------------------------------
// Stockham stages in LDS
// ---------------------------------------------------------------------------

// Each radix's stage is a call of its own, so the kernel's register budget is
// that of its largest stage rather than of all nine inlined into one switch
// (210 registers, two waves per SIMD).  The kernel is built for three waves
// per SIMD (168 registers) and four (128 registers); fft_lds_wpe picks one per
// launch.
template <int R, int NB, int WPE>
__device__ __noinline__ void stage(double2 *__restrict__ buf, int N, int p, int i, int NT,
                                      const double2 *__restrict__ tw, double s)
{
    const int nb = N / R;
    const int twstep = N / (p * R);
    const float rp = __builtin_amdgcn_rcpf((float)p);
    double2 v[NB][R];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int b = i + c * NT;
        if (b < nb) {
#pragma unroll
            for (int q = 0; q < R; ++q) v[c][q] = buf[b + q * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int b = i + c * NT;
        if (b < nb) {
            const int k = mod_small(b, p, rp);   // b < N / R <= 4096
            if (p > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) v[c][q] = cmul(v[c][q], tw_sign(tw[q * k * twstep], s));
            }
            dft<R>(v[c], s);
            const int j = (b - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[j + q * p] = v[c][q];
        }
    }
    __syncthreads();
}

// WPE only separates the stage instantiations of the two kernels, so that each
// is compiled for its own kernel's register budget
template <int WPE>
__device__ void run_stages(double2 *buf, const FftParams &P, int i, double s)
{
    int p = 1;
    for (int st = 0; st < P.nst; ++st) {
        const int r = P.radix[st];
        switch (r) {
            case 2: stage<2, nb_max(2), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
            case 3: stage<3, nb_max(3), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
            case 4: stage<4, nb_max(4), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
            case 5: stage<5, nb_max(5), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
            case 7: stage<7, nb_max(7), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
            case 8: stage<8, nb_max(8), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
            case 11: stage<11, nb_max(11), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
            case 13: stage<13, nb_max(13), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
            default: stage<16, nb_max(16), WPE>(buf, P.N, p, i, P.NT, P.tw, s); break;
        }
        p *= r;
    }
}

__device__ __forceinline__ void put_c(const FftParams &P, long long off, double2 y)
{
    gst_c(P.dst, P.dst32, off, y);
}

// store local point n of transform t; `buf` is the transform's LDS row
__device__ __forceinline__ void store_point(const FftParams &P, const TfInfo &T, const double2 *buf, int n)
{
    const Side &S = P.out;
    const long long pos = T.op0 + (long long)n * S.pst;
    switch (P.store_mode) {
        case ST_C2C: {
            if (pos >= P.olim) return;
            put_c(P, T.oa + pos * S.sax, cscale(buf[n], P.fct));
            return;
        }
        case ST_TWIDDLE: {
            double2 w = tw_sign(twolevel(P.tw_lo, P.tw_hi, P.tw_logS, (long long)P.tw_mul * T.j * (long long)n), (double)P.sign);
            put_c(P, T.oa + pos * S.sax, cmul(buf[n], w));
            return;
        }
        case ST_MUL_TABLE: {
            put_c(P, T.oa + pos * S.sax, cmul(buf[n], P.mtab[pos]));
            return;
        }
        case ST_R2C_HALF: {
            const int M = P.M;
            double2 zk = buf[n % M], zm = cconj(buf[(M - n) % M]);
            double2 e = cscale(cadd(zk, zm), 0.5);
            double2 d = csub(zk, zm);
            double2 o = make_double2(0.5 * d.y, -0.5 * d.x);  // -i/2 (zk - zm*)
            double2 X = cadd(e, cmul(o, real_tw(P, n)));
            X = cscale(X, (P.adj_w && n != 0 && n != M) ? 2.0 * P.fct : P.fct);
            if (!P.fwd_flag) X = cconj(X);
            put_c(P, T.oa + pos * S.sax, X);
            return;
        }
        case ST_R2C_PAIR: {
            const int L = P.M;
            double2 zk = buf[n], zm = cconj(buf[(L - n) % L]);
            const double f = (P.adj_w && n != 0) ? 2.0 * P.fct : P.fct;
            double2 xa = cscale(cadd(zk, zm), 0.5 * f);
            double2 d = csub(zk, zm);
            double2 xb = make_double2(0.5 * f * d.y, -0.5 * f * d.x);
            if (!P.fwd_flag) { xa = cconj(xa); xb = cconj(xb); }
            put_c(P, T.oa + pos * S.sax, xa);
            if (T.flags & 4) put_c(P, T.ob + pos * S.sax, xb);
            return;
        }
        case ST_C2R_HALF: {
            if (pos >= P.olim) return;
            double2 z = buf[n];
            gst_r(P.dst, P.dst32, T.oa + 2 * pos * S.sax, z.x * P.fct);
            gst_r(P.dst, P.dst32, T.oa + (2 * pos + 1) * S.sax, z.y * P.fct);
            return;
        }
        default: {  // ST_C2R_PAIR
            if (pos >= P.olim) return;
            double2 z = buf[n];
            gst_r(P.dst, P.dst32, T.oa + pos * S.sax, z.x * P.fct);
            if (T.flags & 4) gst_r(P.dst, P.dst32, T.ob + pos * S.sax, z.y * P.fct);
            return;
        }
    }
}

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------

template <int WPE>
__global__ void __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(WPE))) k_fft_lds(const FftParams P)
{
    extern __shared__ double2 lds[];
    __shared__ TfInfo info[kMaxTB];

    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int TB = P.TB, N = P.N, LD = P.LD;
    const long long g0 = (long long)xcd_block(blockIdx.x, gridDim.x) * TB;

    if (tid < TB) {
        const TfInfo T = make_tfinfo(P, g0 + tid);
        info[tid] = T;
    }
    __syncthreads();

    // global -> LDS, with zero padding beyond the signal and the Bluestein chirp
    for (int e = tid; e < TB * N; e += nthr) {
        int t, n;
        if (P.col_load) { t = e % TB; n = e / TB; }
        else { n = e % N; t = e / N; }
        const TfInfo &T = info[t];
        double2 z = make_double2(0.0, 0.0);
        const long long pos = T.ip0 + (long long)n * P.in.pst;
        if ((T.flags & 1) && n < P.nload && pos < P.ilim) {
            z = load_point(P, T, pos);
            if (P.conj_in) z = cconj(z);
            if (P.chirp_in) z = cmul(z, P.chirp[pos]);
        }
        lds[t * LD + n] = z;
    }
    __syncthreads();

    const int t = tid / P.NT, i = tid % P.NT;
    double2 *buf = lds + t * LD;
    run_stages<WPE>(buf, P, i, (double)P.sign);

    if (P.conv) {
        // Bluestein: multiply by the chirp kernel's spectrum, transform back
        for (int e = tid; e < TB * N; e += nthr) {
            const int tt = e / N, n = e % N;
            lds[tt * LD + n] = cmul(lds[tt * LD + n], P.spec[n]);
        }
        __syncthreads();
        run_stages<WPE>(buf, P, i, 1.0);
    }
    if (P.chirp_out || P.conj_out) {
        for (int e = tid; e < TB * N; e += nthr) {
            const int tt = e / N, n = e % N;
            const long long pos = info[tt].op0 + (long long)n * P.out.pst;
            double2 z = lds[tt * LD + n];
            if (P.chirp_out && pos < P.olim) z = cmul(z, P.chirp[pos]);
            if (P.conj_out) z = cconj(z);
            lds[tt * LD + n] = z;
        }
        __syncthreads();
    }

    // LDS -> global
    for (int e = tid; e < TB * P.nstore; e += nthr) {
        int tt, n;
        if (P.col_store) { tt = e % TB; n = e / TB; }
        else { n = e % P.nstore; tt = e / P.nstore; }
        const TfInfo &T = info[tt];
        if (T.flags & 1) store_point(P, T, lds + tt * LD, n);
    }
}

// ---------------------------------------------------------------------------
// four-step real post-process: Z -> half spectra
// ---------------------------------------------------------------------------

// Each thread takes the bin pair (k, M - k) (half) or (k, L - k) (pair), so
// every Z point is read once and both outputs of the pair are written.
__global__ void __launch_bounds__(256) k_real_post(const RealPostParams P)
{
    const long long L = P.M;
    const long long per = L / 2 + 1;   // pairs k = 0 .. L/2 cover all outputs
    const long long total = per * P.ntf;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long u = e / per, k = e % per;
        const long long km = (L - k) % L;   // mirror index of Z
        const double2 *Z = P.Z + u * P.zlen;
        long long oa, ob;
        int hb;
        side_base(P.out, (unsigned)u, oa, ob, hb);
        const double2 zk = Z[k], zm = Z[km];
        if (P.half) {
            // X[k] = (zk + zm*)/2 - i/2 W^k (zk - zm*),  W = W_{2M}; X[M-k] from the
            // same pair with the roles swapped and W^{M-k} = -conj(W^k)
            const double2 w = P.two_level ? twolevel(P.tw_lo, P.tw_hi, P.tw_logS, k) : P.rtw[k];
            const double2 e1 = cscale(cadd(zk, cconj(zm)), 0.5), d1 = csub(zk, cconj(zm));
            const double2 o1 = make_double2(0.5 * d1.y, -0.5 * d1.x);
            const double fk = (P.adj_w && k != 0) ? 2.0 * P.fct : P.fct;
            const double fm = (P.adj_w && L - k != L && L - k != 0) ? 2.0 * P.fct : P.fct;
            double2 Xk = cscale(cadd(e1, cmul(o1, w)), fk);
            const double2 e2 = cscale(cadd(zm, cconj(zk)), 0.5), d2 = csub(zm, cconj(zk));
            const double2 o2 = make_double2(0.5 * d2.y, -0.5 * d2.x);
            double2 Xm = cscale(cadd(e2, cmul(o2, make_double2(-w.x, w.y))), fm);
            if (!P.fwd_flag) { Xk = cconj(Xk); Xm = cconj(Xm); }
            gst_c(P.dst, P.dst32, oa + k * P.out.sax, Xk);
            if (L - k != k) gst_c(P.dst, P.dst32, oa + (L - k) * P.out.sax, Xm);
        } else {
            // odd pair: Xa = (zk + zm*)/2, Xb = -i/2 (zk - zm*) for k = 0 .. L/2
            const double2 za = cconj(zm);
            const double f = (P.adj_w && k != 0) ? 2.0 * P.fct : P.fct;
            double2 xa = cscale(cadd(zk, za), 0.5 * f);
            const double2 d = csub(zk, za);
            double2 xb = make_double2(0.5 * f * d.y, -0.5 * f * d.x);
            if (!P.fwd_flag) { xa = cconj(xa); xb = cconj(xb); }
            gst_c(P.dst, P.dst32, oa + k * P.out.sax, xa);
            if (hb) gst_c(P.dst, P.dst32, ob + k * P.out.sax, xb);
        }
    }
}

// ---------------------------------------------------------------------------
// C-path layout kernels: complex bins <-> FFTPACK halfcomplex
// ---------------------------------------------------------------------------

// bins[0..n/2] -> hc[0..n): [r0, r1, i1, r2, i2, ... (, r_{n/2})]
__global__ void k_pack_halfcomplex(const double2 *__restrict__ bins, double *__restrict__ hc, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        if (i == 0) hc[0] = bins[0].x;
        else {
            const long long k = (i + 1) / 2;
            hc[i] = (i & 1) ? bins[k].x : bins[k].y;
        }
    }
}

// hc[0..n) -> bins[0..n/2] (imaginary parts of bin 0 and of bin n/2 for even n are zero)
__global__ void k_unpack_halfcomplex(const double *__restrict__ hc, double2 *__restrict__ bins, long long n)
{
    const long long nb = n / 2 + 1;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nb;
         k += (long long)gridDim.x * blockDim.x) {
        double2 z;
        if (k == 0) z = make_double2(hc[0], 0.0);
        else {
            z.x = hc[2 * k - 1];
            z.y = (2 * k < n) ? hc[2 * k] : 0.0;
        }
        bins[k] = z;
    }
}

// ---------------------------------------------------------------------------
// host-side launchers (called from the planner / C ABI)
// ---------------------------------------------------------------------------

// Measured choice (MI355X): passes along rows with one or two transforms per
// workgroup run faster with the larger register budget (2112-point c2c rows
// 1.29 -> 1.07 ms over 16 x 2112^2; r2c of 4096 x 4096 rows 0.185 -> 0.162 ms;
// 256 x 4095 0.030 -> 0.028 ms), wider row tiles and passes across transforms
// with the higher occupancy (the rows of the 2112^2 r2c, three per workgroup:
// 1.96 -> 1.71 ms with both passes; the 2112^2 columns 1.79 -> 1.51 ms).
int fft_lds_wpe(const FftParams &P)
{
    return (!P.col_load && !P.col_store && P.TB <= 2) ? 3 : 4;
}

hipError_t launch_fft_lds(const FftParams &P, unsigned grid, unsigned block, size_t lds_bytes, hipStream_t s)
{
    static bool attr_done = false;
    if (!attr_done) {
        for (const void *f : {(const void *)k_fft_lds<3>, (const void *)k_fft_lds<4>}) {
            hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               160 * 1024 - (int)(sizeof(TfInfo) * kMaxTB) - 1024);
            if (e != hipSuccess) (void)hipGetLastError();
        }
        attr_done = true;
    }
    if (fft_lds_wpe(P) == 4) hipLaunchKernelGGL(k_fft_lds<4>, dim3(grid), dim3(block), lds_bytes, s, P);
    else hipLaunchKernelGGL(k_fft_lds<3>, dim3(grid), dim3(block), lds_bytes, s, P);
    return hipGetLastError();
}

hipError_t launch_real_post(const RealPostParams &P, hipStream_t s)
{
    const long long per = (long long)P.M / 2 + 1;
    long long total = per * P.ntf;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_real_post, dim3((unsigned)blocks), dim3(256), 0, s, P);
    return hipGetLastError();
}

hipError_t launch_pack_halfcomplex(const double2 *bins, double *hc, long long n, hipStream_t s)
{
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pack_halfcomplex, dim3((unsigned)blocks), dim3(256), 0, s, bins, hc, n);
    return hipGetLastError();
}

hipError_t launch_unpack_halfcomplex(const double *hc, double2 *bins, long long n, hipStream_t s)
{
    long long blocks = (n / 2 + 1 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_unpack_halfcomplex, dim3((unsigned)blocks), dim3(256), 0, s, hc, bins, n);
    return hipGetLastError();
}

}  // namespace ifft

/* CPU oracle for the impulse FFT path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C with long-double arithmetic, what the
 * reference's FFT path computes.  It is the checker for the HIP engine: only
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it.
 * The product library never links, loads or calls it.
 *
 * Semantics followed (reference files, relative to the impulse repository):
 *   - n-D c2c / r2c / c2r with byte strides, axes order, `fct` applied once on
 *     the first transformed axis:      impulse/fft/cpp_pocketfft/pocketfft_hdronly.h:3011-3050,3272-3390
 *   - r2c keeps n/2+1 bins; forward=false conjugates the r2c output:  pocketfft_hdronly.h:3125-3179
 *   - c2r ignores Im of DC (and of Nyquist for even n); forward=true
 *     conjugates its input:                                    pocketfft_hdronly.h:3180-3250
 *   - argument checks (ndim, stride ranks, axes):             pocketfft_hdronly.h:446-477
 *   - C path (`rfft_forward/backward`, FFTPACK halfcomplex order
 *     [r0, r1, i1, r2, i2, ...]), `cfft_forward/backward`:     impulse/fft/c_pocketfft/pocketfft.h:18-32,
 *                                                              impulse/fft/c_pocketfft/pocketfft.nim:224-243
 *   - forward transform = exp(-2 pi i jk/n), unnormalised:     impulse/fft/cpp_pocketfft/README_pocketfft.md:102-116
 *
 * The arithmetic is a textbook recursive mixed-radix DFT (smallest prime factor
 * first, naive prime-size butterflies) with Bluestein's chirp-z algorithm for
 * prime lengths above 64, all in long double with octant-reduced twiddles, so
 * it is independent of the GPU engine's radix order and of PocketFFT's own
 * FFTPACK/Bluestein choice (impulse/fft/cpp_pocketfft/pocketfft_hdronly.h:2349-2417).
 *
 * Return codes: 0 ok, 1 invalid argument, 2 out of memory.
 */
#ifndef IMPULSE_FFT_ORACLE_H
#define IMPULSE_FFT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* long-double naive O(n^2) DFT of interleaved complex doubles (brute force pin). */
int oracle_dft_naive(size_t n, const double *in, double *out, int forward);

/* 1-D complex FFT, in place on interleaved (re,im) doubles, result scaled by fct
 * (the C path leaves length-1 data untouched, fct included: pocketfft.c:874). */
int oracle_cfft_forward(size_t n, double *c, double fct);
int oracle_cfft_backward(size_t n, double *c, double fct);

/* 1-D real FFT in place in FFTPACK halfcomplex order, scaled by fct (not for n == 1,
 * pocketfft.c:1703,1739). */
int oracle_rfft_forward(size_t n, double *c, double fct);
int oracle_rfft_backward(size_t n, double *c, double fct);

/* n-D transforms; strides in bytes, `shape` is the real array's shape for r2c/c2r. */
int oracle_c2c(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
               const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
               int forward, const double *data_in, double *data_out, double fct);
int oracle_r2c(size_t ndim, const size_t *shape_in, const ptrdiff_t *stride_in,
               const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
               int forward, const double *data_in, double *data_out, double fct);
int oracle_c2r(size_t ndim, const size_t *shape_out, const ptrdiff_t *stride_in,
               const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
               int forward, const double *data_in, double *data_out, double fct);

/* DCT / DST of types 1-4 along `axes` (pocketfft::dct / pocketfft::dst,
 * impulse/fft/cpp_pocketfft/pocketfft_hdronly.h:3284-3318 over T_dct1, T_dst1,
 * T_dcst23, T_dcst4 at :2424-2648), as the naive sums they compute, in long
 * double, unnormalised (N = length along the axis):
 *   DCT-I   y[k] = x[0] + (-1)^k x[N-1] + 2 sum_{n=1}^{N-2} x[n] cos(pi k n / (N-1))      (N >= 2)
 *   DCT-II  y[k] = 2 sum_n x[n] cos(pi k (2n+1) / 2N)
 *   DCT-III y[k] = x[0] + 2 sum_{n>=1} x[n] cos(pi n (2k+1) / 2N)
 *   DCT-IV  y[k] = 2 sum_n x[n] cos(pi (2k+1)(2n+1) / 4N)
 *   DST-I   y[k] = 2 sum_n x[n] sin(pi (k+1)(n+1) / (N+1))
 *   DST-II  y[k] = 2 sum_n x[n] sin(pi (k+1)(2n+1) / 2N)
 *   DST-III y[k] = (-1)^k x[N-1] + 2 sum_{n<N-1} x[n] sin(pi (2k+1)(n+1) / 2N)
 *   DST-IV  y[k] = 2 sum_n x[n] sin(pi (2k+1)(2n+1) / 4N)
 * `ortho` applies exactly the end-point scalings that code applies (the
 * overall 1/sqrt factor is the caller's fct): DCT-I x[0], x[N-1] times sqrt2
 * before and y[0], y[N-1] times sqrt2/2 after (:2438,2448); DCT-II and DST-II
 * y[0] times sqrt2/2 (:2522, after DST-II's reversal); DCT-III and DST-III
 * x[0] times sqrt2 (:2526, before DST-III's reversal); DST-I, DCT-IV and
 * DST-IV ignore it.  fct scales the first axis only.  Real doubles with byte
 * strides; in == out allowed with equal strides. */
int oracle_dct(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
               size_t naxes, const size_t *axes, int type, const double *data_in, double *data_out, double fct,
               int ortho);
int oracle_dst(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
               size_t naxes, const size_t *axes, int type, const double *data_in, double *data_out, double fct,
               int ortho);

/* upfirdn (impulse/signal.nim:545-594): t upsampled by `up` (up-1 zeros after
 * every sample, length (nt-1) up + 1, the zero insertion the reference's KATs
 * at tests/test_signal.nim:67-85 fix), convolved in full with the nh taps h,
 * every down-th sample from the first kept; y holds
 * ceil(((nt-1) up + nh) / down) values, each a long-double direct sum. */
/* pocketfft::r2r_fftpack (pocketfft_hdronly.h:3392-3403 with ExecR2R, :3252-3270):
 * along each axis the direction picks the operation -- forward: the FFTPACK
 * halfcomplex [r0, r1, i1, r2, i2, ... (, r_{n/2})] of X[k] = sum x[n]
 * e^{-2 pi i k n/N}; backward: halfcomplex input summed back as x[n] = r0 +
 * 2 sum_k (r_k cos - i_k sin)(2 pi k n/N) (+ r_{N/2} (-1)^n) -- and
 * real2hermitian only decides the sign flips of positions 2, 4, ... (the
 * imaginary slots): of the input before a forward call when it is 0, of the
 * output after a backward call when it is 1 (the reference passes `forward`,
 * not `real2hermitian`, as the plan's direction, :3264).
 * Separable Hartley (:3405-3415, copy_hartley :3066-3095): H[k] = Re X[k] +
 * Im X[k] along each axis.  Genuine Hartley (:3417-3445): the same with X the
 * DFT over all axes at once.  fct scales once.  Naive long-double sums. */
int oracle_r2r_fftpack(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                       size_t naxes, const size_t *axes, int real2hermitian, int forward, const double *data_in,
                       double *data_out, double fct);
int oracle_r2r_separable_hartley(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                                 const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                                 const double *data_in, double *data_out, double fct);
int oracle_r2r_genuine_hartley(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                               const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                               const double *data_in, double *data_out, double fct);

size_t oracle_upfirdn_length(size_t nt, size_t nh, int up, int down);
int oracle_upfirdn(size_t nt, const double *t, size_t nh, const double *h, int up, int down, double *y);

/* Twiddle exp(sign * 2 pi i k / n) in long double, exact at multiples of pi/4
 * (returned through re/im as doubles of the long-double value). */
void oracle_twiddle(size_t k, size_t n, int sign, double *re, double *im);

#ifdef __cplusplus
}
#endif
#endif

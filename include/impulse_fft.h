/* impulse_fft.h -- C ABI of the MI355X (gfx950) FFT engine for impulse.
 *
 * Plain C types only (pointers, sizes, ints); no exceptions cross this ABI.
 * Every entry point returns an impulse_fft_status (0 = success) unless noted.
 *
 * Two groups of entry points replace the two FFI surfaces impulse binds today:
 *
 *  (1) the C++ path, `pocketfft::c2c / r2c / c2r` as bound by
 *      impulse/fft/cpp_pocketfft/pocketfft.nim:35-69 and called from `apply`
 *      (impulse/fft/cpp_pocketfft/pocketfft.nim:235-277), defined in
 *      impulse/fft/cpp_pocketfft/pocketfft_hdronly.h:3272-3390;
 *  (2) the C path, `make_/destroy_/_forward/_backward/_length` for cfft and
 *      rfft plans, impulse/fft/c_pocketfft/pocketfft.h:18-32, bound by
 *      impulse/fft/c_pocketfft/pocketfft.nim:71-81.
 *
 * Semantics follow impulse/fft/cpp_pocketfft/README_pocketfft.md:87-147:
 * strides are in BYTES (negative allowed), `forward` selects exp(-2 pi i ...),
 * transforms are unnormalised and `fct` scales the result once, r2c keeps
 * shape[axis]/2+1 bins along the last entry of `axes`, r2c with forward=0
 * conjugates its output, c2r with forward=1 conjugates its input, and c2r
 * ignores the imaginary part of bin 0 (and of bin n/2 for even n).
 * Arrays hold double / complex<double> (IMPULSE_FFT_F64) or, for the c2c /
 * r2c / c2r kinds, float / complex<float> (IMPULSE_FFT_F32, the reference's
 * T = float instantiation); the arithmetic is FP64 on the GPU either way.
 */
#ifndef IMPULSE_FFT_H
#define IMPULSE_FFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMPULSE_FFT_ABI_VERSION 1

typedef enum impulse_fft_status {
    IMPULSE_FFT_OK = 0,
    IMPULSE_FFT_ERR_INVALID_ARG = 1, /* pocketfft's runtime_error / invalid_argument cases */
    IMPULSE_FFT_ERR_NOMEM = 2,       /* host or device allocation failed (pocketfft bad_alloc / NULL plan) */
    IMPULSE_FFT_ERR_HIP = 3,         /* a HIP runtime call failed */
    IMPULSE_FFT_ERR_UNSUPPORTED = 4, /* valid request this engine does not implement */
    IMPULSE_FFT_ERR_NO_DEVICE = 5    /* no usable gfx950 device */
} impulse_fft_status;

typedef enum impulse_fft_kind {
    IMPULSE_FFT_C2C = 0,
    IMPULSE_FFT_R2C = 1,
    IMPULSE_FFT_C2R = 2,
    /* Adjoints for reverse-mode differentiation (no PocketFFT counterpart).
     * With respect to the real inner products <x,y> = sum x y and
     * <X,Y> = Re sum conj(X) Y over the stored half spectrum:
     *   R2C_ADJOINT maps a half spectrum G (shape[last]/2+1 bins) to the real
     *   array r2c*(G); it equals c2r of the opposite direction with bins
     *   1..(n-1)/2 weighted 1/2 (`shape` is the REAL shape, as for C2R);
     *   C2R_ADJOINT maps a real array g to c2r*(g); it equals r2c of the
     *   opposite direction with those bins weighted 2 (`shape` is the real shape).
     * The adjoint of C2C is C2C of the opposite direction. */
    IMPULSE_FFT_R2C_ADJOINT = 3,
    IMPULSE_FFT_C2R_ADJOINT = 4,
    /* Real-to-real transforms (impulse_fft_r2r_plan_create, impulse_fft_dct/dst). */
    IMPULSE_FFT_DCT = 5,
    IMPULSE_FFT_DST = 6,
    /* FFTPACK-order and Hartley real-to-real transforms (impulse_fft_fftpack_plan_create,
     * impulse_fft_r2r_fftpack, impulse_fft_r2r_separable_hartley, impulse_fft_r2r_genuine_hartley). */
    IMPULSE_FFT_FFTPACK = 7,
    IMPULSE_FFT_HARTLEY = 8,
    IMPULSE_FFT_HARTLEY_GENUINE = 9
} impulse_fft_kind;

typedef enum impulse_fft_dtype {
    IMPULSE_FFT_F64 = 0, /* double / interleaved complex<double> (Nim float64 / Complex64) */
    IMPULSE_FFT_F32 = 1  /* float / interleaved complex<float> (Nim float32 / Complex32): the
                            c2c / r2c / c2r plans and one-shots and the FFT convolution;
                            inputs are widened to FP64, outputs rounded to nearest float */
} impulse_fft_dtype;

/* ------------------------------------------------------------------------- */
/* n-D plans (C++ path)                                                       */
/* ------------------------------------------------------------------------- */

typedef struct impulse_fft_plan_s *impulse_fft_plan;

/* Plan an n-D transform.  `shape` is the array shape; for R2C it is the real
 * INPUT shape and for C2R the real OUTPUT shape (README_pocketfft.md:87-92).
 * The strides describe the input and output arrays this plan will be executed
 * on.  Argument checks mirror pocketfft_hdronly.h:446-477 (ndim >= 1, equal
 * stride ranks, axis < ndim, no repeated axis).
 * Replaces: the plan lookup + checks inside pocketfft::c2c/r2c/c2r
 * (pocketfft_hdronly.h:3272-3390, get_plan at :2655-2706). */
int impulse_fft_plan_create(impulse_fft_plan *plan, int kind, int dtype, size_t ndim,
                            const size_t *shape, const ptrdiff_t *stride_in,
                            const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                            int forward);

/* Execute on DEVICE pointers, asynchronously on `stream` (a hipStream_t, NULL =
 * the default stream).  Calls of one plan share its workspace: calls of the
 * same plan that may overlap must be ordered by the caller (same stream).
 * in == out is allowed for C2C when both stride sets are equal. */
int impulse_fft_execute(impulse_fft_plan plan, const void *data_in, void *data_out, double fct,
                        void *stream);

/* The same as impulse_fft_execute (device pointers, asynchronous on `stream`,
 * same kernels and results) with the workspace taken from a static device
 * array of the library (one per GPU, 64 MiB, held until the process ends)
 * instead of the plan's buffers: the call allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory at call time (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Plans whose workspace exceeds the
 * scratch, and DCT / DST / FFTPACK / Hartley plans, return
 * IMPULSE_FFT_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's array:
 * calls that may overlap on one GPU (other streams, other threads) race and
 * must be ordered by the caller. */
int impulse_fft_execute_noalloc(impulse_fft_plan plan, const void *data_in, void *data_out, double fct,
                                void *stream);

/* Execute on HOST pointers: copies in, runs, copies out, synchronous.  This is
 * the drop-in for pocketfft's host-memory calls.  The copies go through a
 * pinned staging ring of the plan, on a stream of the plan's own; concurrent
 * host calls of one plan are serialised.  in == out needs equal strides and
 * a C2C (or real-to-real) plan, as pocketfft's sanity_check requires. */
int impulse_fft_execute_host(impulse_fft_plan plan, const void *data_in, void *data_out,
                             double fct);

/* Bytes of device workspace this plan holds (twiddles, chirps, scratch). */
size_t impulse_fft_plan_device_bytes(impulse_fft_plan plan);

/* Number of kernel launches one execute issues (HBM passes plus helpers). */
int impulse_fft_plan_num_launches(impulse_fft_plan plan);

/* Describe kernel launch `pass` (0 <= pass < num_launches) of one execute:
 * kernel name (the template instantiation of the last execute once the plan
 * has run), grid and block size, LDS bytes, and the bytes the launch must
 * read (`bytes_in`) and write (`bytes_out`) if every byte moves once. */
int impulse_fft_plan_pass_info(impulse_fft_plan plan, int pass, char *name, size_t name_len,
                               size_t *bytes_in, size_t *bytes_out, unsigned *grid, unsigned *block,
                               size_t *lds_bytes);

/* Per-launch timing: while enabled, the next `max_executes` executes record a
 * HIP event pair around every launch on the execute's stream (0 disables and
 * frees the events).  `profile_read` waits for the recorded events and writes
 * the summed milliseconds of each launch into pass_ms[0..npass) and the
 * number of executes recorded into *executes. */
int impulse_fft_plan_profile(impulse_fft_plan plan, int max_executes);
int impulse_fft_plan_profile_read(impulse_fft_plan plan, double *pass_ms, int npass, int *executes);

int impulse_fft_plan_destroy(impulse_fft_plan plan);

/* One-shot host-pointer calls with the exact argument list of pocketfft::c2c,
 * r2c and c2r (pocketfft_hdronly.h:3272, :3334, :3366 as bound by
 * cpp_pocketfft/pocketfft.nim:35-69), plus `dtype`.  Plans are taken from a
 * thread-safe LRU cache keyed by (kind, dtype, shape, strides, axes,
 * direction, device).  `nthreads` is accepted and ignored: the GPU grid
 * replaces PocketFFT's thread pool (pocketfft_hdronly.h:499-705). */
int impulse_fft_c2c(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                    const ptrdiff_t *stride_out, size_t naxes, const size_t *axes, int forward,
                    int dtype, const void *data_in, void *data_out, double fct, size_t nthreads);
int impulse_fft_r2c(size_t ndim, const size_t *shape_in, const ptrdiff_t *stride_in,
                    const ptrdiff_t *stride_out, size_t naxes, const size_t *axes, int forward,
                    int dtype, const void *data_in, void *data_out, double fct, size_t nthreads);
int impulse_fft_c2r(size_t ndim, const size_t *shape_out, const ptrdiff_t *stride_in,
                    const ptrdiff_t *stride_out, size_t naxes, const size_t *axes, int forward,
                    int dtype, const void *data_in, void *data_out, double fct, size_t nthreads);

/* DCT / DST of types 1-4 over `axes` with the argument list of pocketfft::dct
 * and pocketfft::dst (pocketfft_hdronly.h:3284-3318, bound at
 * cpp_pocketfft/pocketfft.nim:108-132 and applied by `apply(DCTDesc, ...)`,
 * cpp_pocketfft/pocketfft.nim:279-295), plus `dtype` (IMPULSE_FFT_F64).  Real
 * doubles with byte strides on HOST memory; in == out is allowed with equal
 * strides.  Unnormalised; `ortho` applies only the end-point scalings
 * PocketFFT applies (DCT-I: x[0], x[n-1] times sqrt2 before and y[0],
 * y[n-1] times sqrt2/2 after; DCT-II and DST-II: y[0] times sqrt2/2;
 * DCT-III and DST-III: x[0] times sqrt2; ignored for DST-I and type 4) and
 * the 1/sqrt normalisation is the caller's `fct`, which scales the first
 * axis only.  DCT-I needs at least 2 points per axis. */
int impulse_fft_dct(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                    const ptrdiff_t *stride_out, size_t naxes, const size_t *axes, int type,
                    int dtype, const void *data_in, void *data_out, double fct, int ortho,
                    size_t nthreads);
int impulse_fft_dst(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                    const ptrdiff_t *stride_out, size_t naxes, const size_t *axes, int type,
                    int dtype, const void *data_in, void *data_out, double fct, int ortho,
                    size_t nthreads);
/* The same as a plan (kind IMPULSE_FFT_DCT or IMPULSE_FFT_DST), executed with
 * impulse_fft_execute (device pointers) or impulse_fft_execute_host and freed
 * with impulse_fft_plan_destroy.  Per-launch profiling is not kept for these
 * plans (IMPULSE_FFT_ERR_UNSUPPORTED); num_launches counts every launch. */
int impulse_fft_r2r_plan_create(impulse_fft_plan *plan, int kind, int type, int ortho, int dtype,
                                size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                                const ptrdiff_t *stride_out, size_t naxes, const size_t *axes);

/* pocketfft::r2r_fftpack (pocketfft_hdronly.h:3392-3403), which impulse declares
 * but does not export (cpp_pocketfft/pocketfft.nim:71-106).  Along each axis in
 * turn, between real data and FFTPACK halfcomplex [r0, r1, i1, r2, i2, ...
 * (, r_{n/2})].  As in the reference (ExecR2R, :3252-3270), `forward` picks
 * the operation: forward != 0 writes the halfcomplex X[k] = sum_n x[n]
 * e^{-2 pi i k n / n_axis}; forward == 0 reads halfcomplex data and writes the
 * unnormalised inverse of its Hermitian extension.  `real2hermitian` only
 * adds sign flips of positions 2, 4, ... (the imaginary slots): on the input
 * of a forward call when it is 0, on the output of a backward call when it is
 * 1.  Real doubles with byte strides on HOST memory, in == out allowed with
 * equal strides; fct scales the first axis. */
int impulse_fft_r2r_fftpack(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                            const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                            int real2hermitian, int forward, int dtype, const void *data_in,
                            void *data_out, double fct, size_t nthreads);
/* pocketfft::r2r_separable_hartley (pocketfft_hdronly.h:3405-3415): along each
 * axis H[k] = Re X[k] + Im X[k] with X the forward DFT along that axis. */
int impulse_fft_r2r_separable_hartley(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                                      const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                                      int dtype, const void *data_in, void *data_out, double fct,
                                      size_t nthreads);
/* pocketfft::r2r_genuine_hartley (pocketfft_hdronly.h:3417-3445): H = Re X + Im X
 * with X the forward DFT over all `axes` at once (one axis: separable). */
int impulse_fft_r2r_genuine_hartley(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                                    const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                                    int dtype, const void *data_in, void *data_out, double fct,
                                    size_t nthreads);
/* The same three as plans: kind IMPULSE_FFT_FFTPACK (with real2hermitian and
 * forward), IMPULSE_FFT_HARTLEY or IMPULSE_FFT_HARTLEY_GENUINE (flags ignored),
 * run by impulse_fft_execute (device pointers) or impulse_fft_execute_host. */
int impulse_fft_fftpack_plan_create(impulse_fft_plan *plan, int kind, int real2hermitian, int forward,
                                    int dtype, size_t ndim, const size_t *shape,
                                    const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                                    size_t naxes, const size_t *axes);

/* Plan cache control (PocketFFT's is a fixed 16-entry LRU, pocketfft_hdronly.h:50-52).
 * The one-shot calls and the C path's make_cfft_plan / make_rfft_plan take their
 * device plans from it (the C path's real plans from a second LRU of the same
 * capacity keyed by length).  cache_stats reports the plan lookups that found
 * a plan (hits) and that had to build one (misses) since the last clear. */
int impulse_fft_cache_set_capacity(size_t entries);
size_t impulse_fft_cache_size(void);
int impulse_fft_cache_clear(void);
int impulse_fft_cache_stats(size_t *hits, size_t *misses);

/* ------------------------------------------------------------------------- */
/* FFT convolution with a backward pass (no PocketFFT counterpart)            */
/* ------------------------------------------------------------------------- */

/* "same" 2-D convolution of a batch of images [batch, height, width] with one
 * kernel [kh, kw] (origin at its centre ((kh-1)/2, (kw-1)/2), zero outside the
 * image), through r2c / spectrum product / c2r on a 13-smooth grid of at
 * least (height+kh-1) x (width+kw-1).  dtype IMPULSE_FFT_F32 or _F64 gives
 * the element type of x, k, y and the gradients; arithmetic is FP64.  All
 * pointers are DEVICE pointers, C-contiguous; calls are asynchronous on
 * `stream` and calls on one conv object must be ordered by the caller. */
typedef struct impulse_fft_conv2d_s *impulse_fft_conv2d;
int impulse_fft_conv2d_create(impulse_fft_conv2d *conv, int batch, int height, int width, int kh, int kw,
                              int dtype);
int impulse_fft_conv2d_forward(impulse_fft_conv2d conv, const void *x, const void *k, void *y, void *stream);
/* Gradients of a scalar loss: given gy = dL/dy, writes gx = dL/dx and
 * gk = dL/dk (either may be NULL).  Built from the adjoint transforms; the
 * spectra of x and k are recomputed rather than kept from the forward call. */
int impulse_fft_conv2d_backward(impulse_fft_conv2d conv, const void *x, const void *k, const void *gy, void *gx,
                                void *gk, void *stream);
int impulse_fft_conv2d_grid(impulse_fft_conv2d conv, size_t *P, size_t *Q);
int impulse_fft_conv2d_destroy(impulse_fft_conv2d conv);

/* ------------------------------------------------------------------------- */
/* upfirdn: upsample, FIR filter, downsample (impulse/signal.nim:545-594)     */
/* ------------------------------------------------------------------------- */

/* Replaces the body of impulse/signal.nim:545-594, `upfirdn(t, h, up, down)`
 * = t.upsample(up).convolve(h, mode = full, down = down), which resample
 * (impulse/signal.nim:690-744, 746-809) calls.  For each of `batch` signals
 * of length nt, the upsampled signal (up-1 zeros after every sample, length
 * (nt-1) up + 1) is convolved in full with the nh taps h, and every down-th
 * sample from the first is kept: nout = ceil(((nt-1) up + nh) / down)
 * outputs per signal.  FP64 throughout.  The taps are copied at creation
 * (host pointer).  algorithm: IMPULSE_FFT_UPFIRDN_AUTO picks by a cost
 * model, _DIRECT sums the polyphase taps per output, _FFT runs overlap-save
 * blocks through the engine's r2c / c2r plans. */
enum impulse_fft_upfirdn_algorithm {
    IMPULSE_FFT_UPFIRDN_AUTO = 0,
    IMPULSE_FFT_UPFIRDN_DIRECT = 1,
    IMPULSE_FFT_UPFIRDN_FFT = 2
};
typedef struct impulse_fft_upfirdn_s *impulse_fft_upfirdn;
int impulse_fft_upfirdn_create(impulse_fft_upfirdn *plan, size_t batch, size_t nt, const double *h, size_t nh,
                               int up, int down, int algorithm);
/* Outputs per signal, the algorithm chosen, and for _FFT the block FFT length and blocks per signal (0 otherwise). */
int impulse_fft_upfirdn_info(impulse_fft_upfirdn plan, size_t *nout, int *algorithm, size_t *fft_length,
                             size_t *blocks);
/* t: [batch, nt], y: [batch, nout], C-contiguous DEVICE pointers; asynchronous on `stream`.
 * Calls on one plan must be ordered by the caller (the FFT algorithm keeps its workspace in the plan). */
int impulse_fft_upfirdn_execute(impulse_fft_upfirdn plan, const double *t, double *y, void *stream);
/* Same on HOST pointers; synchronous. */
int impulse_fft_upfirdn_execute_host(impulse_fft_upfirdn plan, const double *t, double *y);
int impulse_fft_upfirdn_destroy(impulse_fft_upfirdn plan);

/* ------------------------------------------------------------------------- */
/* Ragged batches of real transforms (no PocketFFT counterpart)               */
/* ------------------------------------------------------------------------- */

/* `count` float32 signals of lengths[i] (any length >= 1) at in_offsets[i]
 * floats into one input buffer; each gets its lengths[i]/2+1 complex128 bins
 * at out_offsets[i] elements into one output buffer -- the same bins as
 * impulse_fft_r2c on the signal widened to float64 (forward = 0 conjugates
 * them, as PocketFFT's r2c does), scaled by `fct`.  Arithmetic is FP64.
 * Nearest reference: one r2c call per signal, cpp_pocketfft/pocketfft.nim:235-277
 * over pocketfft_hdronly.h:3320-3350, each with its own plan lookup
 * (pocketfft_hdronly.h:2655-2706).
 *
 * One execute is one hipGraph launch (a queue reset and one persistent
 * kernel in which each workgroup claims signals, longest first).  The graph is
 * captured at the first execute for given (in, out, fct) and replayed after.
 * Buffers are DEVICE pointers and must not overlap. */
typedef struct impulse_fft_ragged_s *impulse_fft_ragged;
int impulse_fft_ragged_create(impulse_fft_ragged *batch, size_t count, const size_t *lengths,
                              const size_t *in_offsets, const size_t *out_offsets, int forward);
int impulse_fft_ragged_execute(impulse_fft_ragged batch, const float *in, void *out, double fct, void *stream);
/* Autotuner: runs every candidate configuration (workgroup size, LDS budget
 * per transform, radix order, register budget) `reps` times on these buffers,
 * keeps the fastest and reports its time per execute (ms).  Synchronizes
 * `stream`; `out` is overwritten. */
int impulse_fft_ragged_tune(impulse_fft_ragged batch, const float *in, void *out, void *stream, int reps,
                            double *best_ms);
/* Pick a configuration by hand: threads 256/512/1024, lds_points
 * 2048/4096/8192 (longest transform kept in LDS; at most 8 per thread),
 * radix_order 0-3 (bit 0: odd radices before the powers of two; bit 1: powers
 * of two as radix-8 rather than radix-16 stages),
 * waves_per_simd 0 (registers as the workgroup size allows), 4 (at most 128
 * registers per lane: more resident workgroups, some spills) or, with 256
 * threads, 3 (at most 168). */
int impulse_fft_ragged_configure(impulse_fft_ragged batch, int threads, int lds_points, int radix_order,
                                 int waves_per_simd);
/* Arithmetic of the LDS stages (no PocketFFT counterpart):
 *   IMPULSE_FFT_RAGGED_FP64        every stage in FP64 (the default);
 *   IMPULSE_FFT_RAGGED_MFMA_BF16X  the radix-16 stages of transforms that span
 *       whole wavefronts (64 or more threads a transform) run as exact
 *       bf16-slice products on the MFMA units -- the values of a stage cut into
 *       five 8-bit slices, products accumulated exactly in FP32 and combined in
 *       FP64 (error ~1e-12 relative) -- with powers of two planned as radix-16
 *       stages whatever the radix order; twiddles, the Bluestein chirp and its
 *       spectrum, and the other radices stay FP64.
 * Applies from the next execute (the graph is re-captured). */
typedef enum impulse_fft_ragged_precision {
    IMPULSE_FFT_RAGGED_FP64 = 0,
    IMPULSE_FFT_RAGGED_MFMA_BF16X = 1
} impulse_fft_ragged_precision;
int impulse_fft_ragged_set_precision(impulse_fft_ragged batch, int precision);
/* Current configuration, persistent grid size, nominal flops of one execute
 * (2.5 n log2 n per signal) and device scratch bytes.  Any pointer may be NULL. */
int impulse_fft_ragged_info(impulse_fft_ragged batch, int *threads, int *lds_points, int *radix_order,
                            int *waves_per_simd, int *grid, double *flops, size_t *scratch_bytes);
int impulse_fft_ragged_destroy(impulse_fft_ragged batch);

/* ------------------------------------------------------------------------- */
/* Matrix-core complex FFTs (no PocketFFT counterpart)                        */
/* ------------------------------------------------------------------------- */

/* `batch` contiguous complex64 transforms of `length` (a power of two from 256
 * to 2^21) whose 256-point factors run as two radix-16 stages of 16x16x16
 * complex matrix products on the MFMA units: one pass for 256, a four-step
 * 256 x (length/256) otherwise, whose (length/256)-point row transforms run in
 * the same launch up to 4096 (8192 with IMPULSE_FFT_MFMA_F32), on the MFMA
 * units for 65536 and in an FP64 LDS row kernel for the other lengths.
 * Nearest reference: pocketfft::c2c<float> over one axis
 * (pocketfft_hdronly.h:3272-3282, impulse's FFTDesc[Complex32] through
 * cpp_pocketfft/pocketfft.nim:235-277), here with the precision stated:
 *   IMPULSE_FFT_MFMA_F32    FP32 MFMA, FP32 accumulation, FP64-generated
 *                           twiddles applied as split-FP32 products;
 *                           complex64 output (error ~1e-7 relative).
 *   IMPULSE_FFT_MFMA_BF16X  data cut into five 8-bit bf16 slices per stage,
 *                           slice products accumulated exactly in FP32 by
 *                           bf16 MFMA and combined in FP64; FP64 twiddles;
 *                           complex128 output (error ~1e-12 relative).
 * forward = 1 computes exp(-2 pi i jk/n); fct scales the output (in FP64).
 * Buffers are DEVICE pointers; in and out must not overlap.  Two-pass plans
 * (longer than 4096 points, or 8192 with IMPULSE_FFT_MFMA_F32) hold a workspace
 * of one intermediate (batch * length elements of the output type). */
typedef enum impulse_fft_mfma_precision {
    IMPULSE_FFT_MFMA_F32 = 0,
    IMPULSE_FFT_MFMA_BF16X = 1
} impulse_fft_mfma_precision;
typedef struct impulse_fft_mfma_s *impulse_fft_mfma;
int impulse_fft_mfma_create(impulse_fft_mfma *plan, size_t length, size_t batch, int precision);
int impulse_fft_mfma_execute(impulse_fft_mfma plan, const void *in, void *out, int forward, double fct,
                             void *stream);
int impulse_fft_mfma_workspace_bytes(impulse_fft_mfma plan, size_t *bytes);
int impulse_fft_mfma_destroy(impulse_fft_mfma plan);

/* ------------------------------------------------------------------------- */
/* 1-D in-place plans on host memory (C path)                                 */
/* ------------------------------------------------------------------------- */

typedef struct impulse_cfft_plan_s *impulse_cfft_plan;
typedef struct impulse_rfft_plan_s *impulse_rfft_plan;

/* Replaces make_cfft_plan / destroy_cfft_plan / cfft_forward / cfft_backward /
 * cfft_length (c_pocketfft/pocketfft.h:18-23).  `c` holds `length` interleaved
 * complex doubles; the transform is done in place and scaled by `fct`.
 * make_* returns NULL on failure (length 0, no device, out of memory);
 * *_forward/_backward return 0 on success and -1 on failure, as PocketFFT.
 * The device plans come from the plan cache, so building and destroying a
 * plan per call (as impulse's fft() does) costs a lookup after the first. */
impulse_cfft_plan impulse_make_cfft_plan(size_t length);
void impulse_destroy_cfft_plan(impulse_cfft_plan plan);
int impulse_cfft_forward(impulse_cfft_plan plan, double c[], double fct);
int impulse_cfft_backward(impulse_cfft_plan plan, double c[], double fct);
size_t impulse_cfft_length(impulse_cfft_plan plan);

/* Replaces make_rfft_plan / destroy_rfft_plan / rfft_forward / rfft_backward /
 * rfft_length (c_pocketfft/pocketfft.h:25-31).  `c` holds `length` doubles in
 * FFTPACK halfcomplex order [r0, r1, i1, r2, i2, ...(, r_{n/2})], the layout
 * impulse/fft/c_pocketfft/pocketfft.nim:224-243 documents. */
impulse_rfft_plan impulse_make_rfft_plan(size_t length);
void impulse_destroy_rfft_plan(impulse_rfft_plan plan);
int impulse_rfft_forward(impulse_rfft_plan plan, double c[], double fct);
int impulse_rfft_backward(impulse_rfft_plan plan, double c[], double fct);
size_t impulse_rfft_length(impulse_rfft_plan plan);

/* ------------------------------------------------------------------------- */
/* Diagnostics                                                                */
/* ------------------------------------------------------------------------- */

/* Static text for a status code. */
const char *impulse_fft_status_string(int status);
/* Message of the last failing call on this thread ("" if none). */
const char *impulse_fft_last_error(void);
/* 0 if a gfx950 device is usable, else a status code (no kernel is launched). */
int impulse_fft_device_check(void);

#ifdef __cplusplus
}
#endif
#endif /* IMPULSE_FFT_H */

// C-ABI shim over the reference's own PocketFFT C++ header, compiled in place
// from the reference tree by oracle/Makefile into oracle/_ref/ (never copied
// into this repository).  TEST INFRASTRUCTURE ONLY: it lets the tests and the
// benchmark's CPU-baseline leg call pocketfft::c2c/r2c/c2r and dct/dst exactly
// as impulse/fft/cpp_pocketfft/pocketfft.nim:35-69,108-132 binds them, with
// byte strides.
//
// Return codes: 0 ok, 1 exception (invalid argument / runtime error).
#include <complex>
#include <cstddef>
#include <vector>

#include "pocketfft_hdronly.h"

namespace {
pocketfft::shape_t mk_shape(size_t n, const size_t *p) { return pocketfft::shape_t(p, p + n); }
pocketfft::stride_t mk_stride(size_t n, const ptrdiff_t *p) { return pocketfft::stride_t(p, p + n); }
}  // namespace

extern "C" {

int ref_c2c_f64(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout,
                size_t naxes, const size_t *axes, int forward, const double *in, double *out,
                double fct, size_t nthreads)
{
    try {
        pocketfft::c2c<double>(mk_shape(ndim, shape), mk_stride(ndim, sin), mk_stride(ndim, sout),
                               mk_shape(naxes, axes), forward != 0,
                               reinterpret_cast<const std::complex<double> *>(in),
                               reinterpret_cast<std::complex<double> *>(out), fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_r2c_f64(size_t ndim, const size_t *shape_in, const ptrdiff_t *sin, const ptrdiff_t *sout,
                size_t naxes, const size_t *axes, int forward, const double *in, double *out,
                double fct, size_t nthreads)
{
    try {
        pocketfft::r2c<double>(mk_shape(ndim, shape_in), mk_stride(ndim, sin), mk_stride(ndim, sout),
                               mk_shape(naxes, axes), forward != 0, in,
                               reinterpret_cast<std::complex<double> *>(out), fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_c2r_f64(size_t ndim, const size_t *shape_out, const ptrdiff_t *sin, const ptrdiff_t *sout,
                size_t naxes, const size_t *axes, int forward, const double *in, double *out,
                double fct, size_t nthreads)
{
    try {
        pocketfft::c2r<double>(mk_shape(ndim, shape_out), mk_stride(ndim, sin), mk_stride(ndim, sout),
                               mk_shape(naxes, axes), forward != 0,
                               reinterpret_cast<const std::complex<double> *>(in), out, fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_c2c_f32(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout,
                size_t naxes, const size_t *axes, int forward, const float *in, float *out,
                float fct, size_t nthreads)
{
    try {
        pocketfft::c2c<float>(mk_shape(ndim, shape), mk_stride(ndim, sin), mk_stride(ndim, sout),
                              mk_shape(naxes, axes), forward != 0,
                              reinterpret_cast<const std::complex<float> *>(in),
                              reinterpret_cast<std::complex<float> *>(out), fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_r2c_f32(size_t ndim, const size_t *shape_in, const ptrdiff_t *sin, const ptrdiff_t *sout,
                size_t naxes, const size_t *axes, int forward, const float *in, float *out,
                float fct, size_t nthreads)
{
    try {
        pocketfft::r2c<float>(mk_shape(ndim, shape_in), mk_stride(ndim, sin), mk_stride(ndim, sout),
                              mk_shape(naxes, axes), forward != 0, in,
                              reinterpret_cast<std::complex<float> *>(out), fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_c2r_f32(size_t ndim, const size_t *shape_out, const ptrdiff_t *sin, const ptrdiff_t *sout,
                size_t naxes, const size_t *axes, int forward, const float *in, float *out,
                float fct, size_t nthreads)
{
    try {
        pocketfft::c2r<float>(mk_shape(ndim, shape_out), mk_stride(ndim, sin), mk_stride(ndim, sout),
                              mk_shape(naxes, axes), forward != 0,
                              reinterpret_cast<const std::complex<float> *>(in), out, fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_dct_f64(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout, size_t naxes,
                const size_t *axes, int type, const double *in, double *out, double fct, int ortho, size_t nthreads)
{
    try {
        pocketfft::dct<double>(mk_shape(ndim, shape), mk_stride(ndim, sin), mk_stride(ndim, sout),
                               mk_shape(naxes, axes), type, in, out, fct, ortho != 0, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_dst_f64(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout, size_t naxes,
                const size_t *axes, int type, const double *in, double *out, double fct, int ortho, size_t nthreads)
{
    try {
        pocketfft::dst<double>(mk_shape(ndim, shape), mk_stride(ndim, sin), mk_stride(ndim, sout),
                               mk_shape(naxes, axes), type, in, out, fct, ortho != 0, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_r2r_fftpack_f64(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout,
                        size_t naxes, const size_t *axes, int r2h, int forward, const double *in, double *out,
                        double fct, size_t nthreads)
{
    try {
        pocketfft::r2r_fftpack<double>(mk_shape(ndim, shape), mk_stride(ndim, sin), mk_stride(ndim, sout),
                                       mk_shape(naxes, axes), r2h != 0, forward != 0, in, out, fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_r2r_separable_hartley_f64(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout,
                                  size_t naxes, const size_t *axes, const double *in, double *out, double fct,
                                  size_t nthreads)
{
    try {
        pocketfft::r2r_separable_hartley<double>(mk_shape(ndim, shape), mk_stride(ndim, sin), mk_stride(ndim, sout),
                                                 mk_shape(naxes, axes), in, out, fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

int ref_r2r_genuine_hartley_f64(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout,
                                size_t naxes, const size_t *axes, const double *in, double *out, double fct,
                                size_t nthreads)
{
    try {
        pocketfft::r2r_genuine_hartley<double>(mk_shape(ndim, shape), mk_stride(ndim, sin), mk_stride(ndim, sout),
                                               mk_shape(naxes, axes), in, out, fct, nthreads);
    } catch (...) { return 1; }
    return 0;
}

}  // extern "C"

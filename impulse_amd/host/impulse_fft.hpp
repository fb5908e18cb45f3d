// impulse_fft.hpp -- C++ host mirror of impulse's FFT API over the C ABI of
// the MI355X engine (include/impulse_fft.h).  Header-only; link against
// impulse_amd/libimpulse_fft.so.
//
// Names, argument meaning and error messages follow the reference's Nim API:
//   C path   impulse/fft/c_pocketfft/pocketfft.nim:84-360 (NormalizeKind,
//            initNormalize, FFTPlanReal/Complex, rfft, fft_impl, fft, rfft_packed,
//            ifft, unpackFFT, symmetrize, symmTargetSize)
//   C++ path impulse/fft/cpp_pocketfft/pocketfft.nim:137-277 (DataDesc, FFTDesc,
//            apply)
// Nim's `fft(var openArray)` (in place) is `fft_inplace` here, since C++
// cannot overload on var-ness the way Nim does.  Failures throw
// std::runtime_error with the reference's messages ("Forward FFT calculation
// failed." / "Backward FFT calculation failed.", pocketfft.nim:206-214), or
// with the engine's message for argument errors (pocketfft_hdronly.h:446-477).
#pragma once

#include <cmath>
#include <complex>
#include <cstddef>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/impulse_fft.h"

namespace impulse {
namespace fft {

using Complex64 = std::complex<double>;

// ---------------------------------------------------------------------------
// normalisation (c_pocketfft/pocketfft.nim:189-204)
// ---------------------------------------------------------------------------

enum NormalizeKind { nkBackward, nkOrtho, nkForward, nkCustom };

struct Normalize {
    NormalizeKind kind;
    double value;
};

inline Normalize initNormalize(NormalizeKind kind, bool forward, double value, size_t length)
{
    const double n = (double)length;
    switch (kind) {
        case nkBackward: return {kind, forward ? 1.0 : 1.0 / n};
        case nkForward: return {kind, forward ? 1.0 / n : 1.0};
        case nkOrtho: return {kind, 1.0 / std::sqrt(n)};
        default: return {nkCustom, value};
    }
}

inline void check_status(int rc)
{
    if (rc != IMPULSE_FFT_OK)
        throw std::runtime_error(std::string(impulse_fft_status_string(rc)) + ": " + impulse_fft_last_error());
}

// ---------------------------------------------------------------------------
// 1-D plans (c_pocketfft/pocketfft.nim:84-114)
// ---------------------------------------------------------------------------

struct FFTPlanReal {
    impulse_rfft_plan pocket = nullptr;
    size_t length = 0;
    static FFTPlanReal init(size_t length)
    {
        FFTPlanReal p;
        p.pocket = impulse_make_rfft_plan(length);
        if (!p.pocket) throw std::runtime_error(std::string("make_rfft_plan failed: ") + impulse_fft_last_error());
        p.length = length;
        return p;
    }
    FFTPlanReal() = default;
    FFTPlanReal(FFTPlanReal &&o) noexcept : pocket(o.pocket), length(o.length) { o.pocket = nullptr; }
    FFTPlanReal &operator=(FFTPlanReal &&o) noexcept
    {
        std::swap(pocket, o.pocket);
        std::swap(length, o.length);
        return *this;
    }
    FFTPlanReal(const FFTPlanReal &) = delete;
    ~FFTPlanReal()
    {
        if (pocket) impulse_destroy_rfft_plan(pocket);
    }
};

struct FFTPlanComplex {
    impulse_cfft_plan pocket = nullptr;
    size_t length = 0;
    static FFTPlanComplex init(size_t length)
    {
        FFTPlanComplex p;
        p.pocket = impulse_make_cfft_plan(length);
        if (!p.pocket) throw std::runtime_error(std::string("make_cfft_plan failed: ") + impulse_fft_last_error());
        p.length = length;
        return p;
    }
    FFTPlanComplex() = default;
    FFTPlanComplex(FFTPlanComplex &&o) noexcept : pocket(o.pocket), length(o.length) { o.pocket = nullptr; }
    FFTPlanComplex &operator=(FFTPlanComplex &&o) noexcept
    {
        std::swap(pocket, o.pocket);
        std::swap(length, o.length);
        return *this;
    }
    FFTPlanComplex(const FFTPlanComplex &) = delete;
    ~FFTPlanComplex()
    {
        if (pocket) impulse_destroy_cfft_plan(pocket);
    }
};

namespace detail {
inline void call(const FFTPlanReal &p, double *data, bool forward, double fct)
{
    const int err = forward ? impulse_rfft_forward(p.pocket, data, fct) : impulse_rfft_backward(p.pocket, data, fct);
    if (err != 0)
        throw std::runtime_error(forward ? "Forward FFT calculation failed." : "Backward FFT calculation failed.");
}
inline void call(const FFTPlanComplex &p, Complex64 *data, bool forward, double fct)
{
    double *c = reinterpret_cast<double *>(data);
    const int err = forward ? impulse_cfft_forward(p.pocket, c, fct) : impulse_cfft_backward(p.pocket, c, fct);
    if (err != 0)
        throw std::runtime_error(forward ? "Forward FFT calculation failed." : "Backward FFT calculation failed.");
}
}  // namespace detail

constexpr double kInf = std::numeric_limits<double>::infinity();

// In place, FFTPACK halfcomplex result (pocketfft.nim:216-288).
inline void rfft(double *data, size_t length, bool forward, NormalizeKind normalize = nkBackward,
                 double normValue = kInf)
{
    const Normalize norm = initNormalize(normalize, forward, normValue, length);
    FFTPlanReal plan = FFTPlanReal::init(length);
    detail::call(plan, data, forward, norm.value);
}

// In place on float (packed) or complex data (pocketfft.nim:290-303).
template <class T>
void fft_impl(T *data, size_t length, bool forward, NormalizeKind normalize = nkBackward, double normValue = kInf)
{
    const Normalize norm = initNormalize(normalize, forward, normValue, length);
    if constexpr (std::is_same<T, double>::value) {
        FFTPlanReal plan = FFTPlanReal::init(length);
        detail::call(plan, data, forward, norm.value);
    } else {
        static_assert(std::is_same<T, Complex64>::value, "fft_impl takes double or std::complex<double>");
        FFTPlanComplex plan = FFTPlanComplex::init(length);
        detail::call(plan, data, forward, norm.value);
    }
}

// Nim `fft(data: var openArray[T], forward)` (pocketfft.nim:305-310).
template <class T>
void fft_inplace(std::vector<T> &data, bool forward = true, NormalizeKind normalize = nkBackward,
                 double normValue = kInf)
{
    fft_impl(data.data(), data.size(), forward, normalize, normValue);
}

inline std::vector<Complex64> unpackFFT(const std::vector<double> &data)
{
    const size_t n = data.size();
    const size_t out_len = (n % 2) ? (n + 1) / 2 : (n + 2) / 2;
    std::vector<Complex64> res(out_len);
    res[0] = Complex64(data[0], 0.0);
    for (size_t k = 1; k < out_len; ++k)
        res[k] = Complex64(data[2 * k - 1], (2 * k < n) ? data[2 * k] : 0.0);
    return res;
}

template <class T> size_t symmTargetSize(const std::vector<T> &data)
{
    if constexpr (std::is_same<T, double>::value) return data.size();
    else return data.size() * 2 - (data.back().imag() == 0.0 ? 2 : 1);
}

template <class T> std::vector<Complex64> symmetrize(const std::vector<T> &data)
{
    const size_t out_len = symmTargetSize(data);
    std::vector<Complex64> res(out_len);
    if constexpr (std::is_same<T, double>::value) {
        auto u = unpackFFT(data);
        for (size_t i = 0; i < u.size() && i < out_len; ++i) res[i] = u[i];
    } else {
        for (size_t i = 0; i < data.size() && i < out_len; ++i) res[i] = data[i];
    }
    size_t k = out_len - 1;
    for (size_t i = 1; i < (out_len + 1) / 2; ++i) res[k--] = std::conj(res[i]);
    return res;
}

inline std::vector<double> rfft_packed(const std::vector<double> &data, bool forward = true,
                                       NormalizeKind normalize = nkBackward, double normValue = kInf)
{
    std::vector<double> res(data);
    rfft(res.data(), res.size(), forward, normalize, normValue);
    return res;
}

inline std::vector<Complex64> rfft(const std::vector<double> &data, bool forward = true,
                                   NormalizeKind normalize = nkBackward, double normValue = kInf)
{
    std::vector<double> res(data);
    fft_impl(res.data(), res.size(), forward, normalize, normValue);
    return unpackFFT(res);
}

// Out of place (pocketfft.nim:333-347; normalisation honoured as the Tensor overloads do).
template <class T>
std::vector<Complex64> fft(const std::vector<T> &data, bool forward = true, NormalizeKind normalize = nkBackward,
                           double normValue = kInf)
{
    if constexpr (std::is_same<T, double>::value) {
        return symmetrize(rfft_packed(data, forward, normalize, normValue));
    } else {
        std::vector<Complex64> res(data.begin(), data.end());
        fft_impl(res.data(), res.size(), forward, normalize, normValue);
        return res;
    }
}

template <class T>
std::vector<Complex64> ifft(const std::vector<T> &data, bool backward = true, NormalizeKind normalize = nkBackward,
                            double normValue = kInf)
{
    return fft(data, !backward, normalize, normValue);
}

// ---------------------------------------------------------------------------
// C++ path (cpp_pocketfft/pocketfft.nim:137-277)
// ---------------------------------------------------------------------------

template <class T> struct DataDesc {
    std::vector<size_t> shape;
    std::vector<ptrdiff_t> stride;   // bytes
    T *buf = nullptr;

    // stride in elements of T between neighbours along each axis (pocketfft.nim:158-176)
    template <class S1, class S2>
    static DataDesc init(T *buffer, const std::vector<S1> &shape_, const std::vector<S2> &stride_)
    {
        if (shape_.size() != stride_.size()) throw std::runtime_error("stride dimension mismatch");
        if (!buffer) throw std::runtime_error("buffer is nil");
        DataDesc d;
        d.buf = buffer;
        for (size_t i = 0; i < shape_.size(); ++i) {
            d.shape.push_back((size_t)shape_[i]);
            d.stride.push_back((ptrdiff_t)stride_[i] * (ptrdiff_t)sizeof(T));
        }
        return d;
    }
    // C-contiguous (pocketfft.nim:178-199)
    template <class S1> static DataDesc init(T *buffer, const std::vector<S1> &shape_)
    {
        if (!buffer) throw std::runtime_error("buffer is nil");
        DataDesc d;
        d.buf = buffer;
        d.shape.assign(shape_.begin(), shape_.end());
        d.stride.assign(shape_.size(), 0);
        ptrdiff_t acc = sizeof(T);
        for (size_t i = shape_.size(); i-- > 0;) {
            d.stride[i] = acc;
            acc *= (ptrdiff_t)shape_[i];
        }
        return d;
    }
};

template <class T> struct FFTDesc {
    std::vector<size_t> axes;
    T scalingFactor = 1;
    size_t nthreads = 1;   // accepted, ignored: the GPU grid replaces PocketFFT's thread pool
    bool forward = true;

    static FFTDesc init(const std::vector<int> &axes_, bool forward_, T scalingFactor_ = 1, size_t nthreads_ = 1)
    {
        FFTDesc f;
        for (int a : axes_) f.axes.push_back((size_t)a);
        f.forward = forward_;
        f.scalingFactor = scalingFactor_;
        f.nthreads = nthreads_;
        return f;
    }
};

template <class T> struct is_complex : std::false_type {};
template <class T> struct is_complex<std::complex<T>> : std::true_type {};
template <class T> struct scalar_of { using type = T; };
template <class T> struct scalar_of<std::complex<T>> { using type = T; };

// IMPULSE_FFT_F64 for double / complex<double>, IMPULSE_FFT_F32 for float /
// complex<float> (the reference's `T: SomeFloat`, cpp_pocketfft/pocketfft.nim:35-69)
template <class T> constexpr int dtype_of()
{
    using S = typename scalar_of<T>::type;
    static_assert(std::is_same<S, double>::value || std::is_same<S, float>::value,
                  "the FFT takes double / complex<double> or float / complex<float>");
    return std::is_same<S, float>::value ? IMPULSE_FFT_F32 : IMPULSE_FFT_F64;
}

// c2c / r2c / c2r by element types.  c2r gets the real OUTPUT shape
// (README_pocketfft.md:89-92).  Input and output share one precision; the
// engine computes in FP64 for both.
template <class F, class In, class Out> void apply(const FFTDesc<F> &fft, DataDesc<Out> &descOut, const DataDesc<In> &descIn)
{
    static_assert(std::is_same<typename scalar_of<In>::type, typename scalar_of<Out>::type>::value,
                  "input and output must have the same precision");
    constexpr int dt = dtype_of<In>();
    const double fct = (double)fft.scalingFactor;
    const int fwd = fft.forward ? 1 : 0;
    if constexpr (is_complex<In>::value && is_complex<Out>::value) {
        check_status(impulse_fft_c2c(descIn.shape.size(), descIn.shape.data(), descIn.stride.data(),
                                     descOut.stride.data(), fft.axes.size(), fft.axes.data(), fwd, dt,
                                     descIn.buf, descOut.buf, fct, fft.nthreads));
    } else if constexpr (is_complex<Out>::value) {
        check_status(impulse_fft_r2c(descIn.shape.size(), descIn.shape.data(), descIn.stride.data(),
                                     descOut.stride.data(), fft.axes.size(), fft.axes.data(), fwd, dt,
                                     descIn.buf, descOut.buf, fct, fft.nthreads));
    } else if constexpr (is_complex<In>::value) {
        check_status(impulse_fft_c2r(descOut.shape.size(), descOut.shape.data(), descIn.stride.data(),
                                     descOut.stride.data(), fft.axes.size(), fft.axes.data(), fwd, dt,
                                     descIn.buf, descOut.buf, fct, fft.nthreads));
    } else {
        static_assert(is_complex<In>::value || is_complex<Out>::value, "Not implemented");
    }
}

// DCT of dctType (1-4) over axes (cpp_pocketfft/pocketfft.nim:151-156,217-233);
// `sine` selects the DST of the same type (pocketfft::dst).
template <class T> struct DCTDesc {
    std::vector<size_t> axes;
    int dctType = 2;
    T scalingFactor = 1;
    size_t nthreads = 1;   // accepted, ignored
    bool ortho = false;
    bool sine = false;

    static DCTDesc init(const std::vector<int> &axes_, int dctType_ = 2, bool ortho_ = false, T scalingFactor_ = 1,
                        size_t nthreads_ = 1, bool sine_ = false)
    {
        DCTDesc d;
        for (int a : axes_) d.axes.push_back((size_t)a);
        d.dctType = dctType_;
        d.ortho = ortho_;
        d.scalingFactor = scalingFactor_;
        d.nthreads = nthreads_;
        d.sine = sine_;
        return d;
    }
};

// cpp_pocketfft/pocketfft.nim:279-295: pocketfft::dct (or dst) over descIn's shape.
template <class F, class T> void apply(const DCTDesc<F> &dct, DataDesc<T> &descOut, const DataDesc<T> &descIn)
{
    static_assert(std::is_same<T, double>::value && std::is_same<F, double>::value, "the engine computes in FP64");
    auto fn = dct.sine ? impulse_fft_dst : impulse_fft_dct;
    check_status(fn(descIn.shape.size(), descIn.shape.data(), descIn.stride.data(), descOut.stride.data(),
                    dct.axes.size(), dct.axes.data(), dct.dctType, IMPULSE_FFT_F64, descIn.buf, descOut.buf,
                    (double)dct.scalingFactor, dct.ortho ? 1 : 0, dct.nthreads));
}

}  // namespace fft
}  // namespace impulse

// The reference's own FFT tests, restated against the C++ host mirror
// (impulse_amd/host/impulse_fft.hpp) of the MI355X engine:
//   tests/test_fft.nim   -- real round trip for every length 1..8191, rel-L2 <= eps
//                           (raw C-path plan, in-place fft on a vector)
//   tests/test_fft2.nim  -- exact known answers for fft([0,1,2,3]) and normalisations
//   README.md:41-98      -- fft([1,2,1,-1,1.5]) and the C++ DataDesc/FFTDesc/apply example
//   cpp_pocketfft/pocketfft.nim:322-337 -- DCTDesc(axes=[0]).apply on [4, 3, 5, 10]
//                           (checked against the defining sum in long double)
// Exit status 0 = all passed.  Usage: test_fft_cpp [maxlen] [eps]
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../impulse_amd/host/impulse_fft.hpp"

using namespace impulse::fft;

static uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void fill_random(std::vector<double> &d, uint64_t &s)
{
    for (auto &v : d) v = -0.5 + (double)(splitmix(s) >> 11) * 0x1.0p-53;
}

static double errcalc(const double *data, const double *odata, size_t n)
{
    double sum = 0, errsum = 0;
    for (size_t m = 0; m < n; ++m) {
        errsum += (data[m] - odata[m]) * (data[m] - odata[m]);
        sum += odata[m] * odata[m];
    }
    return std::sqrt(errsum / sum);
}

static int failures = 0;
#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                \
        }                                                              \
    } while (0)

int main(int argc, char **argv)
{
    const size_t maxlen = argc > 1 ? (size_t)std::atoll(argv[1]) : 8192;
    const double epsilon = argc > 2 ? std::atof(argv[2]) : 2e-15;
    uint64_t seed = 0;

    // test_real (test_fft.nim:29-49): raw plan, forward then backward with 1/length
    {
        std::vector<double> odata(maxlen), data(maxlen);
        fill_random(odata, seed);
        odata[0] = 0.340188;
        double errsum = 0, worst = 0;
        size_t worst_len = 0;
        for (size_t length = 1; length < maxlen; ++length) {
            std::copy(odata.begin(), odata.begin() + length, data.begin());
            impulse_rfft_plan plan = impulse_make_rfft_plan(length);
            CHECK(plan != nullptr);
            CHECK(impulse_rfft_forward(plan, data.data(), 1.0) == 0);
            CHECK(impulse_rfft_backward(plan, data.data(), 1.0 / (double)length) == 0);
            impulse_destroy_rfft_plan(plan);
            const double err = errcalc(data.data(), odata.data(), length);
            if (err > worst) { worst = err; worst_len = length; }
            if (err > epsilon) {
                std::printf("problem at real length %zu : %g\n", length, err);
                ++failures;
            }
            errsum += err;
        }
        std::printf("test_real errsum: %g (worst %g at %zu)\n", errsum, worst, worst_len);
    }
    // test_real_hl_seq (test_fft.nim:70-89): in-place fft on a vector
    {
        double worst = 0;
        for (size_t length = 1; length < maxlen; length += 37) {
            std::vector<double> odata(length);
            fill_random(odata, seed);
            std::vector<double> data = odata;
            fft_inplace(data, true);
            fft_inplace(data, false);
            const double err = errcalc(data.data(), odata.data(), length);
            worst = std::max(worst, err);
            if (err > epsilon) {
                std::printf("problem at real length %zu : %g\n", length, err);
                ++failures;
            }
        }
        std::printf("test_real_hl_seq worst: %g\n", worst);
    }
    // test_fft2.nim:7-15 -- exact
    {
        const std::vector<Complex64> expected = {{6, 0}, {-2, 2}, {-2, 0}, {-2, -2}};
        const std::vector<double> x = {0, 1, 2, 3};
        const std::vector<Complex64> xc(x.begin(), x.end());
        CHECK(fft(xc) == expected);
        CHECK(fft(x) == expected);
        CHECK(fft(x, true) == expected);
        auto back = ifft(fft(x));
        for (size_t i = 0; i < 4; ++i) {
            CHECK(back[i].real() == x[i]);
            CHECK(back[i].imag() == 0.0);
        }
        auto fwd = fft(x, true, nkForward);
        auto ort = fft(x, true, nkOrtho);
        for (size_t i = 0; i < 4; ++i) {
            CHECK(fwd[i] == expected[i] / Complex64(4.0));
            CHECK(ort[i] == expected[i] / Complex64(2.0));
        }
        CHECK(ifft(x) == fft(x, false));
    }
    // README.md:41-98
    {
        const std::vector<double> dIn = {1.0, 2.0, 1.0, -1.0, 1.5};
        auto y = fft(dIn);
        const Complex64 e1(2.081559480312316, -1.651098762732523), e2(-1.831559480312316, 1.608220406444071);
        CHECK(std::abs(y[0] - 4.5) < 1e-15 && std::abs(y[1] - e1) < 1e-15 && std::abs(y[2] - e2) < 1e-15);
        CHECK(std::abs(y[3] - std::conj(e2)) < 1e-15 && std::abs(y[4] - std::conj(e1)) < 1e-15);
        std::vector<Complex64> dOut(dIn.size());
        auto dInDesc = DataDesc<double>::init(const_cast<double *>(dIn.data()), std::vector<size_t>{dIn.size()});
        auto dOutDesc = DataDesc<Complex64>::init(dOut.data(), std::vector<size_t>{dOut.size()});
        auto f = FFTDesc<double>::init({0}, true);
        apply(f, dOutDesc, dInDesc);
        CHECK(std::abs(dOut[0] - 4.5) < 1e-15 && std::abs(dOut[1] - e1) < 1e-15 && std::abs(dOut[2] - e2) < 1e-15);
        CHECK(dOut[3] == Complex64(0, 0) && dOut[4] == Complex64(0, 0));
        // inverse through c2r with the README's full-length complex buffer convention
        std::vector<double> back(dIn.size());
        auto backDesc = DataDesc<double>::init(back.data(), std::vector<size_t>{back.size()});
        auto inv = FFTDesc<double>::init({0}, false, 1.0 / 5.0);
        apply(inv, backDesc, dOutDesc);
        CHECK(errcalc(back.data(), dIn.data(), 5) < 1e-15);
    }
    // the same README example through FFTDesc[float32] / DataDesc[float32] (T = float):
    // float inputs, complex<float> outputs, values rounded from the FP64 result
    {
        const std::vector<float> dIn = {1.0f, 2.0f, 1.0f, -1.0f, 1.5f};
        std::vector<std::complex<float>> dOut(dIn.size());
        auto dInDesc = DataDesc<float>::init(const_cast<float *>(dIn.data()), std::vector<size_t>{dIn.size()});
        auto dOutDesc = DataDesc<std::complex<float>>::init(dOut.data(), std::vector<size_t>{dOut.size()});
        apply(FFTDesc<float>::init({0}, true), dOutDesc, dInDesc);
        const std::complex<float> e1((float)2.081559480312316, (float)-1.651098762732523),
            e2((float)-1.831559480312316, (float)1.608220406444071);
        CHECK(dOut[0] == std::complex<float>(4.5f, 0.0f) && dOut[1] == e1 && dOut[2] == e2);
        CHECK(dOut[3] == std::complex<float>(0, 0) && dOut[4] == std::complex<float>(0, 0));
        std::vector<float> back(dIn.size());
        auto backDesc = DataDesc<float>::init(back.data(), std::vector<size_t>{back.size()});
        apply(FFTDesc<float>::init({0}, false, 1.0f / 5.0f), backDesc, dOutDesc);
        for (size_t i = 0; i < 5; ++i) CHECK(std::fabs(back[i] - dIn[i]) <= 4e-7f * 2.0f);
        // c2c of fft([0,1,2,3]) in complex<float>: exact
        std::vector<std::complex<float>> a = {{0, 0}, {1, 0}, {2, 0}, {3, 0}}, b(4);
        auto da = DataDesc<std::complex<float>>::init(a.data(), std::vector<size_t>{4});
        auto db = DataDesc<std::complex<float>>::init(b.data(), std::vector<size_t>{4});
        apply(FFTDesc<float>::init({0}, true), db, da);
        CHECK(b[0] == std::complex<float>(6, 0) && b[1] == std::complex<float>(-2, 2));
        CHECK(b[2] == std::complex<float>(-2, 0) && b[3] == std::complex<float>(-2, -2));
    }
    // cpp_pocketfft/pocketfft.nim:322-337: default DCT (type 2) of [4, 3, 5, 10]
    {
        const std::vector<double> dIn = {4.0, 3.0, 5.0, 10.0};
        std::vector<double> dOut(dIn.size());
        auto dInDesc = DataDesc<double>::init(const_cast<double *>(dIn.data()), std::vector<size_t>{dIn.size()});
        auto dOutDesc = DataDesc<double>::init(dOut.data(), std::vector<size_t>{dOut.size()});
        apply(DCTDesc<double>::init({0}), dOutDesc, dInDesc);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (size_t k = 0; k < 4; ++k) {
            long double want = 0;
            for (size_t n = 0; n < 4; ++n) want += 2 * dIn[n] * std::cos(pi * k * (2 * n + 1) / 8.0L);
            CHECK(std::fabs(dOut[k] - (double)want) < 1e-13);
        }
        CHECK(dOut[0] == 44.0);
        // DST-II of the same values, then DST-III back: 2N x
        std::vector<double> s(4), back(4);
        auto sDesc = DataDesc<double>::init(s.data(), std::vector<size_t>{4});
        auto bDesc = DataDesc<double>::init(back.data(), std::vector<size_t>{4});
        apply(DCTDesc<double>::init({0}, 2, false, 1.0, 1, true), sDesc, dInDesc);
        apply(DCTDesc<double>::init({0}, 3, false, 1.0 / 8.0, 1, true), bDesc, sDesc);
        CHECK(errcalc(back.data(), dIn.data(), 4) < 1e-15);
    }
    // argument errors surface as exceptions (pocketfft_hdronly.h:446-477)
    {
        std::vector<Complex64> a(4), b(4);
        auto da = DataDesc<Complex64>::init(a.data(), std::vector<size_t>{4});
        auto db = DataDesc<Complex64>::init(b.data(), std::vector<size_t>{4});
        bool threw = false;
        try {
            apply(FFTDesc<double>::init({1}, true), db, da);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        CHECK(threw);
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}

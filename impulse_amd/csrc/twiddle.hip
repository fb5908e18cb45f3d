// Host FP64 twiddle generation.
//
// exp(-2 pi i k / n) is evaluated in long double after an exact integer
// reduction of the angle to the first octant, then rounded to double.  Angles
// that are multiples of pi/4 therefore come out exactly (0, +-1, +-sqrt(1/2)),
// which the reference's exact known-answer tests rely on
// (tests/test_fft2.nim:7-15); PocketFFT obtains the same property by its own
// octant reduction (impulse/fft/cpp_pocketfft/pocketfft_hdronly.h:277-349).
#include <cmath>

#include "engine.h"

namespace ifft {

static const long double kPiL = 3.141592653589793238462643383279502884L;

double2 twiddle_at(uint64_t k, uint64_t n)
{
    k %= n;
    const unsigned __int128 x = (unsigned __int128)8u * k;
    const uint64_t oct = (uint64_t)(x / n);
    const uint64_t r = (uint64_t)(x - (unsigned __int128)oct * n);
    long double c, s;
    if ((oct & 1u) == 0) {
        const long double a = (kPiL / 4.0L) * ((long double)r / (long double)n);
        c = cosl(a);
        s = sinl(a);
    } else if (r == 0) {
        c = s = sqrtl(0.5L);
    } else {
        const long double a = (kPiL / 4.0L) * ((long double)(n - r) / (long double)n);
        c = cosl(a);
        s = sinl(a);
    }
    long double re, im;  // cos / sin of 2 pi k / n
    switch (oct) {
        case 0: re = c; im = s; break;
        case 1: re = s; im = c; break;
        case 2: re = -s; im = c; break;
        case 3: re = -c; im = s; break;
        case 4: re = -c; im = -s; break;
        case 5: re = -s; im = -c; break;
        case 6: re = s; im = -c; break;
        default: re = c; im = -s; break;
    }
    return make_double2((double)re, (double)-im);
}

// out[j] = W_n^{j * step}, j < count
void twiddle_table(std::vector<double2> &out, uint64_t n, uint64_t count, uint64_t step)
{
    out.resize(count);
    for (uint64_t j = 0; j < count; ++j)
        out[j] = twiddle_at((uint64_t)(((unsigned __int128)j * step) % n), n);
}

}  // namespace ifft

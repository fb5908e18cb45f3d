/* Host-only driver for the AddressSanitizer / UndefinedBehaviorSanitizer build
 * of the oracle (oracle/fft_oracle.c).  It walks every entry point over the
 * index arithmetic the checks depend on -- prime, smooth and Bluestein lengths,
 * n-D views with negative and padded byte strides, in-place calls, the C path's
 * halfcomplex order, r2r and upfirdn -- into buffers sized exactly to what each
 * call may touch, so that any read or write outside them stops the run.  The
 * numbers are checked loosely (round trips, naive DFT at small n) so that a
 * sanitizer-clean run is also a working one.  Exit code 0: clean. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fft_oracle.h"

static uint64_t rng_state = 0x9e3779b97f4a7c15ULL;
static double urand(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) / 9007199254740992.0 * 2.0 - 1.0;
}

static int failures = 0;
static void expect(int cond, const char *what, size_t n)
{
    if (!cond) {
        fprintf(stderr, "FAIL %s n=%zu\n", what, n);
        ++failures;
    }
}

static double rel(const double *a, const double *b, size_t m)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < m; ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return sqrt(num / (den > 0 ? den : 1));
}

static void one_d(size_t n)
{
    /* complex: forward vs naive DFT (small n), forward + backward round trip */
    double *x = malloc(2 * n * sizeof(double)), *y = malloc(2 * n * sizeof(double)),
           *z = malloc(2 * n * sizeof(double));
    for (size_t i = 0; i < 2 * n; ++i) x[i] = urand();
    memcpy(y, x, 2 * n * sizeof(double));
    expect(oracle_cfft_forward(n, y, 1.0) == 0, "cfft_forward", n);
    if (n <= 300) {
        expect(oracle_dft_naive(n, x, z, 1) == 0, "dft_naive", n);
        expect(rel(y, z, 2 * n) < 1e-14, "cfft vs naive", n);
    }
    expect(oracle_cfft_backward(n, y, 1.0 / (double)n) == 0, "cfft_backward", n);
    expect(rel(y, x, 2 * n) < 1e-14, "cfft round trip", n);
    /* real, C path halfcomplex */
    if (n > 1) {
        memcpy(y, x, n * sizeof(double));
        expect(oracle_rfft_forward(n, y, 1.0) == 0, "rfft_forward", n);
        expect(oracle_rfft_backward(n, y, 1.0 / (double)n) == 0, "rfft_backward", n);
        expect(rel(y, x, n) < 1e-14, "rfft round trip", n);
    }
    /* r2c then c2r, unit strides, exact-size buffers */
    {
        const size_t nb = n / 2 + 1;
        double *bins = malloc(2 * nb * sizeof(double)), *back = malloc(n * sizeof(double));
        const size_t shape[1] = {n}, axes[1] = {0};
        const ptrdiff_t sr[1] = {8}, sc[1] = {16};
        expect(oracle_r2c(1, shape, sr, sc, 1, axes, 1, x, bins, 1.0) == 0, "r2c", n);
        expect(oracle_c2r(1, shape, sc, sr, 1, axes, 0, bins, back, 1.0 / (double)n) == 0, "c2r", n);
        expect(rel(back, x, n) < 1e-14, "r2c/c2r round trip", n);
        free(bins);
        free(back);
    }
    free(x);
    free(y);
    free(z);
}

/* a [a, b, c] complex array viewed with the middle axis reversed and the last
 * axis every other element: negative and padded strides into one exact buffer */
static void nd_views(void)
{
    const size_t a = 5, b = 7, c = 12;
    const size_t elems = a * b * c;
    double *buf = malloc(2 * elems * sizeof(double)), *out = malloc(2 * a * b * (c / 2) * sizeof(double)),
           *back = malloc(2 * a * b * (c / 2) * sizeof(double));
    for (size_t i = 0; i < 2 * elems; ++i) buf[i] = urand();
    const size_t shape[3] = {a, b, c / 2};
    /* view start: element (0, b-1, 0); strides in bytes */
    const double *start = buf + 2 * ((b - 1) * c);
    const ptrdiff_t sin[3] = {(ptrdiff_t)(b * c * 16), -(ptrdiff_t)(c * 16), 32};
    const ptrdiff_t sout[3] = {(ptrdiff_t)(b * (c / 2) * 16), (ptrdiff_t)((c / 2) * 16), 16};
    const size_t axes[3] = {2, 0, 1};
    expect(oracle_c2c(3, shape, sin, sout, 3, axes, 1, start, out, 1.0) == 0, "c2c 3-D view", elems);
    expect(oracle_c2c(3, shape, sout, sout, 3, axes, 0, out, back, 1.0 / (double)(a * b * (c / 2))) == 0,
           "c2c 3-D back", elems);
    for (size_t i = 0; i < a; ++i)
        for (size_t j = 0; j < b; ++j)
            for (size_t k = 0; k < c / 2; ++k) {
                const double *s = buf + 2 * (i * b * c + (b - 1 - j) * c + 2 * k);
                const double *t = back + 2 * ((i * b + j) * (c / 2) + k);
                if (fabs(s[0] - t[0]) > 1e-13 || fabs(s[1] - t[1]) > 1e-13) {
                    expect(0, "c2c 3-D view round trip", elems);
                    i = a;
                    j = b;
                    break;
                }
            }
    /* in place, equal strides */
    expect(oracle_c2c(3, shape, sout, sout, 2, axes, 1, back, back, 1.0) == 0, "c2c in place", elems);
    /* r2c over a reversed real axis and c2r back into a padded real layout */
    {
        const size_t rs[2] = {6, 9}, ax[2] = {0, 1};   /* the real axis is the last one listed */
        double *r = malloc(6 * 9 * sizeof(double)), *cb = malloc(2 * 6 * 5 * sizeof(double)),
               *rb = malloc(6 * 11 * sizeof(double));
        for (size_t i = 0; i < 54; ++i) r[i] = urand();
        const ptrdiff_t srr[2] = {-(ptrdiff_t)(9 * 8), 8}, scc[2] = {5 * 16, 16}, srb[2] = {11 * 8, 8};
        expect(oracle_r2c(2, rs, srr, scc, 2, ax, 1, r + 5 * 9, cb, 1.0) == 0, "r2c reversed", 54);
        expect(oracle_c2r(2, rs, scc, srb, 2, ax, 0, cb, rb, 1.0 / 54.0) == 0, "c2r padded", 54);
        for (size_t i = 0; i < 6; ++i)
            for (size_t j = 0; j < 9; ++j)
                expect(fabs(rb[i * 11 + j] - r[(5 - i) * 9 + j]) < 1e-13, "r2c/c2r 2-D round trip", 54);
        free(r);
        free(cb);
        free(rb);
    }
    free(buf);
    free(out);
    free(back);
}

static void r2r_and_signal(void)
{
    const size_t lens[] = {1, 2, 3, 5, 8, 17, 64, 97};
    for (size_t q = 0; q < sizeof(lens) / sizeof(lens[0]); ++q) {
        const size_t n = lens[q];
        double *x = malloc(n * sizeof(double)), *y = malloc(n * sizeof(double));
        for (size_t i = 0; i < n; ++i) x[i] = urand();
        const size_t shape[1] = {n}, axes[1] = {0};
        const ptrdiff_t s[1] = {8};
        for (int t = 1; t <= 4; ++t) {
            if (!(t == 1 && n == 1)) expect(oracle_dct(1, shape, s, s, 1, axes, t, x, y, 1.0, t == 2) == 0, "dct", n);
            expect(oracle_dst(1, shape, s, s, 1, axes, t, x, y, 1.0, t == 3) == 0, "dst", n);
        }
        for (int r2h = 0; r2h < 2; ++r2h)
            for (int fw = 0; fw < 2; ++fw)
                expect(oracle_r2r_fftpack(1, shape, s, s, 1, axes, r2h, fw, x, y, 1.0) == 0, "r2r_fftpack", n);
        expect(oracle_r2r_separable_hartley(1, shape, s, s, 1, axes, x, y, 1.0) == 0, "hartley", n);
        expect(oracle_r2r_genuine_hartley(1, shape, s, s, 1, axes, x, y, 1.0) == 0, "genuine hartley", n);
        free(x);
        free(y);
    }
    /* upfirdn: output sized by oracle_upfirdn_length exactly */
    const int rates[][2] = {{1, 1}, {3, 1}, {1, 3}, {3, 2}, {7, 5}};
    for (size_t q = 0; q < sizeof(rates) / sizeof(rates[0]); ++q)
        for (size_t nt = 1; nt < 40; nt += 13)
            for (size_t nh = 1; nh < 30; nh += 7) {
                double *t = malloc(nt * sizeof(double)), *h = malloc(nh * sizeof(double));
                for (size_t i = 0; i < nt; ++i) t[i] = urand();
                for (size_t i = 0; i < nh; ++i) h[i] = urand();
                const size_t ny = oracle_upfirdn_length(nt, nh, rates[q][0], rates[q][1]);
                double *yy = malloc((ny ? ny : 1) * sizeof(double));
                expect(oracle_upfirdn(nt, t, nh, h, rates[q][0], rates[q][1], yy) == 0, "upfirdn", nt);
                free(t);
                free(h);
                free(yy);
            }
}

static void errors(void)
{
    double d[4] = {0, 0, 0, 0};
    const size_t shape[1] = {2}, bad_axes[1] = {3};
    const ptrdiff_t s[1] = {16};
    expect(oracle_c2c(1, shape, s, s, 1, bad_axes, 1, d, d, 1.0) != 0, "bad axis rejected", 0);
    expect(oracle_c2c(0, shape, s, s, 0, bad_axes, 1, d, d, 1.0) != 0, "ndim 0 rejected", 0);
}

int main(void)
{
    /* smooth, prime (recursive naive butterflies up to 64, Bluestein above), and mixed lengths */
    const size_t lens[] = {1, 2, 3, 4, 5, 6, 7, 8, 11, 13, 16, 17, 31, 32, 60, 64, 67, 97, 100, 127, 128,
                           243, 255, 256, 1009, 2310, 4095, 4096, 10007};
    for (size_t q = 0; q < sizeof(lens) / sizeof(lens[0]); ++q) one_d(lens[q]);
    nd_views();
    r2r_and_signal();
    errors();
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("oracle sanitizer run clean\n");
    return 0;
}

/* CPU oracle for the impulse FFT path -- TEST INFRASTRUCTURE ONLY (see fft_oracle.h).
 *
 * Everything is computed in long double (64-bit mantissa on x86-64), so the
 * oracle's own rounding error (~1e-18 relative) is far below the 1e-12 parity
 * bar it is used to enforce.
 */
#include "fft_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { long double r, i; } lc;

static const long double PI_L = 3.141592653589793238462643383279502884L;

static inline lc lc_mul(lc a, lc b) { lc z = {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; return z; }
static inline lc lc_add(lc a, lc b) { lc z = {a.r + b.r, a.i + b.i}; return z; }

/* exp(sign * 2 pi i k / n), octant-reduced so that multiples of pi/4 are exact. */
static lc twiddle_ld(uint64_t k, uint64_t n, int sign)
{
    k %= n;
    /* angle = (pi/4) * (8k / n); octant o = floor(8k/n), remainder r */
    unsigned __int128 x = (unsigned __int128)8u * k;
    uint64_t o = (uint64_t)(x / n);
    uint64_t r = (uint64_t)(x - (unsigned __int128)o * n);
    long double c, s;
    if ((o & 1u) == 0) {
        long double phi = (PI_L / 4.0L) * ((long double)r / (long double)n);
        c = cosl(phi); s = sinl(phi);
    } else {
        uint64_t rr = n - r;  /* complement within the octant */
        if (rr == n) {        /* exactly on an odd multiple of pi/4 */
            c = s = sqrtl(0.5L);
        } else {
            long double phi = (PI_L / 4.0L) * ((long double)rr / (long double)n);
            c = cosl(phi); s = sinl(phi);
        }
    }
    long double cr, ci;
    switch (o) {
        case 0: cr = c;  ci = s;  break;
        case 1: cr = s;  ci = c;  break;
        case 2: cr = -s; ci = c;  break;
        case 3: cr = -c; ci = s;  break;
        case 4: cr = -c; ci = -s; break;
        case 5: cr = -s; ci = -c; break;
        case 6: cr = s;  ci = -c; break;
        default: cr = c; ci = -s; break;
    }
    lc z = {cr, sign < 0 ? -ci : ci};
    return z;
}

void oracle_twiddle(size_t k, size_t n, int sign, double *re, double *im)
{
    lc z = twiddle_ld(k, n, sign);
    *re = (double)z.r;
    *im = (double)z.i;
}

static size_t smallest_prime_factor(size_t n)
{
    if (n % 2 == 0) return 2;
    for (size_t p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* -------------------------------------------------------------------------- */
/* 1-D complex DFT in long double                                              */
/* -------------------------------------------------------------------------- */

static int fft_ld(size_t n, lc *data, int sign);

/* Recursive decimation in time over the smallest prime factor.  `tw` holds
 * exp(sign 2 pi i j / N) for the top-level N; a sub-problem of size n uses
 * stride N/n into it. */
static int rec(size_t n, const lc *in, size_t is, lc *out, const lc *tw, size_t twstep,
               int sign, lc *tmp)
{
    if (n == 1) { out[0] = in[0]; return 0; }
    size_t p = smallest_prime_factor(n);
    size_t m = n / p;
    if (p == n && n > 64) {
        /* large prime: Bluestein on a private copy */
        lc *buf = (lc *)malloc(n * sizeof(lc));
        if (!buf) return 2;
        for (size_t j = 0; j < n; ++j) buf[j] = in[j * is];
        int rc = fft_ld(n, buf, sign);
        memcpy(out, buf, n * sizeof(lc));
        free(buf);
        return rc;
    }
    for (size_t q = 0; q < p; ++q) {
        int rc = rec(m, in + q * is, is * p, out + q * m, tw, twstep * p, sign, tmp + p);
        if (rc) return rc;
    }
    /* combine: X[k + s m] = sum_q W_n^{q (k + s m)} Y_q[k] */
    for (size_t k = 0; k < m; ++k) {
        for (size_t q = 0; q < p; ++q) tmp[q] = out[q * m + k];
        for (size_t s = 0; s < p; ++s) {
            lc acc = tmp[0];
            size_t idx_step = (k + s * m) % n;
            size_t idx = 0;
            for (size_t q = 1; q < p; ++q) {
                idx += idx_step;
                if (idx >= n) idx -= n;
                acc = lc_add(acc, lc_mul(tmp[q], tw[idx * twstep]));
            }
            out[k + s * m] = acc;
        }
    }
    return 0;
}

static int bluestein_ld(size_t n, lc *data, int sign)
{
    size_t n2 = 1;
    while (n2 < 2 * n - 1) n2 <<= 1;
    lc *w = (lc *)malloc(n * sizeof(lc));
    lc *a = (lc *)calloc(n2, sizeof(lc));
    lc *b = (lc *)calloc(n2, sizeof(lc));
    if (!w || !a || !b) { free(w); free(a); free(b); return 2; }
    /* w[j] = exp(sign i pi j^2 / n) = W_{2n}^{j^2 mod 2n} */
    for (size_t j = 0; j < n; ++j) {
        uint64_t idx = (uint64_t)(((unsigned __int128)j * j) % (2 * (unsigned __int128)n));
        w[j] = twiddle_ld(idx, 2 * n, sign);
    }
    for (size_t j = 0; j < n; ++j) {
        a[j] = lc_mul(data[j], w[j]);
        lc cw = {w[j].r, -w[j].i};
        b[j] = cw;
        if (j) b[n2 - j] = cw;
    }
    int rc = fft_ld(n2, a, -1);
    if (!rc) rc = fft_ld(n2, b, -1);
    if (!rc) {
        for (size_t j = 0; j < n2; ++j) a[j] = lc_mul(a[j], b[j]);
        rc = fft_ld(n2, a, +1);
    }
    if (!rc) {
        long double inv = 1.0L / (long double)n2;
        for (size_t k = 0; k < n; ++k) {
            lc z = lc_mul(a[k], w[k]);
            data[k].r = z.r * inv;
            data[k].i = z.i * inv;
        }
    }
    free(w); free(a); free(b);
    return rc;
}

/* In-place DFT of `data` with exp(sign 2 pi i jk/n). */
static int fft_ld(size_t n, lc *data, int sign)
{
    if (n <= 1) return 0;
    size_t p = smallest_prime_factor(n);
    if (p == n && n > 64) return bluestein_ld(n, data, sign);
    lc *tw = (lc *)malloc(n * sizeof(lc));
    lc *in = (lc *)malloc(n * sizeof(lc));
    /* tmp: one slot of size p per recursion level; sum of factors <= n + 64 */
    size_t tmpn = 0;
    for (size_t m = n; m > 1;) { size_t q = smallest_prime_factor(m); tmpn += q; m /= q; }
    lc *tmp = (lc *)malloc((tmpn + 1) * sizeof(lc));
    if (!tw || !in || !tmp) { free(tw); free(in); free(tmp); return 2; }
    for (size_t j = 0; j < n; ++j) tw[j] = twiddle_ld(j, n, sign);
    memcpy(in, data, n * sizeof(lc));
    int rc = rec(n, in, 1, data, tw, 1, sign, tmp);
    free(tw); free(in); free(tmp);
    return rc;
}

int oracle_dft_naive(size_t n, const double *in, double *out, int forward)
{
    if (n == 0) return 1;
    int sign = forward ? -1 : 1;
    lc *res = (lc *)malloc(n * sizeof(lc));
    if (!res) return 2;
    for (size_t k = 0; k < n; ++k) {
        lc acc = {0.0L, 0.0L};
        for (size_t j = 0; j < n; ++j) {
            uint64_t idx = (uint64_t)(((unsigned __int128)j * k) % n);
            lc x = {in[2 * j], in[2 * j + 1]};
            acc = lc_add(acc, lc_mul(x, twiddle_ld(idx, n, sign)));
        }
        res[k] = acc;
    }
    for (size_t k = 0; k < n; ++k) { out[2 * k] = (double)res[k].r; out[2 * k + 1] = (double)res[k].i; }
    free(res);
    return 0;
}

/* -------------------------------------------------------------------------- */
/* 1-D C-path entry points                                                     */
/* -------------------------------------------------------------------------- */

/* The C path returns before scaling when n == 1, so fct is ignored there
 * (impulse/fft/c_pocketfft/pocketfft.c:874,1703,1739); the C++ path scales
 * (pocketfft_hdronly.h:1309,2114). */
static int cfft_1d(size_t n, double *c, double fct, int sign)
{
    if (n == 0) return 1;
    if (n == 1) return 0;
    lc *buf = (lc *)malloc(n * sizeof(lc));
    if (!buf) return 2;
    for (size_t j = 0; j < n; ++j) { buf[j].r = c[2 * j]; buf[j].i = c[2 * j + 1]; }
    int rc = fft_ld(n, buf, sign);
    if (!rc)
        for (size_t j = 0; j < n; ++j) {
            c[2 * j] = (double)(buf[j].r * fct);
            c[2 * j + 1] = (double)(buf[j].i * fct);
        }
    free(buf);
    return rc;
}

int oracle_cfft_forward(size_t n, double *c, double fct) { return cfft_1d(n, c, fct, -1); }
int oracle_cfft_backward(size_t n, double *c, double fct) { return cfft_1d(n, c, fct, +1); }

/* Forward real transform of n reals -> n/2+1 complex bins (long double). */
static int r2hc_ld(size_t n, const long double *x, lc *bins)
{
    lc *buf = (lc *)malloc(n * sizeof(lc));
    if (!buf) return 2;
    for (size_t j = 0; j < n; ++j) { buf[j].r = x[j]; buf[j].i = 0.0L; }
    int rc = fft_ld(n, buf, -1);
    if (!rc) for (size_t k = 0; k <= n / 2; ++k) bins[k] = buf[k];
    free(buf);
    return rc;
}

/* Backward real transform: n/2+1 Hermitian bins -> n reals.  Im of bin 0 (and
 * of bin n/2 for even n) is ignored, as FFTPACK's halfcomplex format has no
 * slot for it. */
static int hc2r_ld(size_t n, const lc *bins, long double *x)
{
    lc *buf = (lc *)malloc(n * sizeof(lc));
    if (!buf) return 2;
    buf[0].r = bins[0].r; buf[0].i = 0.0L;
    for (size_t k = 1; k < n; ++k) {
        if (2 * k < n) buf[k] = bins[k];
        else if (2 * k == n) { buf[k].r = bins[k].r; buf[k].i = 0.0L; }
        else { buf[k].r = bins[n - k].r; buf[k].i = -bins[n - k].i; }
    }
    int rc = fft_ld(n, buf, +1);
    if (!rc) for (size_t j = 0; j < n; ++j) x[j] = buf[j].r;
    free(buf);
    return rc;
}

int oracle_rfft_forward(size_t n, double *c, double fct)
{
    if (n == 0) return 1;
    if (n == 1) return 0;   /* pocketfft.c:1703 */
    long double *x = (long double *)malloc(n * sizeof(long double));
    lc *bins = (lc *)malloc((n / 2 + 1) * sizeof(lc));
    if (!x || !bins) { free(x); free(bins); return 2; }
    for (size_t j = 0; j < n; ++j) x[j] = c[j];
    int rc = r2hc_ld(n, x, bins);
    if (!rc) {
        c[0] = (double)(bins[0].r * fct);
        size_t i = 1, k = 1;
        for (; i + 1 < n; i += 2, ++k) {
            c[i] = (double)(bins[k].r * fct);
            c[i + 1] = (double)(bins[k].i * fct);
        }
        if (i < n) c[i] = (double)(bins[k].r * fct);
    }
    free(x); free(bins);
    return rc;
}

int oracle_rfft_backward(size_t n, double *c, double fct)
{
    if (n == 0) return 1;
    if (n == 1) return 0;   /* pocketfft.c:1739 */
    long double *x = (long double *)malloc(n * sizeof(long double));
    lc *bins = (lc *)calloc(n / 2 + 1, sizeof(lc));
    if (!x || !bins) { free(x); free(bins); return 2; }
    bins[0].r = c[0];
    size_t i = 1, k = 1;
    for (; i + 1 < n; i += 2, ++k) { bins[k].r = c[i]; bins[k].i = c[i + 1]; }
    if (i < n) bins[k].r = c[i];
    int rc = hc2r_ld(n, bins, x);
    if (!rc) for (size_t j = 0; j < n; ++j) c[j] = (double)(x[j] * fct);
    free(x); free(bins);
    return rc;
}

/* -------------------------------------------------------------------------- */
/* n-D driver                                                                  */
/* -------------------------------------------------------------------------- */

static int check_args(size_t ndim, const size_t *shape, size_t naxes, const size_t *axes)
{
    if (ndim < 1 || !shape || naxes < 1 || !axes) return 1;
    for (size_t a = 0; a < naxes; ++a) {
        if (axes[a] >= ndim) return 1;
        for (size_t b = 0; b < a; ++b) if (axes[b] == axes[a]) return 1;
    }
    return 0;
}

static size_t prod(size_t ndim, const size_t *shape)
{
    size_t p = 1;
    for (size_t d = 0; d < ndim; ++d) p *= shape[d];
    return p;
}

/* Iterate over every line along `axis`: calls fn with the byte offsets of
 * element 0 of the line in the input and output arrays. */
typedef int (*line_fn)(void *ctx, ptrdiff_t ioff, ptrdiff_t ooff);

static int for_each_line(size_t ndim, const size_t *shape, const ptrdiff_t *sin,
                         const ptrdiff_t *sout, size_t axis, line_fn fn, void *ctx)
{
    size_t idx[64];
    if (ndim > 64) return 1;
    memset(idx, 0, sizeof(idx));
    size_t lines = 1;
    for (size_t d = 0; d < ndim; ++d) if (d != axis) lines *= shape[d];
    for (size_t l = 0; l < lines; ++l) {
        ptrdiff_t io = 0, oo = 0;
        for (size_t d = 0; d < ndim; ++d) { io += (ptrdiff_t)idx[d] * sin[d]; oo += (ptrdiff_t)idx[d] * sout[d]; }
        int rc = fn(ctx, io, oo);
        if (rc) return rc;
        for (size_t d = ndim; d-- > 0;) {
            if (d == axis) continue;
            if (++idx[d] < shape[d]) break;
            idx[d] = 0;
        }
    }
    return 0;
}

typedef struct {
    const char *in; char *out; ptrdiff_t is, os; size_t n; int sign; long double fct;
    lc *buf; long double *rbuf; int forward;
} line_ctx;

static int c2c_line(void *vctx, ptrdiff_t io, ptrdiff_t oo)
{
    line_ctx *c = (line_ctx *)vctx;
    for (size_t j = 0; j < c->n; ++j) {
        const double *p = (const double *)(c->in + io + (ptrdiff_t)j * c->is);
        c->buf[j].r = p[0]; c->buf[j].i = p[1];
    }
    int rc = fft_ld(c->n, c->buf, c->sign);
    if (rc) return rc;
    for (size_t j = 0; j < c->n; ++j) {
        double *p = (double *)(c->out + oo + (ptrdiff_t)j * c->os);
        p[0] = (double)(c->buf[j].r * c->fct); p[1] = (double)(c->buf[j].i * c->fct);
    }
    return 0;
}

/* One c2c pass along `axis`, done through a long-double staging copy of the
 * whole array so that in == out aliasing is harmless. */
static int c2c_axis(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout,
                    size_t axis, int forward, const char *in, char *out, long double fct)
{
    line_ctx c;
    c.in = in; c.out = out; c.is = sin[axis]; c.os = sout[axis]; c.n = shape[axis];
    c.sign = forward ? -1 : 1; c.fct = fct; c.forward = forward;
    c.buf = (lc *)malloc(c.n * sizeof(lc)); c.rbuf = NULL;
    if (!c.buf) return 2;
    int rc = for_each_line(ndim, shape, sin, sout, axis, c2c_line, &c);
    free(c.buf);
    return rc;
}

int oracle_c2c(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
               const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
               int forward, const double *data_in, double *data_out, double fct)
{
    if (check_args(ndim, shape, naxes, axes)) return 1;
    if (prod(ndim, shape) == 0) return 0;
    if (data_in == data_out)
        for (size_t d = 0; d < ndim; ++d) if (stride_in[d] != stride_out[d]) return 1;
    /* first axis reads `in`, later axes work in place on `out`; fct once */
    int rc = c2c_axis(ndim, shape, stride_in, stride_out, axes[0], forward,
                      (const char *)data_in, (char *)data_out, fct);
    for (size_t a = 1; a < naxes && !rc; ++a)
        rc = c2c_axis(ndim, shape, stride_out, stride_out, axes[a], forward,
                      (const char *)data_out, (char *)data_out, 1.0L);
    return rc;
}

static int r2c_line(void *vctx, ptrdiff_t io, ptrdiff_t oo)
{
    line_ctx *c = (line_ctx *)vctx;
    for (size_t j = 0; j < c->n; ++j) c->rbuf[j] = *(const double *)(c->in + io + (ptrdiff_t)j * c->is);
    int rc = r2hc_ld(c->n, c->rbuf, c->buf);
    if (rc) return rc;
    for (size_t k = 0; k <= c->n / 2; ++k) {
        double *p = (double *)(c->out + oo + (ptrdiff_t)k * c->os);
        p[0] = (double)(c->buf[k].r * c->fct);
        p[1] = (double)((c->forward ? c->buf[k].i : -c->buf[k].i) * c->fct);
    }
    return 0;
}

static int c2r_line(void *vctx, ptrdiff_t io, ptrdiff_t oo)
{
    line_ctx *c = (line_ctx *)vctx;
    for (size_t k = 0; k <= c->n / 2; ++k) {
        const double *p = (const double *)(c->in + io + (ptrdiff_t)k * c->is);
        c->buf[k].r = p[0];
        c->buf[k].i = c->forward ? -p[1] : p[1];
    }
    int rc = hc2r_ld(c->n, c->buf, c->rbuf);
    if (rc) return rc;
    for (size_t j = 0; j < c->n; ++j)
        *(double *)(c->out + oo + (ptrdiff_t)j * c->os) = (double)(c->rbuf[j] * c->fct);
    return 0;
}

int oracle_r2c(size_t ndim, const size_t *shape_in, const ptrdiff_t *stride_in,
               const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
               int forward, const double *data_in, double *data_out, double fct)
{
    if (check_args(ndim, shape_in, naxes, axes)) return 1;
    if (prod(ndim, shape_in) == 0) return 0;
    size_t axis = axes[naxes - 1];
    size_t shape_out[64];
    if (ndim > 64) return 1;
    memcpy(shape_out, shape_in, ndim * sizeof(size_t));
    shape_out[axis] = shape_in[axis] / 2 + 1;
    line_ctx c;
    c.in = (const char *)data_in; c.out = (char *)data_out; c.is = stride_in[axis]; c.os = stride_out[axis];
    c.n = shape_in[axis]; c.sign = -1; c.fct = fct; c.forward = forward;
    c.buf = (lc *)malloc((c.n / 2 + 1) * sizeof(lc));
    c.rbuf = (long double *)malloc(c.n * sizeof(long double));
    if (!c.buf || !c.rbuf) { free(c.buf); free(c.rbuf); return 2; }
    /* iterate lines over the input shape; output offsets use the same indices */
    int rc = for_each_line(ndim, shape_in, stride_in, stride_out, axis, r2c_line, &c);
    free(c.buf); free(c.rbuf);
    for (size_t a = 0; a + 1 < naxes && !rc; ++a)
        rc = c2c_axis(ndim, shape_out, stride_out, stride_out, axes[a], forward,
                      (const char *)data_out, (char *)data_out, 1.0L);
    return rc;
}

int oracle_c2r(size_t ndim, const size_t *shape_out, const ptrdiff_t *stride_in,
               const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
               int forward, const double *data_in, double *data_out, double fct)
{
    if (check_args(ndim, shape_out, naxes, axes)) return 1;
    if (prod(ndim, shape_out) == 0) return 0;
    if (ndim > 64) return 1;
    size_t axis = axes[naxes - 1];
    size_t shape_in[64];
    memcpy(shape_in, shape_out, ndim * sizeof(size_t));
    shape_in[axis] = shape_out[axis] / 2 + 1;
    const char *src = (const char *)data_in;
    const ptrdiff_t *src_str = stride_in;
    double *tmp = NULL;
    ptrdiff_t tstr[64];
    if (naxes > 1) {
        /* c2c over the other axes into a C-contiguous temporary, fct = 1 */
        size_t nval = prod(ndim, shape_in);
        tmp = (double *)malloc(nval * 2 * sizeof(double));
        if (!tmp) return 2;
        tstr[ndim - 1] = 16;
        for (size_t d = ndim - 1; d-- > 0;) tstr[d] = tstr[d + 1] * (ptrdiff_t)shape_in[d + 1];
        int rc = c2c_axis(ndim, shape_in, stride_in, tstr, axes[0], forward,
                          (const char *)data_in, (char *)tmp, 1.0L);
        for (size_t a = 1; a + 1 < naxes && !rc; ++a)
            rc = c2c_axis(ndim, shape_in, tstr, tstr, axes[a], forward, (const char *)tmp, (char *)tmp, 1.0L);
        if (rc) { free(tmp); return rc; }
        src = (const char *)tmp;
        src_str = tstr;
    }
    line_ctx c;
    c.in = src; c.out = (char *)data_out; c.is = src_str[axis]; c.os = stride_out[axis];
    c.n = shape_out[axis]; c.sign = 1; c.fct = fct; c.forward = forward;
    c.buf = (lc *)malloc((c.n / 2 + 1) * sizeof(lc));
    c.rbuf = (long double *)malloc(c.n * sizeof(long double));
    int rc = 2;
    if (c.buf && c.rbuf)
        rc = for_each_line(ndim, shape_out, src_str, stride_out, axis, c2r_line, &c);
    free(c.buf); free(c.rbuf); free(tmp);
    return rc;
}


/* ---- DCT / DST (naive sums, long double) ------------------------------- */

typedef struct {
    const char *in; char *out; ptrdiff_t is, os; size_t n; int type, cosine, ortho; long double fct;
    long double *x, *y;
} r2r_ctx;

/* cos / sin (pi a / b) from the octant-exact twiddle: exp(2 pi i a / 2b) */
static long double cospi_q(uint64_t a, uint64_t b) { return twiddle_ld(a % (2 * b), 2 * b, 1).r; }
static long double sinpi_q(uint64_t a, uint64_t b) { return twiddle_ld(a % (2 * b), 2 * b, 1).i; }

static int r2r_line(void *vctx, ptrdiff_t io, ptrdiff_t oo)
{
    r2r_ctx *c = (r2r_ctx *)vctx;
    const size_t N = c->n;
    const long double s2 = 1.414213562373095048801688724209698L;
    long double *x = c->x, *y = c->y;
    for (size_t j = 0; j < N; ++j) x[j] = *(const double *)(c->in + io + (ptrdiff_t)j * c->is);
    if (c->cosine) {
        switch (c->type) {
            case 1:
                if (c->ortho) { x[0] *= s2; x[N - 1] *= s2; }
                for (size_t k = 0; k < N; ++k) {
                    long double a = x[0] + ((k & 1) ? -x[N - 1] : x[N - 1]);
                    for (size_t m = 1; m + 1 < N; ++m) a += 2 * x[m] * cospi_q(k * m, N - 1);
                    y[k] = a;
                }
                if (c->ortho) { y[0] *= s2 * 0.5L; y[N - 1] *= s2 * 0.5L; }
                break;
            case 2:
                for (size_t k = 0; k < N; ++k) {
                    long double a = 0;
                    for (size_t m = 0; m < N; ++m) a += x[m] * cospi_q(k * (2 * m + 1), 2 * N);
                    y[k] = 2 * a;
                }
                if (c->ortho) y[0] *= s2 * 0.5L;
                break;
            case 3:
                if (c->ortho) x[0] *= s2;
                for (size_t k = 0; k < N; ++k) {
                    long double a = x[0];
                    for (size_t m = 1; m < N; ++m) a += 2 * x[m] * cospi_q(m * (2 * k + 1), 2 * N);
                    y[k] = a;
                }
                break;
            default:
                for (size_t k = 0; k < N; ++k) {
                    long double a = 0;
                    for (size_t m = 0; m < N; ++m) a += x[m] * cospi_q((2 * k + 1) * (2 * m + 1), 4 * N);
                    y[k] = 2 * a;
                }
        }
    } else {
        switch (c->type) {
            case 1:
                for (size_t k = 0; k < N; ++k) {
                    long double a = 0;
                    for (size_t m = 0; m < N; ++m) a += x[m] * sinpi_q((k + 1) * (m + 1), N + 1);
                    y[k] = 2 * a;
                }
                break;
            case 2:
                for (size_t k = 0; k < N; ++k) {
                    long double a = 0;
                    for (size_t m = 0; m < N; ++m) a += x[m] * sinpi_q((k + 1) * (2 * m + 1), 2 * N);
                    y[k] = 2 * a;
                }
                if (c->ortho) y[0] *= s2 * 0.5L;
                break;
            case 3:
                if (c->ortho) x[0] *= s2;
                for (size_t k = 0; k < N; ++k) {
                    long double a = (k & 1) ? -x[N - 1] : x[N - 1];
                    for (size_t m = 0; m + 1 < N; ++m) a += 2 * x[m] * sinpi_q((2 * k + 1) * (m + 1), 2 * N);
                    y[k] = a;
                }
                break;
            default:
                for (size_t k = 0; k < N; ++k) {
                    long double a = 0;
                    for (size_t m = 0; m < N; ++m) a += x[m] * sinpi_q((2 * k + 1) * (2 * m + 1), 4 * N);
                    y[k] = 2 * a;
                }
        }
    }
    for (size_t k = 0; k < N; ++k) *(double *)(c->out + oo + (ptrdiff_t)k * c->os) = (double)(y[k] * c->fct);
    return 0;
}

static int r2r_nd(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout, size_t naxes,
                  const size_t *axes, int type, int cosine, const double *in, double *out, double fct, int ortho)
{
    if (type < 1 || type > 4) return 1;
    if (check_args(ndim, shape, naxes, axes)) return 1;
    if (prod(ndim, shape) == 0) return 0;
    if (in == out)
        for (size_t d = 0; d < ndim; ++d) if (sin[d] != sout[d]) return 1;
    for (size_t a = 0; a < naxes; ++a)
        if (cosine && type == 1 && shape[axes[a]] < 2) return 1;   /* DCT-I needs N >= 2 */
    int rc = 0;
    for (size_t a = 0; a < naxes && !rc; ++a) {
        r2r_ctx c;
        const size_t ax = axes[a];
        c.in = a ? (const char *)out : (const char *)in;
        c.out = (char *)out;
        c.is = a ? sout[ax] : sin[ax];
        c.os = sout[ax];
        c.n = shape[ax];
        c.type = type; c.cosine = cosine; c.ortho = ortho;
        c.fct = a ? 1.0L : (long double)fct;
        c.x = (long double *)malloc(c.n * sizeof(long double));
        c.y = (long double *)malloc(c.n * sizeof(long double));
        if (!c.x || !c.y) { free(c.x); free(c.y); return 2; }
        rc = for_each_line(ndim, shape, a ? sout : sin, sout, ax, r2r_line, &c);
        free(c.x); free(c.y);
    }
    return rc;
}

int oracle_dct(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
               size_t naxes, const size_t *axes, int type, const double *data_in, double *data_out, double fct,
               int ortho)
{
    return r2r_nd(ndim, shape, stride_in, stride_out, naxes, axes, type, 1, data_in, data_out, fct, ortho);
}

int oracle_dst(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
               size_t naxes, const size_t *axes, int type, const double *data_in, double *data_out, double fct,
               int ortho)
{
    return r2r_nd(ndim, shape, stride_in, stride_out, naxes, axes, type, 0, data_in, data_out, fct, ortho);
}

size_t oracle_upfirdn_length(size_t nt, size_t nh, int up, int down)
{
    if (nt < 1 || nh < 1 || up < 1 || down < 1) return 0;
    const size_t lc = (nt - 1) * (size_t)up + nh;
    return (lc + (size_t)down - 1) / (size_t)down;
}

int oracle_upfirdn(size_t nt, const double *t, size_t nh, const double *h, int up, int down, double *y)
{
    const size_t nout = oracle_upfirdn_length(nt, nh, up, down);
    if (!nout || !t || !h || !y) return 1;
    for (size_t j = 0; j < nout; ++j) {
        const size_t m = j * (size_t)down;          /* index into the full convolution */
        long double acc = 0.0L;
        for (size_t i = 0; i < nh && i <= m; ++i) {  /* c[m] = sum_i u[m - i] h[i] */
            const size_t mu = m - i;
            if (mu % (size_t)up) continue;
            const size_t q = mu / (size_t)up;
            if (q < nt) acc += (long double)t[q] * (long double)h[i];
        }
        y[j] = (double)acc;
    }
    return 0;
}

/* ---- r2r_fftpack and the Hartley transforms (pocketfft_hdronly.h:3252-3270,
 * 3066-3110, 3392-3445), naive long-double sums per line ---- */
typedef struct {
    const char *in; char *out; ptrdiff_t is, os; size_t n; int hartley, r2h, forward; long double fct;
    long double *x, *y;
} rfp_ctx;

static int rfp_line(void *vctx, ptrdiff_t io, ptrdiff_t oo)
{
    rfp_ctx *c = (rfp_ctx *)vctx;
    const size_t N = c->n;
    long double *x = c->x, *y = c->y;
    for (size_t j = 0; j < N; ++j) x[j] = *(const double *)(c->in + io + (ptrdiff_t)j * c->is);
    if (c->hartley) {
        /* H[k] = Re X[k] + Im X[k], X[k] = sum x[n] e^{-2 pi i k n / N} */
        for (size_t k = 0; k < N; ++k) {
            long double acc = 0.0L;
            for (size_t n = 0; n < N; ++n)
                acc += x[n] * (cospi_q(2 * (uint64_t)((k * n) % N), N) - sinpi_q(2 * (uint64_t)((k * n) % N), N));
            y[k] = acc;
        }
    } else {
        /* ExecR2R: the direction picks the operation (forward: real -> halfcomplex, else
         * halfcomplex -> real); real2hermitian only adds the sign flips of the positions
         * 2, 4, ... (the imaginary slots): before a forward call without it, after a
         * backward call with it */
        if (!c->r2h && c->forward)
            for (size_t i = 2; i < N; i += 2) x[i] = -x[i];
        if (c->forward) {
            for (size_t k = 0; 2 * k <= N; ++k) {
                long double re = 0.0L, im = 0.0L;
                for (size_t n = 0; n < N; ++n) {
                    const uint64_t a = 2 * (uint64_t)((k * n) % N);
                    re += x[n] * cospi_q(a, N);
                    im -= x[n] * sinpi_q(a, N);
                }
                if (k == 0) y[0] = re;
                else if (2 * k == N) y[N - 1] = re;
                else { y[2 * k - 1] = re; y[2 * k] = im; }
            }
        } else {
            for (size_t n = 0; n < N; ++n) {
                long double acc = x[0];
                for (size_t k = 1; 2 * k < N; ++k) {
                    const uint64_t a = 2 * (uint64_t)((k * n) % N);
                    acc += 2.0L * (x[2 * k - 1] * cospi_q(a, N) - x[2 * k] * sinpi_q(a, N));
                }
                if (N % 2 == 0 && N > 1) acc += x[N - 1] * ((n & 1) ? -1.0L : 1.0L);
                y[n] = acc;
            }
        }
        if (c->r2h && !c->forward)
            for (size_t i = 2; i < N; i += 2) y[i] = -y[i];
    }
    for (size_t k = 0; k < N; ++k) *(double *)(c->out + oo + (ptrdiff_t)k * c->os) = (double)(y[k] * c->fct);
    return 0;
}

static int rfp_nd(size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout, size_t naxes,
                  const size_t *axes, int hartley, int r2h, int forward, const double *in, double *out, double fct)
{
    if (check_args(ndim, shape, naxes, axes)) return 1;
    if (prod(ndim, shape) == 0) return 0;
    if (in == out)
        for (size_t d = 0; d < ndim; ++d) if (sin[d] != sout[d]) return 1;
    int rc = 0;
    for (size_t a = 0; a < naxes && !rc; ++a) {
        rfp_ctx c;
        const size_t ax = axes[a];
        c.in = a ? (const char *)out : (const char *)in;
        c.out = (char *)out;
        c.is = a ? sout[ax] : sin[ax];
        c.os = sout[ax];
        c.n = shape[ax];
        c.hartley = hartley; c.r2h = r2h; c.forward = forward;
        c.fct = a ? 1.0L : (long double)fct;
        c.x = (long double *)malloc(c.n * sizeof(long double));
        c.y = (long double *)malloc(c.n * sizeof(long double));
        if (!c.x || !c.y) { free(c.x); free(c.y); return 2; }
        rc = for_each_line(ndim, shape, a ? sout : sin, sout, ax, rfp_line, &c);
        free(c.x); free(c.y);
    }
    return rc;
}

int oracle_r2r_fftpack(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                       size_t naxes, const size_t *axes, int real2hermitian, int forward, const double *data_in,
                       double *data_out, double fct)
{
    return rfp_nd(ndim, shape, stride_in, stride_out, naxes, axes, 0, real2hermitian != 0, forward != 0, data_in,
                  data_out, fct);
}

int oracle_r2r_separable_hartley(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                                 const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                                 const double *data_in, double *data_out, double fct)
{
    return rfp_nd(ndim, shape, stride_in, stride_out, naxes, axes, 1, 1, 1, data_in, data_out, fct);
}

int oracle_r2r_genuine_hartley(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                               const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                               const double *data_in, double *data_out, double fct)
{
    if (naxes == 1)
        return oracle_r2r_separable_hartley(ndim, shape, stride_in, stride_out, naxes, axes, data_in, data_out, fct);
    if (check_args(ndim, shape, naxes, axes)) return 1;
    if (prod(ndim, shape) == 0) return 0;
    if (ndim > 64) return 1;
    /* H[k] = sum_n x[n] (cos - sin)(2 pi sum_d k_d n_d / N_d) over the transformed dims, the
     * other indices shared between input and output; computed into a buffer first (in == out) */
    size_t total = prod(ndim, shape);
    long double *res = (long double *)malloc(total * sizeof(long double));
    if (!res) return 2;
    int isax[64] = {0};
    for (size_t a = 0; a < naxes; ++a) isax[axes[a]] = 1;
    size_t ko[64], ni[64];
    for (size_t e = 0; e < total; ++e) {
        size_t rem = e;
        for (size_t d = ndim; d-- > 0;) { ko[d] = rem % shape[d]; rem /= shape[d]; }
        /* sum over the input points sharing the non-transformed indices */
        size_t span = 1;
        for (size_t d = 0; d < ndim; ++d) if (isax[d]) span *= shape[d];
        long double acc = 0.0L;
        for (size_t f = 0; f < span; ++f) {
            size_t r2 = f;
            for (size_t d = ndim; d-- > 0;) {
                if (isax[d]) { ni[d] = r2 % shape[d]; r2 /= shape[d]; }
                else ni[d] = ko[d];
            }
            /* phase as a fraction of a turn, exact per axis */
            long double turn = 0.0L;
            ptrdiff_t off = 0;
            for (size_t d = 0; d < ndim; ++d) {
                off += (ptrdiff_t)ni[d] * stride_in[d];
                if (isax[d]) turn += (long double)((ko[d] * ni[d]) % shape[d]) / (long double)shape[d];
            }
            turn -= floorl(turn);
            const long double ang = 6.283185307179586476925286766559005768L * turn;
            acc += (long double)*(const double *)((const char *)data_in + off) * (cosl(ang) - sinl(ang));
        }
        res[e] = acc * (long double)fct;
    }
    for (size_t e = 0; e < total; ++e) {
        size_t rem = e;
        ptrdiff_t off = 0;
        for (size_t d = ndim; d-- > 0;) { off += (ptrdiff_t)(rem % shape[d]) * stride_out[d]; rem /= shape[d]; }
        *(double *)((char *)data_out + off) = (double)res[e];
    }
    free(res);
    return 0;
}

"""Pins of the CPU oracle (oracle/fft_oracle.c) -- no GPU needed.

Every oracle function is checked against (a) the golden vectors the
reference's own PocketFFT wrote (tests/golden/golden.npz and, for the DCT /
DST, tests/golden/golden_r2r.npz), (b) the reference
library built in place (oracle/_ref, when present), (c) a long-double naive
DFT, and (d) closed forms and the reference's known-answer tests, chosen so
that a wrong sign, a dropped term, a transposed operand or a wrong scale
factor fails at least one of them.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import ref
from tests.conftest import ROOT, rel_l2

TOL = 1e-14


def test_kat_fft2_nim():
    # tests/test_fft2.nim:7-10: fft([0,1,2,3]) == [6, -2+2i, -2, -2-2i] exactly
    x = np.arange(4.0)
    exp = np.array([6, -2 + 2j, -2, -2 - 2j])
    assert np.array_equal(oracle.c2c(x.astype(complex), [0]), exp)
    assert np.array_equal(oracle.r2c(x, [0]), exp[:3])
    # normalisations (test_fft2.nim:13-14): /4 and /2
    assert np.array_equal(oracle.c2c(x.astype(complex), [0], True, 0.25), exp / 4)
    assert np.array_equal(oracle.c2c(x.astype(complex), [0], True, 0.5), exp / 2)


def test_kat_readme():
    # README.md:41,49,68 -- fft([1, 2, 1, -1, 1.5])
    y = oracle.c2c(np.array([1, 2, 1, -1, 1.5], complex), [0])
    exp = np.array([4.5, 2.081559480312316 - 1.651098762732523j, -1.831559480312316 + 1.608220406444071j,
                    -1.831559480312316 - 1.608220406444071j, 2.081559480312316 + 1.651098762732523j])
    assert np.max(np.abs(y - exp)) < 1e-15


@pytest.mark.parametrize("n", list(range(1, 65)) + [97, 128, 255, 256, 257])
def test_naive_long_double(n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    for fwd in (True, False):
        assert rel_l2(oracle.c2c(x, [0], fwd), oracle.dft_naive(x, fwd)) < 1e-15


def test_naive_matches_definition():
    # the naive DFT itself against the textbook sum in numpy (sign and index order)
    n = 12
    x = np.random.default_rng(5).standard_normal(n) + 0j
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert rel_l2(oracle.dft_naive(x, True), W @ x) < 1e-14
    assert rel_l2(oracle.dft_naive(x, False), W.conj() @ x) < 1e-14


def test_twiddle_octants_exact():
    for n in (8, 16, 24, 4096, 4095 * 8):
        for k in range(0, n, n // 8):
            w = oracle.twiddle(k, n)
            assert w.real in (0.0, 1.0, -1.0) or abs(abs(w.real) - np.sqrt(0.5)) < 1e-17
            assert abs(w.real) == abs(w.imag) or 0.0 in (w.real, w.imag)


def test_closed_forms():
    n = 60
    d = np.zeros(n, complex)
    d[0] = 1
    assert np.array_equal(oracle.c2c(d, [0]), np.ones(n))        # impulse -> ones
    c = np.ones(n, complex)
    y = oracle.c2c(c, [0])
    assert y[0] == n and np.max(np.abs(y[1:])) < 1e-13          # constant -> n delta
    k0 = 7
    tone = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    y = oracle.c2c(tone, [0])
    assert abs(y[k0] - n) < 1e-12 and np.max(np.abs(np.delete(y, k0))) < 1e-12
    x = np.random.default_rng(3).standard_normal(n)
    X = oracle.c2c(x.astype(complex), [0])
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(x ** 2)) < 1e-12   # Parseval
    assert rel_l2(X[1:][::-1], X[1:].conj()) < 1e-15                  # Hermitian for real input


def test_golden_1d(golden):
    lengths = sorted({int(k.split("_")[1]) for k in golden.files if k.startswith("c2c_") and k.endswith("_in")})
    for n in lengths:
        x = golden[f"c2c_{n}_in"]
        assert rel_l2(oracle.c2c(x, [0], True), golden[f"c2c_{n}_fwd"]) < TOL, n
        if f"c2c_{n}_bwd" in golden.files:
            assert rel_l2(oracle.c2c(x, [0], False, 0.5), golden[f"c2c_{n}_bwd"]) < TOL, n
        r = golden[f"r2c_{n}_in"]
        assert rel_l2(oracle.r2c(r, [0], True), golden[f"r2c_{n}_out"]) < TOL, n
        assert rel_l2(oracle.rfft_packed(r, True), golden[f"hc_{n}_out"]) < TOL, n
        if f"r2c_{n}_bwd" in golden.files:
            assert rel_l2(oracle.r2c(r, [0], False, 2.0), golden[f"r2c_{n}_bwd"]) < TOL, n
            c = golden[f"c2r_{n}_in"]
            assert rel_l2(oracle.c2r(c, n, [0], False, 1.0 / n), golden[f"c2r_{n}_out"]) < TOL, n
            assert rel_l2(oracle.c2r(c, n, [0], True), golden[f"c2r_{n}_fwdflag"]) < TOL, n


def test_golden_nd(golden):
    a = golden["nd_c2c_in"]
    assert rel_l2(oracle.c2c(a[:, ::2], [0, 1], True), golden["nd_c2c_01"]) < TOL
    assert rel_l2(oracle.c2c(a[:, ::2], [1, 0], False, 0.25), golden["nd_c2c_10_bwd"]) < TOL
    b = golden["nd_r2c_in"]
    assert rel_l2(oracle.r2c(b, [0, 1, 2]), golden["nd_r2c_012"]) < TOL
    assert rel_l2(oracle.r2c(b, [2, 0]), golden["nd_r2c_20"]) < TOL
    assert rel_l2(oracle.r2c(b, [1], False), golden["nd_r2c_1_bwd"]) < TOL
    h = golden["nd_c2r_in"]
    assert rel_l2(oracle.c2r(h, 9, [0, 1, 2], False, 1.0 / 540), golden["nd_c2r_012"]) < TOL
    assert rel_l2(oracle.c2r(h, 8, [2], False), golden["nd_c2r_2"]) < TOL
    for n in (4096, 4095):
        assert rel_l2(oracle.r2c(golden[f"rows_{n}_in"], [1]), golden[f"rows_{n}_out"]) < TOL


def test_halfcomplex_roundtrip_all_lengths():
    # tests/test_fft.nim:29-49 (real round trip, rel-L2 <= 2e-15), on a stride of lengths
    rng = np.random.default_rng(0)
    for n in list(range(1, 300)) + [1009, 2017, 4095, 4096, 8191]:
        x = rng.uniform(-0.5, 0.5, n)
        y = oracle.rfft_packed(oracle.rfft_packed(x, True), False, 1.0 / n)
        assert rel_l2(y, x) <= 2e-15, n


@pytest.mark.skipif(not ref.available(), reason="reference PocketFFT not built here")
def test_against_reference_library():
    rng = np.random.default_rng(11)
    for n in list(range(1, 130)) + [1009, 2017, 4099, 8191, 10007]:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert rel_l2(oracle.c2c(x, [0]), ref.c2c(x, [0])) < 1e-14, n
        r = rng.uniform(-1, 1, n)
        assert rel_l2(oracle.r2c(r, [0], False, 3.0), ref.r2c(r, [0], False, 3.0)) < 1e-14, n
        c = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
        assert rel_l2(oracle.c2r(c, n, [0], True), ref.c2r(c, n, [0], True)) < 1e-14, n
        assert rel_l2(oracle.rfft_packed(r, False), ref.rfft_packed(r, False)) < 1e-14, n
        assert rel_l2(oracle.cfft(x, False, 0.5), ref.cfft(x, False, 0.5)) < 1e-14, n
    # n-D with negative strides
    a = rng.standard_normal((7, 9, 6)) + 1j * rng.standard_normal((7, 9, 6))
    v = a[::-1, :, ::-2]
    assert rel_l2(oracle.c2c(v, [2, 0]), ref.c2c(v, [2, 0])) < 1e-14


def test_oracle_rejects_bad_args():
    x = np.zeros(4, complex)
    with pytest.raises(ValueError):
        oracle.c2c(x, [1])        # bad axis number
    with pytest.raises(ValueError):
        oracle.c2c(np.zeros((2, 2), complex), [0, 0])   # axis specified repeatedly


# ---------------------------------------------------------------------------
# DCT / DST types 1-4 (pocketfft_hdronly.h:3284-3318)
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def golden_r2r():
    import os
    from tests.conftest import ROOT
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_r2r.npz"))


def _r2r_fn(name):
    return oracle.dct if name == "dct" else oracle.dst


def test_r2r_golden(golden_r2r):
    g = golden_r2r
    for key in g.files:
        if not key.startswith(("dct", "dst")):
            continue
        name, t, n = key[:3], int(key[3]), int(key.split("_")[1])
        ortho = key.endswith("_ortho")
        y = _r2r_fn(name)(g[f"r2r_{n}_in"], [0], t, 0.5 if ortho else 1.0, ortho)
        assert rel_l2(y, g[key]) < TOL, key
    v = g["nd_r2r_in"]
    for t in (1, 2, 3, 4):
        assert rel_l2(oracle.dct(v[:, ::2, :], [2, 0], t, 0.25, t == 2), g[f"nd_dct{t}_20"]) < TOL, t
        assert rel_l2(oracle.dst(v[::-1], [0, 1, 2], t, 1.0, t == 3), g[f"nd_dst{t}_012"]) < TOL, t


def test_r2r_kat_nim():
    # cpp_pocketfft/pocketfft.nim:322-337 applies DCTDesc(axes=[0]) (type 2) to [4, 3, 5, 10]
    y = oracle.dct(np.array([4.0, 3.0, 5.0, 10.0]), [0], 2)
    assert y[0] == 44.0
    # X[2] = 2 * sum x[n] cos(pi (2n+1) / 4) = sqrt(2) * (4 - 3 - 5 + 10)
    assert abs(y[2] - 6 * np.sqrt(2)) < 1e-14


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 13, 16, 31])
def test_r2r_inverse_pairs(n):
    """Unnormalised pairs: I.I, II.III, III.II and IV.IV give x times 2(N-1), 2N or 2(N+1)."""
    x = np.random.default_rng(n).uniform(-1, 1, n)
    for name in ("dct", "dst"):
        f = _r2r_fn(name)
        for a, b, scale in ((1, 1, 2 * (n - 1) if name == "dct" else 2 * (n + 1)),
                            (2, 3, 2 * n), (3, 2, 2 * n), (4, 4, 2 * n)):
            y = f(f(x, [0], a), [0], b, 1.0 / scale)
            assert rel_l2(y, x) < 1e-14, (name, a, b, n)


@pytest.mark.parametrize("n", [2, 3, 5, 8, 17])
def test_r2r_ortho_scalings(n):
    """With ortho=True (end-point scalings only) and fct = 1/sqrt(2N') the DCTs of every type and DST-I / DST-IV
    are orthonormal matrices; a missing or misplaced sqrt(2) breaks M^T M = I.  The reference's DST-II / DST-III
    scale element 0 (output 0 of DST-II by 1/sqrt(2), input 0 of DST-III by sqrt(2),
    pocketfft_hdronly.h:2519-2526 after the sine reversal) rather than element N-1, so those two are pinned
    as that exact row / column scaling of the plain matrix."""
    eye = np.eye(n)
    for name in ("dct", "dst"):
        f = _r2r_fn(name)
        for t in (1, 2, 3, 4):
            span = 2 * (n - 1) if (t == 1 and name == "dct") else 2 * (n + 1) if t == 1 else 2 * n
            m = f(eye, [0], t, 1.0 / np.sqrt(span), True)
            if name == "dst" and t in (2, 3):
                plain = f(eye, [0], t, 1.0 / np.sqrt(span), False)
                if t == 2:
                    plain[0, :] *= np.sqrt(0.5)
                else:
                    plain[:, 0] *= np.sqrt(2.0)
                assert np.max(np.abs(m - plain)) < 1e-15, (name, t, n)
            else:
                assert np.max(np.abs(m.T @ m - eye)) < 1e-14, (name, t, n)


def test_r2r_via_fft_of_even_extension():
    """DCT-II as the real part of a 2N-point DFT of [x, reversed x] rotated by e^{-i pi k / 2N}."""
    for n in (1, 2, 7, 64, 97):
        x = np.random.default_rng(n).uniform(-1, 1, n)
        z = oracle.dft_naive(np.concatenate([x, x[::-1]]).astype(complex))[:n]
        want = (np.exp(-1j * np.pi * np.arange(n) / (2 * n)) * z).real
        assert np.max(np.abs(oracle.dct(x, [0], 2) - want)) < 1e-13 * max(1, n), n
        # DST-I from the odd extension [0, x, 0, -reversed x]
        z = oracle.dft_naive(np.concatenate([[0.0], x, [0.0], -x[::-1]]).astype(complex))[1:n + 1]
        assert np.max(np.abs(oracle.dst(x, [0], 1) + z.imag)) < 1e-13 * max(1, n), n


def test_r2r_constant_input():
    n = 9
    y = oracle.dct(np.ones(n), [0], 2)
    assert y[0] == 2 * n and np.max(np.abs(y[1:])) < 1e-14
    y = oracle.dct(np.ones(n), [0], 1)
    assert abs(y[0] - 2 * (n - 1)) < 1e-14 and np.max(np.abs(y[1:])) < 1e-14


def test_r2r_rejects_bad_args():
    with pytest.raises(ValueError):
        oracle.dct(np.zeros(1), [0], 1)     # DCT-I needs N >= 2 (pocketfft_hdronly.h:2431)
    with pytest.raises(ValueError):
        oracle.dct(np.zeros(4), [0], 5)
    with pytest.raises(ValueError):
        oracle.dst(np.zeros(4), [1], 2)


@pytest.mark.skipif(not ref.available(), reason="reference PocketFFT not built here")
def test_r2r_against_reference_library():
    rng = np.random.default_rng(12)
    for n in list(range(1, 70)) + [127, 128, 500, 1023]:
        x = rng.uniform(-1, 1, n)
        for ortho in (False, True):
            for t in (1, 2, 3, 4):
                if not (t == 1 and n == 1):
                    assert rel_l2(oracle.dct(x, [0], t, 0.7, ortho), ref.dct(x, [0], t, 0.7, ortho)) < TOL, (n, t)
                assert rel_l2(oracle.dst(x, [0], t, 0.7, ortho), ref.dst(x, [0], t, 0.7, ortho)) < TOL, (n, t)


# ---------------------------------------------------------------------------
# r2r_fftpack and the Hartley transforms (pocketfft_hdronly.h:3392-3445)
# ---------------------------------------------------------------------------

def test_fftpack_hartley_golden(golden_r2r):
    g = golden_r2r
    n_checked = 0
    for key in g.files:
        if key.startswith("fp") and key[2:4].isdigit() and not key.startswith("fp_"):
            r2h, fw, n = int(key[2]), int(key[3]), int(key.split("_")[1])
            y = oracle.r2r_fftpack(g[f"fp_{n}_in"], [0], r2h, fw, 1.0 if fw else 0.5)
            assert rel_l2(y, g[key]) < TOL, key
            n_checked += 1
        elif key.startswith("hartley_"):
            n = int(key.split("_")[1])
            assert rel_l2(oracle.r2r_separable_hartley(g[f"fp_{n}_in"], [0]), g[key]) < TOL, key
            n_checked += 1
    assert n_checked == 5 * 18
    v = g["nd_r2r_in"]
    assert rel_l2(oracle.r2r_fftpack(v[:, ::2, :], [2, 0], True, True, 0.25), g["nd_fp11_20"]) < TOL
    assert rel_l2(oracle.r2r_fftpack(v[::-1], [1], False, False), g["nd_fp00_1"]) < TOL
    assert rel_l2(oracle.r2r_fftpack(v, [0, 2], False, True), g["nd_fp01_02"]) < TOL
    assert rel_l2(oracle.r2r_separable_hartley(v, [0, 1, 2], 0.5), g["nd_hartley_012"]) < TOL
    assert rel_l2(oracle.r2r_genuine_hartley(v, [1, 2]), g["nd_genuine_12"]) < TOL
    assert rel_l2(oracle.r2r_genuine_hartley(v[:, ::-1, :], [2, 0, 1], 2.0), g["nd_genuine_201"]) < TOL


@pytest.mark.parametrize("n", [1, 2, 3, 8, 9, 30, 31])
def test_fftpack_hartley_closed_forms(n):
    """Independent statements: the forward halfcomplex equals the C path's rfft_packed; backward after
    forward returns N x; the separable Hartley equals Re + Im of the oracle's r2c spectrum extended by
    Hermitian symmetry and is its own inverse up to N."""
    x = np.random.default_rng(n).uniform(-1, 1, n)
    hc = oracle.r2r_fftpack(x, [0], True, True)
    assert rel_l2(hc, oracle.rfft_packed(x, True)) < TOL
    assert rel_l2(oracle.r2r_fftpack(hc, [0], False, False, 1.0 / n), x) < TOL
    X = oracle.c2c(x.astype(complex), [0])
    assert rel_l2(oracle.r2r_separable_hartley(x, [0]), X.real + X.imag) < TOL
    assert rel_l2(oracle.r2r_separable_hartley(oracle.r2r_separable_hartley(x, [0]), [0], 1.0 / n), x) < TOL
    # the mixed flag pairs: sign flips of the positions 2, 4, ... around the plain operations
    flip = np.ones(n)
    flip[2::2] = -1.0
    assert rel_l2(oracle.r2r_fftpack(x, [0], False, True), oracle.r2r_fftpack(x * flip, [0], True, True)) < TOL
    assert rel_l2(oracle.r2r_fftpack(x, [0], True, False), oracle.r2r_fftpack(x, [0], False, False) * flip) < TOL


def test_genuine_hartley_nd_closed_forms():
    v = np.random.default_rng(13).uniform(-1, 1, (4, 6, 5))
    X = oracle.c2c(v.astype(complex), [0, 2])
    assert rel_l2(oracle.r2r_genuine_hartley(v, [0, 2]), X.real + X.imag) < TOL
    # the genuine transform is an involution up to the number of points
    h = oracle.r2r_genuine_hartley(v, [0, 1, 2])
    assert rel_l2(oracle.r2r_genuine_hartley(h, [0, 1, 2], 1.0 / v.size), v) < TOL
    # one axis: the separable transform
    assert rel_l2(oracle.r2r_genuine_hartley(v, [1]), oracle.r2r_separable_hartley(v, [1])) < TOL


@pytest.mark.skipif(not ref.available(), reason="reference PocketFFT not built here")
def test_fftpack_hartley_against_reference_library():
    rng = np.random.default_rng(14)
    for n in list(range(1, 50)) + [64, 97, 128, 243, 1000]:
        x = rng.uniform(-1, 1, n)
        for r2h in (True, False):
            for fw in (True, False):
                assert rel_l2(oracle.r2r_fftpack(x, [0], r2h, fw, 0.7), ref.r2r_fftpack(x, [0], r2h, fw, 0.7)) < TOL
        assert rel_l2(oracle.r2r_separable_hartley(x, [0], 1.3), ref.r2r_separable_hartley(x, [0], 1.3)) < TOL
    v = rng.uniform(-1, 1, (5, 6, 7))
    for axes in ([0, 1], [2, 0], [0, 1, 2]):
        assert rel_l2(oracle.r2r_genuine_hartley(v, axes, 0.5), ref.r2r_genuine_hartley(v, axes, 0.5)) < TOL
        for r2h in (True, False):
            for fw in (True, False):
                assert rel_l2(oracle.r2r_fftpack(v[::-1, :, ::2], axes, r2h, fw),
                              ref.r2r_fftpack(v[::-1, :, ::2], axes, r2h, fw)) < TOL


def test_oracle_vs_reference_float_golden():
    """The reference's T = float instantiation (pocketfft::c2c / r2c / c2r<float>, written by
    tests/golden/make_golden.py --f32) against the oracle on the same float inputs widened to double:
    the two differ only by the reference's float rounding, <= 5e-7 rel-L2 here (measured 2.3e-7)."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "golden_f32.npz"))
    ns = sorted({int(k.split("_")[1]) for k in g.files if k.startswith("c2c_") and k.endswith("_in")})
    assert len(ns) >= 40
    for n in ns:
        x = g[f"c2c_{n}_in"].astype(np.complex128)
        assert rel_l2(g[f"c2c_{n}_fwd"], oracle.c2c(x, [0], True)) < 5e-7, n
        assert rel_l2(g[f"c2c_{n}_bwd"], oracle.c2c(x, [0], False, 0.5)) < 5e-7, n
        r = g[f"r2c_{n}_in"].astype(np.float64)
        assert rel_l2(g[f"r2c_{n}_out"], oracle.r2c(r, [0], True)) < 5e-7, n
        assert rel_l2(g[f"r2c_{n}_bwd"], oracle.r2c(r, [0], False, 2.0)) < 5e-7, n
        c = g[f"c2r_{n}_in"].astype(np.complex128)
        assert rel_l2(g[f"c2r_{n}_out"], oracle.c2r(c, n, [0], False, 1.0 / n)) < 5e-7, n
    a = g["nd_c2c_in"].astype(np.complex128)
    assert rel_l2(g["nd_c2c_01"], oracle.c2c(a[:, ::2], [0, 1], True)) < 5e-7
    assert rel_l2(g["nd_c2c_10_bwd"], oracle.c2c(a[::-1, :], [1, 0], False, 0.25)) < 5e-7
    b = g["nd_r2c_in"].astype(np.float64)
    assert rel_l2(g["nd_r2c_012"], oracle.r2c(b, [0, 1, 2], True)) < 5e-7
    assert rel_l2(g["nd_r2c_20"], oracle.r2c(b, [2, 0], True)) < 5e-7
    h = g["nd_c2r_in"].astype(np.complex128)
    assert rel_l2(g["nd_c2r_012"], oracle.c2r(h, 9, [0, 1, 2], False, 1.0 / 540)) < 5e-7
    for n in (4096, 4095):
        assert rel_l2(g[f"rows_{n}_out"], oracle.r2c(g[f"rows_{n}_in"].astype(np.float64), [1])) < 5e-7

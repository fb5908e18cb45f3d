"""Pins of the CPU oracle (oracle/fft_oracle.c) -- no GPU needed.

Every oracle function is checked against (a) the golden vectors the
reference's own PocketFFT wrote (tests/golden/golden.npz), (b) the reference
library built in place (oracle/_ref, when present), (c) a long-double naive
DFT, and (d) closed forms and the reference's known-answer tests, chosen so
that a wrong sign, a dropped term, a transposed operand or a wrong scale
factor fails at least one of them.
"""
import numpy as np
import pytest

import oracle
from oracle import ref
from tests.conftest import rel_l2

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

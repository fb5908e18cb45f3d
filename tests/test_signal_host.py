"""Host side of resampling (impulse_amd/signal.py) and the oracle's upfirdn -- no GPU needed.

The filter design and resample's alignment logic are checked against the values
the reference's own tests expect (tests/test_signal.nim), with upfirdn replaced
by an independent numpy statement of its definition (zero insertion, then
numpy.convolve, then every down-th sample).  The oracle's upfirdn is pinned by
the reference's exact KATs, numpy.convolve and closed forms.
"""
import json
import os

import numpy as np
import pytest

import oracle
from impulse_amd import signal as S
from tests.conftest import ROOT


def _mae(a, b):
    return float(np.mean(np.abs(np.asarray(a) - np.asarray(b))))


def _np_upfirdn(t, h, up=1, down=1, algorithm=0):
    t = np.asarray(t)
    u = np.zeros((len(t) - 1) * up + 1, dtype=np.result_type(t.dtype, np.float64))
    u[::up] = t
    return np.convolve(u, np.asarray(h))[::down]


# tests/test_signal.nim:5-16
def test_kaiser_kat():
    assert _mae(S.kaiser(4), [0.03671089, 0.7753221, 0.7753221, 0.03671089]) < 1e-8
    assert _mae(S.kaiser(4), S.kaiser(4, 5.0)) < 1e-8
    assert _mae(S.kaiser(4, 8.0), [0.00233883, 0.65247867, 0.65247867, 0.00233883]) < 1e-8
    # signal.nim:216-220: kaiser(12, 14.0)
    assert _mae(S.kaiser(12, 14.0)[:3], [7.72687e-06, 0.00346009, 0.04652]) < 1e-8
    assert S.kaiser(1).tolist() == [1.0] and S.kaiser(0).size == 0
    with pytest.raises(AssertionError):
        S.kaiser(-1)


def test_kaiser_closed_form():
    # beta = 0 is the rectangular window; the window is symmetric with its peak 1 at the centre
    assert np.allclose(S.kaiser(7, 0.0), 1.0)
    w = S.kaiser(9, 6.0)
    assert np.allclose(w, w[::-1]) and w[4] == 1.0


# tests/test_signal.nim:18-65
FIRLS_KATS = [
    (dict(fir_order=4, bands=[[0.0, 0.3], [0.4, 1.0]], desired=[[1.0, 1.0], [0.0, 0.0]]),
     [0.1264964, 0.278552488, 0.34506779765, 0.278552488, 0.1264964]),
    (dict(fir_order=4, bands=[0.0, 0.3, 0.4, 1.0], desired=[1.0, 1.0, 0.0, 0.0]),
     [0.1264964, 0.278552488, 0.34506779765, 0.278552488, 0.1264964]),
    (dict(fir_order=4, bands=[[0.0, 1.5], [2.0, 5.0]], desired=[[1.0, 1.0], [0.0, 0.0]], fs=10.0),
     [0.1264964, 0.278552488, 0.34506779765, 0.278552488, 0.1264964]),
    (dict(fir_order=4, bands=[[0.0, 0.3], [0.4, 1.0]], desired=[[1.0, 1.0], [0.0, 0.0]], weights=[1.0, 0.5]),
     [0.110353444, 0.28447325, 0.36086805, 0.28447325, 0.110353444]),
    (dict(fir_order=5, bands=[[0.0, 0.3], [0.3, 0.6], [0.6, 1.0]], desired=[[1.0, 1.0], [1.0, 0.2], [0.0, 0.0]]),
     [-0.05603328, 0.146107441, 0.43071645, 0.43071645, 0.146107441, -0.05603328]),
    (dict(fir_order=6, bands=[[0.0, 0.4], [0.6, 1.0]], desired=[[0.0, 0.0], [0.9, 1.0]], symmetric=False),
     [-0.13944975, 0.2851858, -0.25859575, 0.0, 0.25859575, -0.2851858, 0.13944975]),
]


@pytest.mark.parametrize("i", range(len(FIRLS_KATS)))
def test_firls_kat(i):
    kw, want = FIRLS_KATS[i]
    assert _mae(S.firls(**kw), want) < 1e-8


def test_firls_least_squares_optimum():
    """With gaps or unequal weights the taps solve the weighted least-squares problem: perturbing any
    coefficient pair (keeping the symmetry) raises the integrated error, checked on a fine grid."""
    bands, desired, weights = [0.0, 0.25, 0.35, 1.0], [1.0, 1.0, 0.0, 0.0], [1.0, 3.0]
    h = S.firls(10, bands, desired, weights)
    assert np.allclose(h, h[::-1])
    f = np.linspace(0, 0.5, 20001)

    def err(taps):
        n = np.arange(taps.size) - (taps.size - 1) / 2.0
        amp = (taps[None, :] * np.cos(2 * np.pi * f[:, None] * n[None, :])).sum(1)
        tot = 0.0
        for b in range(2):
            lo, hi = bands[2 * b] / 2.0, bands[2 * b + 1] / 2.0
            m = (f >= lo) & (f <= hi)
            d = np.interp(f[m], [lo, hi], desired[2 * b:2 * b + 2])
            tot += weights[b] * np.trapezoid((amp[m] - d) ** 2, f[m])
        return tot

    base = err(h)
    for k in range(5):
        e = np.zeros_like(h)
        e[k] = e[-1 - k] = 1e-3
        assert err(h + e) > base and err(h - e) > base


def test_firls_errors():
    with pytest.raises(ValueError):
        S.firls(4, [0.0, 0.3, 0.4, 1.0], [1.0, 1.0, 0.0, 0.0], fs=0.0)
    with pytest.raises(ValueError):
        S.firls(4, [0.0, 0.3, 0.4, 1.5], [1.0, 1.0, 0.0, 0.0])         # above Nyquist
    with pytest.raises(ValueError):
        S.firls(4, [0.0, 0.3, 0.4], [1.0, 1.0, 0.0])                   # odd length
    with pytest.raises(ValueError):
        S.firls(4, [0.0, 0.3, 0.4, 1.0], [1.0, 1.0, 0.0])              # length mismatch
    with pytest.raises(ValueError):
        S.firls(5, [0.0, 0.5, 0.6, 1.0], [0.0, 0.0, 1.0, 1.0], symmetric=False)   # signal.nim:420-424
    with pytest.raises(ValueError):
        S.firls(4, [0.0, 0.3, 0.4, 1.0], [1.0, 1.0, 0.0, 0.0], weights=[1.0])
    with pytest.raises(ValueError):
        S.firls(4, [0.0, 0.4, 0.3, 1.0], [1.0, 1.0, 0.0, 0.0])         # not monotonic


def test_reduce_and_generate():
    assert S.reduce_resampling_rates(6, 4) == (3, 2)
    assert S.reduce_resampling_rates(7, 5) == (7, 5)
    h = S.generate_resampling_filter(3, 2)
    assert h.size == 2 * 10 * 3 + 1 and abs(h.sum() - 3.0) < 1e-12 and np.allclose(h, h[::-1])
    assert S.generate_resampling_filter(3, 2, fir_order_factor=0).tolist() == [1.0, 1.0, 1.0]


T = np.array([1.0, 1.2, -0.3, -1.0, 2.0])
E32 = [1.0006061736, 1.2269863677, 1.0003755977, -0.3001818521, -1.4444776098, 0.1147783685, 2.0012123471,
       0.9175789829]


# tests/test_signal.nim:87-147, with upfirdn replaced by its numpy statement
def test_resample_host_logic(monkeypatch):
    monkeypatch.setattr(S, "upfirdn", _np_upfirdn)
    up3 = [1.0006061736, 1.1652623795, 1.2269863677, 1.2007274083, 1.0003755977, 0.4963084977, -0.3001818521,
           -1.1022817902, -1.4444776098, -1.0006061736, 0.1147783685, 1.3383072463, 2.0012123471, 1.7824018717,
           0.9175789829]
    assert _mae(S.resample(T, up=3), up3) < 1e-8
    assert _mae(S.resample(T, down=2), [0.9812293475, -0.0867375515, 0.5627008880]) < 1e-8
    assert _mae(S.resample(T, up=3, down=2), E32) < 1e-8
    assert np.array_equal(S.resample(T, up=3, down=2), S.resample(T, up=6, down=4))
    assert _mae(S.resample(T, S.generate_resampling_filter(3, 2), up=3, down=2), E32) < 1e-8
    assert _mae(S.resample(T, up=3, down=2, fir_order_factor=30, beta=6.0),
                [1.0000890770, 1.1907035514, 1.0289575738, -0.3000267231, -1.4576402433, 0.1203579771,
                 2.0001781539, 0.9282917306]) < 1e-8
    assert _mae(S.resample(T, up=3, down=2, fir_order_factor=0), [1.0, 1.2, 1.2, -0.3, -1, -1, 2, 0]) < 1e-8
    r = S.resample(T + 2j * T, up=3, down=2)
    assert _mae(r.real, E32) < 1e-8 and _mae(r.imag, 2 * np.array(E32)) < 1e-8
    assert np.array_equal(S.resample(T, up=2, down=2), T)
    with pytest.raises(ValueError):
        S.resample(T, np.ones((2, 3)), up=2)


def test_resample_lengths(monkeypatch):
    """tests/test_signal.nim:149-256: output lengths for inputs 100..119 long and up, down in 2..9."""
    monkeypatch.setattr(S, "upfirdn", _np_upfirdn)
    with open(os.path.join(ROOT, "tests", "golden", "resample_lengths.json")) as f:
        want = json.load(f)["lengths"]
    t = np.arange(-50.0, 100.0)
    got = []
    for n in range(100, 120):
        for up in range(2, 10):
            for down in range(2, 10):
                got.append(len(S.resample(t[:n], up=up, down=down)))
    assert got == want


def test_adjust_filter_position_invariants():
    for n, h, up, down in [(5, 61, 3, 2), (100, 161, 1, 8), (37, 40, 7, 3), (10, 1, 2, 5)]:
        rl = int(np.ceil(n * up / down))
        hh, delay = S.adjust_filter_position(np.ones(h), n, rl, up, down)
        lead = int(np.argmax(hh != 0))
        mid = lead + (h - 1) / 2.0
        # the kept samples land on the filter's middle point, and upfirdn is long enough to cover the result
        assert (np.ceil(mid)) / down >= delay and delay == np.floor(np.ceil(mid) / down)
        assert oracle.upfirdn_length(n, hh.size, up, down) > delay + rl - 1


# ---------------------------------------------------------------------------
# the oracle's upfirdn
# ---------------------------------------------------------------------------

def test_oracle_upfirdn_kats():
    # tests/test_signal.nim:67-85 (exact)
    assert oracle.upfirdn([1, 1, 1], [1, 1, 1]).tolist() == [1, 2, 3, 2, 1]
    assert oracle.upfirdn([1, 2, 3], [1], 3).tolist() == [1, 0, 0, 2, 0, 0, 3]
    assert oracle.upfirdn([1, 2, 3], [1, 1, 1], 3).tolist() == [1, 1, 1, 2, 2, 2, 3, 3, 3]
    assert oracle.upfirdn([1.0, 1.0, 1.0], [0.5, 1.0, 0.5], 2).tolist() == [0.5, 1.0, 1.0, 1.0, 1.0, 1.0, 0.5]
    assert oracle.upfirdn(np.arange(10), [1], 1, 3).tolist() == [0, 3, 6, 9]
    assert oracle.upfirdn(np.arange(10.0), [0.5, 1.0, 0.5], 2, 3).tolist() == [0.0, 1.0, 2.5, 4, 5.5, 7.0, 8.5]


def test_oracle_upfirdn_vs_numpy():
    rng = np.random.default_rng(3)
    for nt, nh, up, down in [(1, 1, 1, 1), (1, 5, 3, 2), (50, 7, 1, 1), (33, 12, 4, 3), (64, 64, 2, 7),
                             (17, 100, 5, 1), (200, 3, 1, 9)]:
        t, h = rng.uniform(-1, 1, nt), rng.uniform(-1, 1, nh)
        want = _np_upfirdn(t, h, up, down)
        got = oracle.upfirdn(t, h, up, down)
        assert got.shape == want.shape and np.max(np.abs(got - want)) < 1e-14, (nt, nh, up, down)


def test_oracle_upfirdn_closed_forms():
    t = np.random.default_rng(4).uniform(-1, 1, 40)
    # a unit filter only upsamples / decimates
    u = oracle.upfirdn(t, [1.0], 3, 1)
    assert np.array_equal(u[::3], t) and not np.any(u[np.arange(u.size) % 3 != 0])
    assert np.array_equal(oracle.upfirdn(t, [1.0], 1, 4), t[::4])
    # a delayed unit filter shifts; linearity in the filter
    assert np.array_equal(oracle.upfirdn(t, [0.0, 0.0, 1.0]), np.concatenate([[0, 0], t]))
    a, b = np.arange(1.0, 6.0), np.arange(5.0, 0.0, -1.0)
    assert np.allclose(oracle.upfirdn(t, a + 2 * b, 2, 3), oracle.upfirdn(t, a, 2, 3) + 2 * oracle.upfirdn(t, b, 2, 3),
                       atol=1e-13)

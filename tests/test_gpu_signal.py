"""GPU parity of upfirdn (impulse/signal.nim:545-594) and resample
(impulse/signal.nim:690-809) through impulse_fft_upfirdn_*: the direct and the
overlap-save FFT algorithms against the oracle's long-double sums and each
other, the reference's own KATs and expected resample outputs
(tests/test_signal.nim:67-147).

Bars: rel-L2 <= 1e-12 against the oracle; the reference's integer and dyadic
KATs bit-exact on the direct algorithm and within 1e-14 on the FFT one; the
reference's resample vectors at its own 1e-8 mean absolute error.
"""
import numpy as np
import pytest

import oracle
from impulse_amd import _lib
from impulse_amd import signal as S
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12
ALGOS = [_lib.UPFIRDN_DIRECT, _lib.UPFIRDN_FFT]


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().impulse_fft_device_check())
    yield
    torch.cuda.synchronize()


def _mae(a, b):
    return float(np.mean(np.abs(np.asarray(a) - np.asarray(b))))


KATS = [(([1, 1, 1], [1, 1, 1], 1, 1), [1, 2, 3, 2, 1]),
        (([1, 2, 3], [1], 3, 1), [1, 0, 0, 2, 0, 0, 3]),
        (([1, 2, 3], [1, 1, 1], 3, 1), [1, 1, 1, 2, 2, 2, 3, 3, 3]),
        (([1.0, 1.0, 1.0], [0.5, 1.0, 0.5], 2, 1), [0.5, 1.0, 1.0, 1.0, 1.0, 1.0, 0.5]),
        ((list(range(10)), [1], 1, 3), [0, 3, 6, 9]),
        ((list(np.arange(10.0)), [0.5, 1.0, 0.5], 2, 3), [0.0, 1.0, 2.5, 4, 5.5, 7.0, 8.5])]


@pytest.mark.parametrize("algo", [_lib.UPFIRDN_AUTO] + ALGOS)
def test_reference_kats(algo):
    # tests/test_signal.nim:67-85
    for (t, h, up, down), want in KATS:
        y = S.upfirdn(np.array(t), np.array(h), up, down, algorithm=algo)
        assert y.shape == (len(want),)
        if algo == _lib.UPFIRDN_FFT:
            assert np.max(np.abs(y - np.array(want, float))) < 1e-14, (t, h, up, down)
        else:
            assert np.array_equal(y, np.array(want)), (t, h, up, down, y)
        if isinstance(t[0], int):
            assert np.issubdtype(y.dtype, np.integer)


CASES = [(1, 1, 1, 1), (1, 7, 3, 2), (2, 1, 1, 1), (5, 61, 3, 2), (100, 7, 1, 1), (33, 12, 4, 3),
         (64, 64, 2, 7), (17, 100, 5, 1), (200, 3, 1, 9), (1000, 161, 1, 8), (777, 301, 7, 3),
         (4096, 33, 2, 1), (5000, 1001, 1, 1), (3001, 2000, 3, 5), (20000, 4097, 1, 2), (2500, 5000, 2, 3)]


@pytest.mark.parametrize("algo", ALGOS)
def test_vs_oracle(algo):
    rng = np.random.default_rng(21 + algo)
    for nt, nh, up, down in CASES:
        t, h = rng.uniform(-1, 1, nt), rng.standard_normal(nh)
        y = S.upfirdn(t, h, up, down, algorithm=algo)
        want = oracle.upfirdn(t, h, up, down)
        assert y.shape == want.shape, (nt, nh, up, down)
        assert rel_l2(y, want) < TOL, (nt, nh, up, down, rel_l2(y, want))


def test_fft_blocks_and_auto_choice():
    """The FFT algorithm splits long signals into overlap-save blocks; the cost model keeps short filters
    direct and sends long ones to the FFT."""
    p = S.UpfirdnPlan(1 << 20, np.ones(2049), 1, 1, algorithm=_lib.UPFIRDN_FFT)
    assert p.blocks > 1 and p.fft_length >= 4 * 2048
    assert S.UpfirdnPlan(1 << 20, np.ones(61), 3, 2).algorithm == _lib.UPFIRDN_DIRECT
    assert S.UpfirdnPlan(1 << 20, np.ones(8001), 1, 1).algorithm == _lib.UPFIRDN_FFT


@pytest.mark.parametrize("nt,nh,up,down", [(1 << 20, 2049, 1, 1), (300000, 1201, 3, 2), (123457, 3001, 2, 5),
                                           (1 << 18, 61, 3, 2)])
def test_large_direct_vs_fft(nt, nh, up, down):
    """Sizes past the oracle: the two algorithms (sums of products vs. overlap-save spectra) agree, and a
    sampled window of outputs matches the oracle on the slice of input it depends on."""
    rng = np.random.default_rng(nt % 1000)
    t, h = rng.uniform(-1, 1, nt), rng.standard_normal(nh) / np.sqrt(nh)
    a = S.upfirdn(t, h, up, down, algorithm=_lib.UPFIRDN_DIRECT)
    b = S.upfirdn(t, h, up, down, algorithm=_lib.UPFIRDN_FFT)
    assert a.shape == b.shape == (oracle.upfirdn_length(nt, nh, up, down),)
    assert rel_l2(b, a) < TOL
    # outputs j0 .. j0+63 depend on t[q] for q in [(j0 down - nh + 1) / up, (j0 + 63) down / up]
    j0 = a.size // 2
    q0 = max(0, (j0 * down - nh + 1) // up)
    q1 = min(nt - 1, ((j0 + 63) * down) // up)
    # re-index: the window's first output sits at m = j0 down, i.e. m - q0 up inside the sliced problem
    off = j0 * down - q0 * up
    sub = oracle.upfirdn(t[q0:q1 + 1], h, up, 1)
    want = sub[off:off + 64 * down:down]
    assert rel_l2(a[j0:j0 + 64], want) < TOL and rel_l2(b[j0:j0 + 64], want) < TOL


def test_batch_and_device_tensors():
    import torch
    rng = np.random.default_rng(5)
    t = rng.uniform(-1, 1, (3, 4000))
    h = rng.standard_normal(257)
    for algo in ALGOS:
        p = S.UpfirdnPlan(4000, h, 4, 3, batch=3, algorithm=algo)
        y = p.execute(torch.from_numpy(t).cuda())
        torch.cuda.synchronize()
        yh = p.execute(t)
        for b in range(3):
            want = oracle.upfirdn(t[b], h, 4, 3)
            assert rel_l2(y[b].cpu().numpy(), want) < TOL and rel_l2(yh[b], want) < TOL
        # repeated execution on the same plan and stream gives the same result
        y2 = p.execute(torch.from_numpy(t).cuda())
        torch.cuda.synchronize()
        assert torch.equal(y, y2)
    # rank-1 device tensors through upfirdn and resample, float64 and complex128
    td = torch.from_numpy(t[0]).cuda()
    assert rel_l2(S.upfirdn(td, h, 2, 5).cpu().numpy(), oracle.upfirdn(t[0], h, 2, 5)) < TOL
    tc = torch.complex(td, -2 * td)
    yc = S.upfirdn(tc, h, 2, 5).cpu().numpy()
    assert rel_l2(yc.real, oracle.upfirdn(t[0], h, 2, 5)) < TOL
    assert rel_l2(yc.imag, -2 * oracle.upfirdn(t[0], h, 2, 5)) < TOL
    rd = S.resample(td, up=3, down=2).cpu().numpy()
    assert rel_l2(rd, S.resample(t[0], up=3, down=2)) < TOL


def test_device_tensor_complex_filter_and_integers():
    """A complex filter on the torch path runs as its real and imaginary taps (the host path's
    rule); integer signals with integer filters round, and keep an integer dtype."""
    import torch
    rng = np.random.default_rng(9)
    t = rng.uniform(-1, 1, 3000)
    h = rng.standard_normal(41) + 1j * rng.standard_normal(41)
    want = oracle.upfirdn(t, h.real, 3, 2) + 1j * oracle.upfirdn(t, h.imag, 3, 2)
    for algo in ALGOS:
        y = S.upfirdn(torch.from_numpy(t).cuda(), torch.from_numpy(h), 3, 2, algorithm=algo)
        assert y.is_complex() and y.is_cuda
        assert rel_l2(y.cpu().numpy(), want) < TOL
        # complex signal, complex filter: (a + ib) * (c + id) by the four real products
        tc = torch.complex(torch.from_numpy(t), torch.from_numpy(-0.5 * t)).cuda()
        yc = S.upfirdn(tc, h, 3, 2, algorithm=algo).cpu().numpy()
        assert rel_l2(yc, want * (1 - 0.5j)) < TOL
    # integer KATs on device tensors: exact integers, integer dtype
    for (ti, hi, up, down), expect in KATS[:3] + KATS[4:5]:
        yi = S.upfirdn(torch.tensor(ti, dtype=torch.int64).cuda(), np.asarray(hi, np.int64), up, down)
        assert yi.dtype == torch.int64
        assert yi.cpu().tolist() == expect
    # negative sums: rounding, not truncation toward zero, and the int32 dtype kept
    ti = torch.tensor([3, -7, 11, 5], dtype=torch.int32).cuda()
    yi = S.upfirdn(ti, np.array([2, -1, 4], np.int32), 2, 1)
    assert yi.dtype == torch.int32
    assert yi.cpu().tolist() == [6, -3, -2, 7, -6, -11, 54, -5, 20]


T = np.array([1.0, 1.2, -0.3, -1.0, 2.0])
E32 = [1.0006061736, 1.2269863677, 1.0003755977, -0.3001818521, -1.4444776098, 0.1147783685, 2.0012123471,
       0.9175789829]


@pytest.mark.parametrize("algo", ALGOS)
def test_resample_reference_vectors(algo):
    # tests/test_signal.nim:87-147
    up3 = [1.0006061736, 1.1652623795, 1.2269863677, 1.2007274083, 1.0003755977, 0.4963084977, -0.3001818521,
           -1.1022817902, -1.4444776098, -1.0006061736, 0.1147783685, 1.3383072463, 2.0012123471, 1.7824018717,
           0.9175789829]
    assert _mae(S.resample(T, up=3, algorithm=algo), up3) < 1e-8
    assert _mae(S.resample(T, down=2, algorithm=algo), [0.9812293475, -0.0867375515, 0.5627008880]) < 1e-8
    r32 = S.resample(T, up=3, down=2, algorithm=algo)
    assert _mae(r32, E32) < 1e-8
    assert np.array_equal(r32, S.resample(T, up=6, down=4, algorithm=algo))
    assert _mae(S.resample(T, S.generate_resampling_filter(3, 2), up=3, down=2, algorithm=algo), E32) < 1e-8
    assert _mae(S.resample(T, up=3, down=2, fir_order_factor=30, beta=6.0, algorithm=algo),
                [1.0000890770, 1.1907035514, 1.0289575738, -0.3000267231, -1.4576402433, 0.1203579771,
                 2.0001781539, 0.9282917306]) < 1e-8
    assert _mae(S.resample(T, up=3, down=2, fir_order_factor=0, algorithm=algo),
                [1.0, 1.2, 1.2, -0.3, -1, -1, 2, 0]) < 1e-8
    rc = S.resample(T + 2j * T, up=3, down=2, algorithm=algo)
    assert _mae(rc.real, E32) < 1e-8 and _mae(rc.imag, 2 * np.array(E32)) < 1e-8


def test_resample_long_signal_vs_oracle():
    """resample = a slice of upfirdn with the adjusted filter; checked against the oracle on a 20,000-sample
    signal at rates 7/3 (generated filter of 141 taps)."""
    t = np.sin(np.linspace(0, 200, 20000)) + 0.1 * np.random.default_rng(9).standard_normal(20000)
    y = S.resample(t, up=7, down=3)
    h = S.generate_resampling_filter(7, 3)
    hh, delay = S.adjust_filter_position(h, t.size, y.size, 7, 3)
    want = oracle.upfirdn(t, hh, 7, 3)[delay:delay + y.size]
    assert y.size == int(np.ceil(20000 * 7 / 3)) and rel_l2(y, want) < TOL


def test_errors():
    with pytest.raises(_lib.FFTError):
        S.UpfirdnPlan(10, np.ones(3), 0, 1)
    with pytest.raises(_lib.FFTError):
        S.UpfirdnPlan(10, np.ones(3), 1, 0)
    with pytest.raises(_lib.FFTError):
        S.UpfirdnPlan(0, np.ones(3), 1, 1)
    with pytest.raises(_lib.FFTError):
        S.UpfirdnPlan(10, np.ones(0), 1, 1)
    with pytest.raises(_lib.FFTError):
        S.UpfirdnPlan(10, np.ones(3), 1, 1, algorithm=3)
    with pytest.raises(ValueError):
        S.upfirdn(np.ones((2, 2)), np.ones(3))
    p = S.UpfirdnPlan(10, np.ones(3))
    with pytest.raises(ValueError):
        p.execute(np.ones(11))

"""GPU parity of the DCT / DST types 1-4 (pocketfft::dct / dst,
pocketfft_hdronly.h:3284-3318; DCTDesc apply, cpp_pocketfft/pocketfft.nim:279-295)
against the golden vectors the reference's own PocketFFT wrote, the oracle and,
where it was built, the reference library, through the host one-shot entry
points and device plans.

Bar: rel-L2 <= 1e-12 per transform.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import ref
from impulse_amd import _lib
from impulse_amd import fft as F
from tests.conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12
KINDS = [("dct", t) for t in (1, 2, 3, 4)] + [("dst", t) for t in (1, 2, 3, 4)]


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().impulse_fft_device_check())
    yield
    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def golden_r2r():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_r2r.npz"))


def _ours(name):
    return F.dct if name == "dct" else F.dst


def _oracle(name):
    return oracle.dct if name == "dct" else oracle.dst


def _span(name, t, n):
    """Length of the c2c each line runs through (2(N-1), 2(N+1) or 2N)."""
    if t == 1:
        return 2 * (n - 1) if name == "dct" else 2 * (n + 1)
    return 2 * n


def test_kat_nim_example():
    # cpp_pocketfft/pocketfft.nim:322-337
    d_in = np.array([4.0, 3.0, 5.0, 10.0])
    d_out = np.zeros(4)
    F.apply(F.DCTDesc.init([0]), F.DataDesc.init(d_out, [4]), F.DataDesc.init(d_in, [4]))
    assert d_out[0] == 44.0
    assert rel_l2(d_out, oracle.dct(d_in, [0], 2)) < 1e-15


def test_golden_1d(golden_r2r):
    g = golden_r2r
    n_checked = 0
    for key in g.files:
        if not key.startswith(("dct", "dst")):
            continue
        name, t, n = key[:3], int(key[3]), int(key.split("_")[1])
        ortho = key.endswith("_ortho")
        y = _ours(name)(g[f"r2r_{n}_in"], [0], t, 0.5 if ortho else 1.0, ortho)
        assert rel_l2(y, g[key]) < TOL, key
        n_checked += 1
    assert n_checked > 150


def test_golden_nd(golden_r2r):
    g = golden_r2r
    v = g["nd_r2r_in"]
    for t in (1, 2, 3, 4):
        assert rel_l2(F.dct(v[:, ::2, :], [2, 0], t, 0.25, t == 2), g[f"nd_dct{t}_20"]) < TOL, t
        assert rel_l2(F.dst(v[::-1], [0, 1, 2], t, 1.0, t == 3), g[f"nd_dst{t}_012"]) < TOL, t


# even lengths with a 7-smooth half (fused kernel, inlined radix-2..8 stages: 6, 16, 64, 100, 128, 500,
# 1000, 1024, 4096) and with 11 or 13 in the half (fused kernel, per-radix stage calls: 22, 52, 286)
LENGTHS = [1, 2, 3, 4, 5, 6, 7, 9, 15, 16, 17, 22, 31, 33, 52, 63, 64, 65, 100, 127, 128, 129, 243, 255, 257,
           286, 500, 1000, 1023, 1024, 1025, 2047, 3001, 4096, 4097]


@pytest.mark.parametrize("name,t", KINDS)
def test_lengths_vs_oracle(name, t):
    rng = np.random.default_rng(100 + t + (10 if name == "dst" else 0))
    for n in LENGTHS:
        if name == "dct" and t == 1 and n == 1:
            continue
        x = rng.uniform(-1, 1, n)
        for ortho in (False, True):
            fct = 1.0 / np.sqrt(_span(name, t, n)) if ortho else 1.7
            y = _ours(name)(x, [0], t, fct, ortho)
            assert rel_l2(y, _oracle(name)(x, [0], t, fct, ortho)) < TOL, (name, t, n, ortho)


@pytest.mark.parametrize("name,t", KINDS)
def test_large_lengths_round_trip(name, t):
    """Lengths the O(N^2) oracle cannot reach: the unnormalised inverse pair brings x back times 2N'
    (I.I, II.III, III.II, IV.IV), and the reference library (when built) gives the same spectrum."""
    inv = {1: 1, 2: 3, 3: 2, 4: 4}[t]
    rng = np.random.default_rng(7 * t)
    for n in (65536, 100003, 131072 + 1, 1 << 20):
        x = rng.uniform(-1, 1, n)
        y = _ours(name)(x, [0], t)
        back = _ours(name)(y, [0], inv, 1.0 / _span(name, t, n))
        assert rel_l2(back, x) < TOL, (name, t, n)
        if ref.available() and n <= 131073:
            assert rel_l2(y, (ref.dct if name == "dct" else ref.dst)(x, [0], t)) < TOL, (name, t, n)


def test_nd_strided_and_negative_strides():
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, (7, 12, 10))
    views = [(a, [0, 1, 2]), (a[:, ::2, :], [2, 0]), (a[::-1, :, ::-3], [1]), (a.transpose(2, 0, 1), [0, 2]),
             (a[1:6, 3:11, 2:9], [2, 1, 0])]
    for name, t in KINDS:
        for i, (v, axes) in enumerate(views):
            ortho = bool(i % 2)
            y = _ours(name)(v, axes, t, 0.3, ortho)
            assert rel_l2(y, _oracle(name)(v, axes, t, 0.3, ortho)) < TOL, (name, t, i)
            # into a strided output view
            out = np.zeros(v.shape[:-1] + (2 * v.shape[-1],))[..., ::2]
            _ours(name)(v, axes, t, 0.3, ortho, out=out)
            assert rel_l2(out, _oracle(name)(v, axes, t, 0.3, ortho)) < TOL, (name, t, i)


def test_in_place_host():
    rng = np.random.default_rng(6)
    for name, t in KINDS:
        x = rng.uniform(-1, 1, (9, 40))
        want = _oracle(name)(x, [1, 0], t, 2.0)
        _ours(name)(x, [1, 0], t, 2.0, out=x)
        assert rel_l2(x, want) < TOL, (name, t)


def test_device_plans_torch():
    """DCTDesc on device tensors: a cached device plan on the current stream (contiguous, strided, in place)."""
    import torch
    rng = np.random.default_rng(8)
    for name, t in KINDS:
        sine = name == "dst"
        x = rng.uniform(-1, 1, (16, 33, 20))
        xd = torch.from_numpy(x).cuda()
        yd = torch.empty_like(xd)
        F.apply(F.DCTDesc.init([2, 0], t, t == 2, 0.5, sine=sine), F.DataDesc.init(yd, list(x.shape)),
                F.DataDesc.init(xd, list(x.shape)))
        torch.cuda.synchronize()
        assert rel_l2(yd.cpu().numpy(), _oracle(name)(x, [2, 0], t, 0.5, t == 2)) < TOL, (name, t)
        # a strided view (every other column) in and out, element strides as the Nim init takes them
        big = torch.from_numpy(rng.uniform(-1, 1, (16, 66))).cuda()
        ob = torch.zeros(16, 66, dtype=torch.float64, device="cuda")
        F.apply(F.DCTDesc.init([1], t, False, 1.0, sine=sine), F.DataDesc.init(ob, [16, 33], [66, 2]),
                F.DataDesc.init(big, [16, 33], [66, 2]))
        torch.cuda.synchronize()
        hv = big.cpu().numpy()[:, ::2]
        assert rel_l2(ob.cpu().numpy()[:, ::2], _oracle(name)(hv, [1], t)) < TOL, (name, t)
        assert torch.count_nonzero(ob[:, 1::2]) == 0
        # in place
        z = torch.from_numpy(x[3]).cuda().contiguous()
        want = _oracle(name)(x[3], [0, 1], t, 1.0, True)
        F.apply(F.DCTDesc.init([0, 1], t, True, 1.0, sine=sine), F.DataDesc.init(z, [33, 20]),
                F.DataDesc.init(z, [33, 20]))
        torch.cuda.synchronize()
        assert rel_l2(z.cpu().numpy(), want) < TOL, (name, t)


def test_c_api_device_plan_reuse():
    """One r2r plan executed repeatedly with different scale factors on the same stream."""
    import torch
    shape = [48, 250]
    st = [250 * 8, 8]
    p = F.Plan(_lib.DST, shape, st, st, [1], True, r2r_type=4)
    assert p.num_launches >= 1
    x = np.random.default_rng(9).uniform(-1, 1, shape)
    xd = torch.from_numpy(x).cuda()
    for fct in (1.0, 0.25, -3.0):
        yd = torch.empty_like(xd)
        p.execute(xd.data_ptr(), yd.data_ptr(), fct, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert rel_l2(yd.cpu().numpy(), oracle.dst(x, [1], 4, fct)) < TOL, fct


def test_zero_sized_and_errors():
    # an empty axis is a no-op, as in pocketfft (the reference returns before planning)
    e = np.zeros((0, 5))
    assert F.dct(e, [1], 2).shape == (0, 5)
    with pytest.raises(Exception):
        F.dct(np.zeros(1), [0], 1)                       # DCT-I needs N >= 2
    with pytest.raises(Exception):
        F.dct(np.zeros(8), [0], 5)                       # type out of range
    with pytest.raises(Exception):
        F.dst(np.zeros((4, 4)), [2], 2)                  # axis out of range
    with pytest.raises(Exception):
        F.dst(np.zeros((4, 4)), [1, 1], 2)               # axis repeated
    with pytest.raises(ValueError):
        F.DCTDesc.init([0], 0)
    with pytest.raises(TypeError):
        F.apply(F.DCTDesc.init([0]), F.DataDesc.init(np.zeros(4, complex), [4]), F.DataDesc.init(np.zeros(4), [4]))

"""GPU parity of r2r_fftpack and the separable / genuine Hartley transforms
(pocketfft_hdronly.h:3392-3445) against the golden vectors the reference's own
PocketFFT wrote, the oracle's long-double sums and, where it was built, the
reference library; through the host one-shot entry points and device plans.

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
FLAGS = [(r2h, fw) for r2h in (True, False) for fw in (True, False)]


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


def test_golden(golden_r2r):
    g = golden_r2r
    n_checked = 0
    for key in g.files:
        if key.startswith("fp") and key[2:4].isdigit():
            r2h, fw, n = int(key[2]), int(key[3]), int(key.split("_")[1])
            y = F.r2r_fftpack(g[f"fp_{n}_in"], [0], r2h, fw, 1.0 if fw else 0.5)
            assert rel_l2(y, g[key]) < TOL, key
            n_checked += 1
        elif key.startswith("hartley_"):
            n = int(key.split("_")[1])
            assert rel_l2(F.r2r_separable_hartley(g[f"fp_{n}_in"], [0]), g[key]) < TOL, key
            n_checked += 1
    assert n_checked == 5 * 18
    v = g["nd_r2r_in"]
    assert rel_l2(F.r2r_fftpack(v[:, ::2, :], [2, 0], True, True, 0.25), g["nd_fp11_20"]) < TOL
    assert rel_l2(F.r2r_fftpack(v[::-1], [1], False, False), g["nd_fp00_1"]) < TOL
    assert rel_l2(F.r2r_fftpack(v, [0, 2], False, True), g["nd_fp01_02"]) < TOL
    assert rel_l2(F.r2r_separable_hartley(v, [0, 1, 2], 0.5), g["nd_hartley_012"]) < TOL
    assert rel_l2(F.r2r_genuine_hartley(v, [1, 2]), g["nd_genuine_12"]) < TOL
    assert rel_l2(F.r2r_genuine_hartley(v[:, ::-1, :], [2, 0, 1], 2.0), g["nd_genuine_201"]) < TOL


# even lengths take the fused kernel: a 7-smooth half (inlined radix-2..8 stages) or one with 11 or 13
# (per-radix stage calls: 22, 52, 286)
LENGTHS = [1, 2, 3, 4, 5, 6, 7, 9, 16, 17, 22, 31, 52, 64, 97, 100, 127, 256, 286, 1000, 1009, 4096, 4097]


@pytest.mark.parametrize("r2h,fw", FLAGS)
def test_fftpack_lengths_vs_oracle(r2h, fw):
    rng = np.random.default_rng(40 + 2 * r2h + fw)
    for n in LENGTHS:
        x = rng.uniform(-1, 1, (3, n))
        y = F.r2r_fftpack(x, [1], r2h, fw, 0.5)
        assert rel_l2(y, oracle.r2r_fftpack(x, [1], r2h, fw, 0.5)) < TOL, (n, r2h, fw)


def test_hartley_lengths_vs_oracle():
    rng = np.random.default_rng(41)
    for n in LENGTHS:
        x = rng.uniform(-1, 1, (3, n))
        assert rel_l2(F.r2r_separable_hartley(x, [1], 1.5), oracle.r2r_separable_hartley(x, [1], 1.5)) < TOL, n
        assert rel_l2(F.r2r_genuine_hartley(x, [1]), oracle.r2r_genuine_hartley(x, [1])) < TOL, n


def test_large_round_trips():
    """Lengths past the O(N^2) oracle: forward then backward returns N x, the separable Hartley is its
    own inverse up to N, and the reference library (when built) gives the same values."""
    rng = np.random.default_rng(42)
    for n in (65536, 100003, 1 << 20):
        x = rng.uniform(-1, 1, n)
        hc = F.r2r_fftpack(x, [0], True, True)
        assert rel_l2(F.r2r_fftpack(hc, [0], False, False, 1.0 / n), x) < TOL, n
        h = F.r2r_separable_hartley(x, [0])
        assert rel_l2(F.r2r_separable_hartley(h, [0], 1.0 / n), x) < TOL, n
        if ref.available() and n <= 100003:
            assert rel_l2(hc, ref.r2r_fftpack(x, [0], True, True)) < TOL, n
            assert rel_l2(h, ref.r2r_separable_hartley(x, [0])) < TOL, n


def test_nd_views_and_in_place():
    rng = np.random.default_rng(43)
    a = rng.uniform(-1, 1, (7, 12, 10))
    views = [(a, [0, 1, 2]), (a[:, ::2, :], [2, 0]), (a[::-1, :, ::-3], [1]), (a.transpose(2, 0, 1), [0, 2])]
    for i, (v, axes) in enumerate(views):
        for r2h, fw in FLAGS:
            assert rel_l2(F.r2r_fftpack(v, axes, r2h, fw, 0.3), oracle.r2r_fftpack(v, axes, r2h, fw, 0.3)) < TOL, \
                (i, r2h, fw)
        assert rel_l2(F.r2r_separable_hartley(v, axes), oracle.r2r_separable_hartley(v, axes)) < TOL, i
        assert rel_l2(F.r2r_genuine_hartley(v, axes, 0.7), oracle.r2r_genuine_hartley(v, axes, 0.7)) < TOL, i
        out = np.zeros(v.shape[:-1] + (2 * v.shape[-1],))[..., ::2]
        F.r2r_genuine_hartley(v, axes, 0.7, out=out)
        assert rel_l2(out, oracle.r2r_genuine_hartley(v, axes, 0.7)) < TOL, i
    for r2h, fw in FLAGS:
        x = rng.uniform(-1, 1, (9, 40))
        want = oracle.r2r_fftpack(x, [1, 0], r2h, fw, 2.0)
        F.r2r_fftpack(x, [1, 0], r2h, fw, 2.0, out=x)
        assert rel_l2(x, want) < TOL, (r2h, fw)
    x = rng.uniform(-1, 1, (9, 40))
    want = oracle.r2r_genuine_hartley(x, [0, 1])
    F.r2r_genuine_hartley(x, [0, 1], out=x)
    assert rel_l2(x, want) < TOL


def test_device_plans():
    import torch
    rng = np.random.default_rng(44)
    x = rng.uniform(-1, 1, (16, 33, 20))
    xd = torch.from_numpy(x).cuda()
    st = [33 * 20 * 8, 20 * 8, 8]
    sp = torch.cuda.current_stream().cuda_stream
    cases = [(_lib.FFTPACK, r2h, fw) for r2h, fw in FLAGS] + [(_lib.HARTLEY, 1, 1), (_lib.HARTLEY_GENUINE, 1, 1)]
    for kind, r2h, fw in cases:
        p = F.FftpackPlan(kind, [16, 33, 20], st, st, [2, 0], r2h, fw)
        assert p.num_launches >= 2
        yd = torch.empty_like(xd)
        p.execute(xd.data_ptr(), yd.data_ptr(), 0.5, sp)
        torch.cuda.synchronize()
        if kind == _lib.FFTPACK:
            want = oracle.r2r_fftpack(x, [2, 0], r2h, fw, 0.5)
        elif kind == _lib.HARTLEY:
            want = oracle.r2r_separable_hartley(x, [2, 0], 0.5)
        else:
            want = oracle.r2r_genuine_hartley(x, [2, 0], 0.5)
        assert rel_l2(yd.cpu().numpy(), want) < TOL, (kind, r2h, fw)


def test_errors():
    assert F.r2r_fftpack(np.zeros((0, 4)), [1], True, True).shape == (0, 4)
    with pytest.raises(Exception):
        F.r2r_fftpack(np.zeros((4, 4)), [2], True, True)          # axis out of range
    with pytest.raises(Exception):
        F.r2r_genuine_hartley(np.zeros((4, 4)), [1, 1])           # axis repeated
    with pytest.raises(TypeError):
        F.r2r_separable_hartley(np.zeros(4, np.float32), [0])

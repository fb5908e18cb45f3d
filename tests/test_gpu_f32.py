"""GPU parity of the float32 drop-in (IMPULSE_FFT_F32: the reference's T = float instantiation,
pocketfft::c2c / r2c / c2r<float>, pocketfft_hdronly.h:3272-3390, reached from impulse through
FFTDesc[float32] / DataDesc[float32], cpp_pocketfft/pocketfft.nim:35-69,137-147).

The engine widens float inputs to FP64, computes in FP64 and rounds each output component to the
nearest float.  Every output component then differs from the exact DFT of the widened input by at
most 2^-24 of its own magnitude plus the FP64 error, so the rel-L2 against the oracle (long double on
the same widened input) is bounded by F32_TOL = 2^-24 + 2e-15.  Against the reference's own float
results (tests/golden/golden_f32.npz) the bar is 5e-7, the reference's own float rounding.
"""
import os

import numpy as np
import pytest

import oracle
from impulse_amd import _lib, synth
from impulse_amd import fft as F
from tests.conftest import ROOT, rel_l2
from tests.test_gpu_parity import LENGTHS

pytestmark = pytest.mark.gpu
F32_TOL = 2.0 ** -24 + 2e-15
REF_TOL = 5e-7


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().impulse_fft_device_check())
    yield
    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def g32():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_f32.npz"))


def _w(a):
    """Widen to the FP64 dtype of the same kind (exact)."""
    return a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)


def test_golden_f32_1d(g32):
    ns = sorted({int(k.split("_")[1]) for k in g32.files if k.startswith("c2c_") and k.endswith("_in")})
    for n in ns:
        print(f"case f32 golden n={n}", flush=True)
        x = g32[f"c2c_{n}_in"]
        y = F.c2c(x, [0], True)
        assert y.dtype == np.complex64
        assert rel_l2(y, g32[f"c2c_{n}_fwd"]) < REF_TOL, n
        assert rel_l2(y, oracle.c2c(_w(x), [0], True)) <= F32_TOL, n
        assert rel_l2(F.c2c(x, [0], False, 0.5), oracle.c2c(_w(x), [0], False, 0.5)) <= F32_TOL, n
        assert rel_l2(F.c2c(x, [0], False, 0.5), g32[f"c2c_{n}_bwd"]) < REF_TOL, n
        r = g32[f"r2c_{n}_in"]
        Y = F.r2c(r, [0], True)
        assert Y.dtype == np.complex64
        assert rel_l2(Y, g32[f"r2c_{n}_out"]) < REF_TOL, n
        assert rel_l2(Y, oracle.r2c(_w(r), [0], True)) <= F32_TOL, n
        assert rel_l2(F.r2c(r, [0], False, 2.0), g32[f"r2c_{n}_bwd"]) < REF_TOL, n
        c = g32[f"c2r_{n}_in"]
        z = F.c2r(c, n, [0], False, 1.0 / n)
        assert z.dtype == np.float32
        assert rel_l2(z, g32[f"c2r_{n}_out"]) < REF_TOL, n
        assert rel_l2(z, oracle.c2r(_w(c), n, [0], False, 1.0 / n)) <= F32_TOL, n


def test_golden_f32_nd(g32):
    a = g32["nd_c2c_in"]
    assert rel_l2(F.c2c(a[:, ::2], [0, 1], True), g32["nd_c2c_01"]) < REF_TOL
    assert rel_l2(F.c2c(a[::-1, :], [1, 0], False, 0.25), g32["nd_c2c_10_bwd"]) < REF_TOL
    assert rel_l2(F.c2c(a[::-1, :], [1, 0], False, 0.25), oracle.c2c(_w(a)[::-1, :], [1, 0], False, 0.25)) <= F32_TOL
    b = g32["nd_r2c_in"]
    assert rel_l2(F.r2c(b, [0, 1, 2]), g32["nd_r2c_012"]) < REF_TOL
    assert rel_l2(F.r2c(b, [2, 0]), g32["nd_r2c_20"]) < REF_TOL
    assert rel_l2(F.r2c(b, [2, 0]), oracle.r2c(_w(b), [2, 0])) <= F32_TOL
    h = g32["nd_c2r_in"]
    assert rel_l2(F.c2r(h, 9, [0, 1, 2], False, 1.0 / 540), g32["nd_c2r_012"]) < REF_TOL
    assert rel_l2(F.c2r(h, 9, [0, 1, 2], False, 1.0 / 540), oracle.c2r(_w(h), 9, [0, 1, 2], False, 1.0 / 540)) \
        <= F32_TOL
    for n in (4096, 4095):
        assert rel_l2(F.r2c(g32[f"rows_{n}_in"], [1]), g32[f"rows_{n}_out"]) < REF_TOL


# every algorithm branch (single pass, LDS Bluestein, four-step, four-step Bluestein), as for FP64
@pytest.mark.parametrize("n", LENGTHS)
def test_f32_lengths_vs_oracle(n):
    print(f"case f32 n={n}", flush=True)
    batch = 3
    x = synth.complex_normal((batch, n), seed=n).astype(np.complex64)
    for fwd in (True, False):
        assert rel_l2(F.c2c(x, [1], fwd, 0.5), oracle.c2c(_w(x), [1], fwd, 0.5)) <= F32_TOL, ("c2c", n, fwd)
    r = synth.uniform((batch, n), -1, 1, seed=n + 1).astype(np.float32)
    for fwd in (True, False):
        assert rel_l2(F.r2c(r, [1], fwd), oracle.r2c(_w(r), [1], fwd)) <= F32_TOL, ("r2c", n, fwd)
    h = synth.complex_normal((batch, n // 2 + 1), seed=n + 2).astype(np.complex64)
    for fwd in (True, False):
        assert rel_l2(F.c2r(h, n, [1], fwd, 1.0 / n), oracle.c2r(_w(h), n, [1], fwd, 1.0 / n)) <= F32_TOL, \
            ("c2r", n, fwd)


@pytest.mark.parametrize("n", [4096, 4095])
def test_f32_c1_rows_round_trip(n):
    """[256, n] float32 r2c along axis 1 and back (the C1 workload in the reference's T = float)."""
    print(f"case f32 C1 n={n}", flush=True)
    x = synth.uniform((256, n), -1.0, 1.0, seed=0).astype(np.float32)
    X = F.r2c(x, [1], True)
    assert rel_l2(X, oracle.r2c(_w(x), [1], True)) <= F32_TOL
    back = F.c2r(X, n, [1], False, 1.0 / n)
    assert rel_l2(back, x) < 1e-6


def test_f32_strided_axes_and_views():
    x = synth.complex_normal((24, 40, 6), seed=3).astype(np.complex64)
    for view, axes in ((x[:, ::2, :], [1]), (x[::-1], [0, 2]), (x[:, :, ::-1], [2, 0, 1]), (x.transpose(2, 0, 1), [1])):
        assert rel_l2(F.c2c(view, axes, True), oracle.c2c(_w(view), axes, True)) <= F32_TOL, axes
    r = synth.uniform((30, 18, 10), -1, 1, seed=4).astype(np.float32)
    for view, axes in ((r[::-1], [0]), (r[:, ::2, :], [1, 2]), (r.transpose(1, 2, 0), [0, 2])):
        assert rel_l2(F.r2c(view, axes), oracle.r2c(_w(view), axes)) <= F32_TOL, axes


def test_f32_long_lengths_device():
    """Four-step (with the fused real post-process), four-step Bluestein and the real-half load on
    torch device tensors; an odd float offset takes the per-component load instead of the paired one."""
    import torch
    cases = [("r2c", 1 << 20, 2), ("c2c", 65536, 3), ("r2c", 100003, 2), ("c2c", 20014, 2), ("c2r", 131072, 2)]
    for kind, n, b in cases:
        print(f"case f32 device {kind} n={n} batch={b}", flush=True)
        if kind == "c2c":
            x = synth.complex_normal((b, n), seed=n).astype(np.complex64)
            xd = torch.from_numpy(x).cuda()
            yd = torch.empty_like(xd)
            F.apply(F.FFTDesc([1], True), F.DataDesc(yd, yd.shape), F.DataDesc(xd, xd.shape))
            want = oracle.c2c(_w(x), [1], True)
        elif kind == "r2c":
            x = synth.uniform((b, n), -1, 1, seed=n).astype(np.float32)
            xd = torch.from_numpy(x).cuda()
            yd = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device="cuda")
            F.apply(F.FFTDesc([1], True), F.DataDesc(yd, yd.shape), F.DataDesc(xd, xd.shape))
            want = oracle.r2c(_w(x), [1], True)
        else:
            x = synth.complex_normal((b, n // 2 + 1), seed=n).astype(np.complex64)
            xd = torch.from_numpy(x).cuda()
            yd = torch.empty((b, n), dtype=torch.float32, device="cuda")
            F.apply(F.FFTDesc([1], False, 1.0 / n), F.DataDesc(yd, yd.shape), F.DataDesc(xd, xd.shape))
            want = oracle.c2r(_w(x), n, [1], False, 1.0 / n)
        torch.cuda.synchronize()
        assert rel_l2(yd.cpu().numpy(), want) <= F32_TOL, (kind, n)
    # even real rows at a 4-byte (not 8-byte) aligned base: one float in
    n = 4096
    x = synth.uniform((3 * n + 1,), -1, 1, seed=9).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    for off in (0, 1):
        view = xd[off:off + 3 * n].view(3, n)
        yd = torch.empty((3, n // 2 + 1), dtype=torch.complex64, device="cuda")
        F.apply(F.FFTDesc([1], True), F.DataDesc(yd, yd.shape), F.DataDesc(view, view.shape))
        torch.cuda.synchronize()
        assert rel_l2(yd.cpu().numpy(), oracle.r2c(_w(x[off:off + 3 * n].reshape(3, n)), [1])) <= F32_TOL, off


def test_f32_plan_bytes_and_errors():
    """Per-launch byte figures count float arrays at 4 bytes; f32 pointers must be 4 / 8-byte aligned."""
    import ctypes
    L = _lib.lib()
    p = F.Plan(_lib.C2C, [64, 4096], [4096 * 8, 8], [4096 * 8, 8], [1], True, dtype=_lib.F32)
    bi, bo = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(L.impulse_fft_plan_pass_info(p.handle, 0, None, 0, ctypes.byref(bi), ctypes.byref(bo), None, None,
                                            None))
    assert bi.value == bo.value == 64 * 4096 * 8
    assert L.impulse_fft_plan_create(ctypes.byref(ctypes.c_void_p()), _lib.C2C, 7, 1, _lib.sizes([8]),
                                     _lib.strides([8]), _lib.strides([8]), 1, _lib.sizes([0]), 1) == 1
    # strides that are not multiples of the float element size
    assert L.impulse_fft_plan_create(ctypes.byref(ctypes.c_void_p()), _lib.R2C, _lib.F32, 1, _lib.sizes([8]),
                                     _lib.strides([6]), _lib.strides([8]), 1, _lib.sizes([0]), 1) == 4


def test_f32_autograd_adjoints():
    """<A x, y> = <x, A* y> for r2c / c2r / c2c on float32 / complex64 (float rounding bar)."""
    import torch
    from impulse_amd import autograd as AG
    torch.manual_seed(0)
    for n in (64, 4095, 10007):
        x = torch.randn(3, n, dtype=torch.float32, device="cuda", requires_grad=True)
        Y = AG.rfftn(x, [1])
        G = torch.randn_like(Y)
        (gx,) = torch.autograd.grad((Y.conj() * G).real.sum(), x)
        # <A x, G> over the half spectrum with interior weights from the r2c adjoint
        lhs = (Y.detach().conj() * G).real.double().sum()
        rhs = (x.detach().double() * gx.double()).sum()
        assert abs(float(lhs - rhs)) <= 1e-5 * float(x.detach().double().norm() * gx.double().norm()), n
        Z = torch.randn(3, n // 2 + 1, dtype=torch.complex64, device="cuda", requires_grad=True)
        z = AG.irfftn(Z, n, [1])
        w = torch.randn_like(z)
        (gZ,) = torch.autograd.grad((z * w).sum(), Z)
        lhs = (z.detach().double() * w.double()).sum()
        rhs = (Z.detach().conj() * gZ).real.double().sum()
        assert abs(float(lhs - rhs)) <= 1e-5 * float(Z.detach().abs().double().norm() * gZ.abs().double().norm()), n


@pytest.mark.parametrize("kind,batch,w4", [("c2c", 2, True), ("c2c", 512, False), ("r2c", 3, True),
                                           ("r2c", 1024, False)])
def test_f32_bluestein_pass_instances(kind, batch, w4):
    """The three-pass Bluestein of n = 16381 on float arrays: the chirp-in pass reads the float
    (or complex float) input directly (IO code 5 + 16, + 8 on the four-wave instances), and for c2c
    the chirp-out pass writes the complex float output directly (6 + 32); r2c writes the workspace
    and its real post-process rounds to float.  First and last signals against the oracle."""
    import ctypes
    import torch
    n = 16381
    print(f"case f32 Bluestein pass instances {kind} [{batch},{n}]", flush=True)
    sp = torch.cuda.current_stream().cuda_stream
    if kind == "c2c":
        x = synth.complex_normal((batch, n), seed=batch).astype(np.complex64)
        yd = torch.empty((batch, n), dtype=torch.complex64, device="cuda")
        lk = _lib.C2C
    else:
        x = synth.uniform((batch, n), -1, 1, seed=batch).astype(np.float32)
        yd = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device="cuda")
        lk = _lib.R2C
    xd = torch.from_numpy(x).cuda()
    st_in = [s * xd.element_size() for s in xd.stride()]
    st_out = [s * yd.element_size() for s in yd.stride()]
    for fwd in (True, False):
        pl = F.Plan(lk, [batch, n], st_in, st_out, [1], fwd, dtype=_lib.F32)
        pl.execute(xd.data_ptr(), yd.data_ptr(), 0.5, sp)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        for row in (0, batch - 1):
            xr = _w(x[row])
            ref = oracle.c2c(xr, [0], fwd, 0.5) if kind == "c2c" else oracle.r2c(xr, [0], fwd, 0.5)
            assert rel_l2(y[row], ref) <= F32_TOL, (kind, batch, fwd, row)
        names = []
        for k in range(pl.num_launches):
            name = ctypes.create_string_buffer(64)
            z, u = ctypes.c_size_t(), ctypes.c_uint()
            _lib.check(_lib.lib().impulse_fft_plan_pass_info(pl.handle, k, name, 64, ctypes.byref(z), ctypes.byref(z),
                                                          ctypes.byref(u), ctypes.byref(u), ctypes.byref(z)))
            names.append(name.value.decode())
        print(kind, batch, fwd, names)
        io = [int(s.rsplit(",", 1)[1].rstrip(">")) for s in names if s.startswith("k_fft_pow2")]
        want = [21, 4, 38] if kind == "c2c" else [21, 4, 6]
        assert io[:3] == ([c + 8 for c in want] if w4 else want), names

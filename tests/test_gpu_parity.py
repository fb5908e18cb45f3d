"""GPU parity of the HIP engine against the oracle and the reference's golden
vectors, through the C ABI (host-pointer and device-pointer entry points).

Bars: rel-L2 <= 1e-12 against the FP64 oracle / PocketFFT golden vectors,
round trips <= 1e-10, the reference's exact known answers bit-exact.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from impulse_amd import _lib, synth
from impulse_amd import fft as F
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().impulse_fft_device_check())
    yield
    torch.cuda.synchronize()


def _torch():
    import torch
    return torch


# ---------------------------------------------------------------------------
# known answers (tests/test_fft2.nim:7-15, README.md:41-98)
# ---------------------------------------------------------------------------

def test_kat_fft2_exact():
    exp = np.array([6, -2 + 2j, -2, -2 - 2j])
    x = np.arange(4.0)
    assert np.array_equal(F.fft(x.astype(np.complex128)), exp)          # complex input
    assert np.array_equal(F.fft(x), exp)                               # real input
    assert np.array_equal(F.fft(x, forward=True), exp)
    assert np.array_equal(F.ifft(F.fft(x)).real, x)
    assert np.array_equal(F.ifft(F.fft(x)).imag, np.zeros(4))
    assert np.array_equal(F.fft(x, normalize=F.nkForward), exp / 4)
    assert np.array_equal(F.fft(x, normalize=F.nkOrtho), exp / 2)
    assert np.array_equal(F.ifft(x), F.fft(x, forward=False))


def test_kat_readme():
    y = F.fft(np.array([1.0, 2.0, 1.0, -1.0, 1.5]))
    exp = np.array([4.5, 2.081559480312316 - 1.651098762732523j, -1.831559480312316 + 1.608220406444071j,
                    -1.831559480312316 - 1.608220406444071j, 2.081559480312316 + 1.651098762732523j])
    assert np.max(np.abs(y - exp)) < 1e-15
    # C++ path r2c into a full-length complex buffer leaves the tail untouched (README.md:98)
    d_in = np.array([1.0, 2.0, 1.0, -1.0, 1.5])
    d_out = np.zeros(5, np.complex128)
    F.apply(F.FFTDesc([0], True), F.DataDesc(d_out, [5]), F.DataDesc(d_in, [5]))
    assert np.max(np.abs(d_out[:3] - exp[:3])) < 1e-15
    assert np.array_equal(d_out[3:], np.zeros(2))


# ---------------------------------------------------------------------------
# golden vectors written by the reference's own PocketFFT
# ---------------------------------------------------------------------------

def test_golden_1d(golden):
    lengths = sorted({int(k.split("_")[1]) for k in golden.files if k.startswith("c2c_") and k.endswith("_in")})
    for n in lengths:
        x = golden[f"c2c_{n}_in"]
        assert rel_l2(F.c2c(x, [0], True), golden[f"c2c_{n}_fwd"]) < TOL, n
        r = golden[f"r2c_{n}_in"]
        assert rel_l2(F.r2c(r, [0], True), golden[f"r2c_{n}_out"]) < TOL, n
        assert rel_l2(F.rfft_packed(r, True), golden[f"hc_{n}_out"]) < TOL, n
        if f"c2c_{n}_bwd" in golden.files:
            assert rel_l2(F.c2c(x, [0], False, 0.5), golden[f"c2c_{n}_bwd"]) < TOL, n
            assert rel_l2(F.r2c(r, [0], False, 2.0), golden[f"r2c_{n}_bwd"]) < TOL, n
            c = golden[f"c2r_{n}_in"]
            assert rel_l2(F.c2r(c, n, [0], False, 1.0 / n), golden[f"c2r_{n}_out"]) < TOL, n
            assert rel_l2(F.c2r(c, n, [0], True), golden[f"c2r_{n}_fwdflag"]) < TOL, n


def test_golden_nd(golden):
    a = golden["nd_c2c_in"]
    assert rel_l2(F.c2c(a[:, ::2], [0, 1], True), golden["nd_c2c_01"]) < TOL
    assert rel_l2(F.c2c(a[:, ::2], [1, 0], False, 0.25), golden["nd_c2c_10_bwd"]) < TOL
    b = golden["nd_r2c_in"]
    assert rel_l2(F.r2c(b, [0, 1, 2]), golden["nd_r2c_012"]) < TOL
    assert rel_l2(F.r2c(b, [2, 0]), golden["nd_r2c_20"]) < TOL
    assert rel_l2(F.r2c(b, [1], False), golden["nd_r2c_1_bwd"]) < TOL
    h = golden["nd_c2r_in"]
    assert rel_l2(F.c2r(h, 9, [0, 1, 2], False, 1.0 / 540), golden["nd_c2r_012"]) < TOL
    assert rel_l2(F.c2r(h, 8, [2], False), golden["nd_c2r_2"]) < TOL
    for n in (4096, 4095):
        assert rel_l2(F.r2c(golden[f"rows_{n}_in"], [1]), golden[f"rows_{n}_out"]) < TOL


# ---------------------------------------------------------------------------
# every algorithm branch against the oracle
# ---------------------------------------------------------------------------

# single pass (smooth <= 8192), LDS Bluestein (n2 <= 8192), four-step, four-step Bluestein
LENGTHS = (list(range(1, 130)) + [143, 169, 243, 360, 1000, 1001, 1024, 2017, 2053, 4095, 4096, 6561, 8192,
                                  8190, 10000, 4099, 10007, 16384, 30030, 65536, 20014, 131072])


@pytest.mark.parametrize("n", LENGTHS)
def test_lengths_vs_oracle(n):
    print(f"case n={n}", flush=True)
    batch = 3
    x = synth.complex_normal((batch, n), seed=n)
    for fwd in (True, False):
        assert rel_l2(F.c2c(x, [1], fwd, 0.5), oracle.c2c(x, [1], fwd, 0.5)) < TOL, ("c2c", n, fwd)
    r = synth.uniform((batch, n), -1, 1, seed=n + 1)
    for fwd in (True, False):
        assert rel_l2(F.r2c(r, [1], fwd), oracle.r2c(r, [1], fwd)) < TOL, ("r2c", n, fwd)
    h = synth.complex_normal((batch, n // 2 + 1), seed=n + 2)
    for fwd in (False, True):
        assert rel_l2(F.c2r(h, n, [1], fwd, 1.0 / n), oracle.c2r(h, n, [1], fwd, 1.0 / n)) < TOL, ("c2r", n, fwd)


@pytest.mark.parametrize("n", [6, 64, 97, 4095, 4096, 10007, 16381, 65536])
def test_strided_axis_vs_oracle(n):
    # transform along axis 0 of [n, 5] (batch contiguous, axis strided) and of a
    # negative-stride view
    print(f"case strided n={n}", flush=True)
    x = synth.complex_normal((n, 5), seed=n)
    assert rel_l2(F.c2c(x, [0]), oracle.c2c(x, [0])) < TOL
    v = x[::-1, ::2]
    assert rel_l2(F.c2c(v, [0], False), oracle.c2c(v, [0], False)) < TOL
    r = synth.uniform((n, 4), -1, 1, seed=n)
    assert rel_l2(F.r2c(r, [0]), oracle.r2c(r, [0])) < TOL
    h = synth.complex_normal((n // 2 + 1, 3), seed=n)
    assert rel_l2(F.c2r(h, n, [0], False), oracle.c2r(h, n, [0], False)) < TOL


def test_nd_mixed():
    x = synth.complex_normal((6, 35, 16), seed=3)
    for axes in ([0], [1, 2], [2, 0, 1], [0, 2]):
        assert rel_l2(F.c2c(x, axes), oracle.c2c(x, axes)) < TOL, axes
    r = synth.uniform((5, 12, 9), -1, 1, seed=4)
    for axes in ([2], [0, 2], [2, 1, 0], [1]):
        assert rel_l2(F.r2c(r, axes, False, 3.0), oracle.r2c(r, axes, False, 3.0)) < TOL, axes
    h = synth.complex_normal((5, 12, 5), seed=5)
    for axes in ([2], [0, 2], [1, 0, 2]):
        assert rel_l2(F.c2r(h, 9, axes, False), oracle.c2r(h, 9, axes, False)) < TOL, axes
    # in place c2c on host memory
    y = x.copy()
    F.c2c(y, [1], True, 1.0, out=y)
    assert rel_l2(y, oracle.c2c(x, [1])) < TOL


# ---------------------------------------------------------------------------
# the reference's C-path round trip (tests/test_fft.nim:29-49) over every length
# ---------------------------------------------------------------------------

def test_c_path_round_trip_all_lengths():
    maxlen = 8192
    odata = synth.uniform((maxlen,), -0.5, 0.5, seed=0)
    odata[0] = 0.340188
    worst, worst_n = 0.0, 0
    for length in range(1, maxlen):
        data = odata[:length].copy()
        plan = F.FFTPlanReal.init(length)
        plan.call(data, True, 1.0)
        plan.call(data, False, 1.0 / length)
        err = rel_l2(data, odata[:length])
        if err > worst:
            worst, worst_n = err, length
    print(f"C-path round trip: worst rel-L2 {worst:.3e} at length {worst_n}")
    assert worst <= 1e-13, (worst, worst_n)


def test_c_path_halfcomplex_vs_oracle():
    for n in (1, 2, 3, 7, 16, 97, 4095, 4096, 10007):
        x = synth.uniform((n,), -1, 1, seed=n)
        for fwd, fct in ((True, 1.0), (False, 0.5)):
            assert rel_l2(F.rfft_packed(x, fwd, F.nkCustom, fct), oracle.rfft_packed(x, fwd, fct)) < TOL, n
        z = synth.complex_normal((n,), seed=n)
        for fwd in (True, False):
            c = z.copy()
            F.fft_inplace(c, fwd, F.nkCustom, 0.25)
            assert rel_l2(c, oracle.cfft(z, fwd, 0.25)) < TOL, n


# ---------------------------------------------------------------------------
# device pointers (C++ path over torch tensors) and the full-size workloads
# ---------------------------------------------------------------------------

def _dev(a):
    return _torch().from_numpy(np.ascontiguousarray(a)).cuda()


def test_device_r2c_rows_4096_4095():
    torch = _torch()
    for n in (4096, 4095):
        print(f"case rows [256,{n}]", flush=True)
        x = synth.uniform((256, n), -1.0, 1.0, seed=0)
        xd = _dev(x)
        yd = torch.empty((256, n // 2 + 1), dtype=torch.complex128, device="cuda")
        F.apply(F.FFTDesc([1], True), F.DataDesc(yd, yd.shape), F.DataDesc(xd, xd.shape))
        zd = torch.empty_like(xd)
        F.apply(F.FFTDesc([1], False, 1.0 / n), F.DataDesc(zd, zd.shape), F.DataDesc(yd, yd.shape))
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        ref = oracle.r2c(x, [1])
        errs = [rel_l2(y[i], ref[i]) for i in range(256)]
        print(f"[256,{n}] r2c rel-L2 max {max(errs):.3e} median {np.median(errs):.3e}")
        assert max(errs) <= TOL
        assert rel_l2(zd.cpu().numpy(), x) <= 1e-10


def test_device_nondefault_stream():
    torch = _torch()
    x = synth.complex_normal((64, 1024), seed=1)
    xd, yd = _dev(x), torch.empty((64, 1024), dtype=torch.complex128, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        F.apply(F.FFTDesc([1], True), F.DataDesc(yd, yd.shape), F.DataDesc(xd, xd.shape))
    s.synchronize()
    assert rel_l2(yd.cpu().numpy(), oracle.c2c(x, [1])) < TOL


def test_c2c_4096x4096_axes01():
    torch = _torch()
    n = 4096
    print("case c2c 4096x4096", flush=True)
    # separability: the 2-D DFT of an outer product u v^T is (F u)(F v)^T
    u = synth.complex_normal((n,), seed=1)
    v = synth.complex_normal((n,), seed=2)
    xd = torch.outer(_dev(u), _dev(v))
    yd = torch.empty_like(xd)
    F.apply(F.FFTDesc([0, 1], True), F.DataDesc(yd, yd.shape), F.DataDesc(xd, xd.shape))
    torch.cuda.synchronize()
    Fu, Fv = oracle.c2c(u, [0]), oracle.c2c(v, [0])
    ref = torch.outer(_dev(Fu), _dev(Fv))
    err = (torch.linalg.vector_norm(yd - ref) / torch.linalg.vector_norm(ref)).item()
    print(f"4096^2 separable rel-L2 {err:.3e}")
    assert err <= TOL
    # random input: exact DFT sums at sampled bins, then the round trip
    x = synth.complex_normal((n, n), seed=0)
    xd = _dev(x)
    F.apply(F.FFTDesc([0, 1], True), F.DataDesc(yd, yd.shape), F.DataDesc(xd, xd.shape))
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    k = np.arange(n)
    rng = np.random.default_rng(0)
    for _ in range(4):
        k1, k2 = rng.integers(0, n, 2)
        w1 = np.exp(-2j * np.pi * ((k1 * k) % n) / n)
        w2 = np.exp(-2j * np.pi * ((k2 * k) % n) / n)
        direct = w1 @ x @ w2
        assert abs(y[k1, k2] - direct) <= 1e-12 * np.linalg.norm(y) / n, (k1, k2)
    zd = torch.empty_like(xd)
    F.apply(F.FFTDesc([0, 1], False, 1.0 / (n * n)), F.DataDesc(zd, zd.shape), F.DataDesc(yd, yd.shape))
    torch.cuda.synchronize()
    rt = (torch.linalg.vector_norm(zd - xd) / torch.linalg.vector_norm(xd)).item()
    print(f"4096^2 round trip rel-L2 {rt:.3e}")
    assert rt <= 1e-12


@pytest.mark.parametrize("n,batch", [(1 << 22, 16), (1000003, 4)])
def test_r2c_long(n, batch):
    torch = _torch()
    print(f"case r2c [{batch},{n}]", flush=True)
    x = synth.uniform((batch, n), -1.0, 1.0, seed=0)
    xd = _dev(x)
    yd = torch.empty((batch, n // 2 + 1), dtype=torch.complex128, device="cuda")
    F.apply(F.FFTDesc([1], True), F.DataDesc(yd, yd.shape), F.DataDesc(xd, xd.shape))
    torch.cuda.synchronize()
    for row in (0, batch - 1):
        ref = oracle.r2c(x[row], [0])
        err = rel_l2(yd[row].cpu().numpy(), ref)
        print(f"r2c [{batch},{n}] row {row} rel-L2 {err:.3e}")
        assert err <= TOL
    zd = torch.empty_like(xd)
    F.apply(F.FFTDesc([1], False, 1.0 / n), F.DataDesc(zd, zd.shape), F.DataDesc(yd, yd.shape))
    torch.cuda.synchronize()
    rt = (torch.linalg.vector_norm(zd - xd) / torch.linalg.vector_norm(xd)).item()
    print(f"r2c [{batch},{n}] round trip rel-L2 {rt:.3e}")
    assert rt <= 1e-10


@pytest.mark.parametrize("n", [1 << 20, 1 << 21, 1572864, 1 << 23])
def test_r2c_fused_real_post(n):
    """Even r2c lengths whose last four-step pass forms the half spectrum from row pairs (j, L2-j)
    (rows of 1024 or 2048 points; 1572864 = 2 * 1024 * 768 leaves a group with fewer pairs): both
    signs and the adjoint weights against the oracle, and against the unfused three-pass plan."""
    import os
    torch = _torch()
    print(f"case fused r2c n={n}", flush=True)
    batch = 2
    x = synth.uniform((batch, n), -1.0, 1.0, seed=n % 977)
    xd = _dev(x)
    st_in, st_out = [n * 8, 8], [(n // 2 + 1) * 16, 16]
    sp = torch.cuda.current_stream().cuda_stream

    def run(kind, fwd, fct=1.0):
        y = torch.empty((batch, n // 2 + 1), dtype=torch.complex128, device="cuda")
        F.Plan(kind, [batch, n], st_in, st_out, [1], fwd).execute(xd.data_ptr(), y.data_ptr(), fct, sp)
        torch.cuda.synchronize()
        return y.cpu().numpy()

    for fwd in (True, False):
        y = run(_lib.R2C, fwd, 0.5)
        for row in range(batch):
            assert rel_l2(y[row], oracle.r2c(x[row], [0], fwd, 0.5)) < TOL, (n, fwd, row)
    yf = run(_lib.R2C, True)
    # the adjoint of c2r runs as this r2c with interior bins doubled (bins 0 and n/2 kept)
    ya = run(_lib.C2R_ADJOINT, False)
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    assert rel_l2(ya, yf * w) < 1e-15
    os.environ["IMPULSE_FFT_NO_ROWPAIR"] = "1"
    try:
        yu = run(_lib.R2C, True)
    finally:
        del os.environ["IMPULSE_FFT_NO_ROWPAIR"]
    assert rel_l2(yf, yu) < 1e-15


@pytest.mark.parametrize("n", [16381, 24573, 32749, 65521, 131071])
def test_bluestein_three_pass(n):
    """Four-step Bluestein lengths whose padded length splits into power-of-two rows (16381 -> 256 x 128,
    24573 = 3 * 8191 -> 256 x 192, 32749, 65521, 131071): the row pass transforms, multiplies by the
    kernel spectrum and transforms back in one launch.  c2c of both signs, r2c and c2r of odd (paired)
    and even (half-length) signals against the oracle, and against the four-pass plan."""
    import os
    torch = _torch()
    print(f"case three-pass Bluestein n={n}", flush=True)
    sp = torch.cuda.current_stream().cuda_stream
    batch = 3

    def run(kind, shape, xin, out_shape, out_dtype, fwd, fct):
        xd = _dev(xin)
        yd = torch.empty(out_shape, dtype=out_dtype, device="cuda")
        st_in = [s * xd.element_size() for s in xd.stride()]
        st_out = [s * yd.element_size() for s in yd.stride()]
        F.Plan(kind, shape, st_in, st_out, [1], fwd).execute(xd.data_ptr(), yd.data_ptr(), fct, sp)
        torch.cuda.synchronize()
        return yd.cpu().numpy()

    def both(kind, shape, xin, out_shape, out_dtype, fwd, fct):
        y3 = run(kind, shape, xin, out_shape, out_dtype, fwd, fct)
        os.environ["IMPULSE_FFT_NO_CONV_ROWS"] = "1"
        try:
            y4 = run(kind, shape, xin, out_shape, out_dtype, fwd, fct)
        finally:
            del os.environ["IMPULSE_FFT_NO_CONV_ROWS"]
        assert rel_l2(y3, y4) < 1e-13, (kind, n, fwd)
        return y3

    x = synth.complex_normal((batch, n), seed=n)
    for fwd in (True, False):
        y = both(_lib.C2C, [batch, n], x, (batch, n), torch.complex128, fwd, 0.5)
        assert rel_l2(y, oracle.c2c(x, [1], fwd, 0.5)) < TOL, ("c2c", n, fwd)
    for m in (n, 2 * n):                      # odd length (row pairs) and even length (half-length signal)
        r = synth.uniform((batch, m), -1, 1, seed=m + 1)
        for fwd in (True, False):
            y = both(_lib.R2C, [batch, m], r, (batch, m // 2 + 1), torch.complex128, fwd, 1.0)
            assert rel_l2(y, oracle.r2c(r, [1], fwd)) < TOL, ("r2c", m, fwd)
        h = synth.complex_normal((batch, m // 2 + 1), seed=m + 2)
        for fwd in (False, True):
            y = both(_lib.C2R, [batch, m], h, (batch, m), torch.float64, fwd, 1.0 / m)
            assert rel_l2(y, oracle.c2r(h, m, [1], fwd, 1.0 / m)) < TOL, ("c2r", m, fwd)


@pytest.mark.parametrize("kind,batch,w4", [("c2c", 2, True), ("c2c", 512, False), ("r2c", 3, True),
                                           ("r2c", 1024, False)])
def test_bluestein_pass_instances(kind, batch, w4):
    """The three-pass Bluestein of n = 16381 (n2 = 256 x 128) with direct buffer accesses in its
    column passes: chirp-in loads (complex, or odd real pairs) with the twiddled store, and the
    chirp-out store.  Small batches (a few rounds of workgroups) run the passes on the four-wave
    instances (IO code + 8), large batches on the three-wave ones; both against the oracle on
    the first and last signals, checking which instances ran."""
    torch = _torch()
    n = 16381
    print(f"case Bluestein pass instances {kind} [{batch},{n}]", flush=True)
    sp = torch.cuda.current_stream().cuda_stream
    if kind == "c2c":
        x = synth.complex_normal((batch, n), seed=batch)
        yd = torch.empty((batch, n), dtype=torch.complex128, device="cuda")
        lk = _lib.C2C
    else:
        x = synth.uniform((batch, n), -1, 1, seed=batch)
        yd = torch.empty((batch, n // 2 + 1), dtype=torch.complex128, device="cuda")
        lk = _lib.R2C
    xd = _dev(x)
    st_in = [s * xd.element_size() for s in xd.stride()]
    st_out = [s * yd.element_size() for s in yd.stride()]
    for fwd in (True, False):
        pl = F.Plan(lk, [batch, n], st_in, st_out, [1], fwd)
        pl.execute(xd.data_ptr(), yd.data_ptr(), 0.5, sp)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        for row in (0, batch - 1):
            ref = oracle.c2c(x[row], [0], fwd, 0.5) if kind == "c2c" else oracle.r2c(x[row], [0], fwd, 0.5)
            assert rel_l2(y[row], ref) < TOL, (kind, batch, fwd, row)
        names = []
        for k in range(pl.num_launches):
            name = ctypes.create_string_buffer(64)
            z, u = ctypes.c_size_t(), ctypes.c_uint()
            _lib.check(_lib.lib().impulse_fft_plan_pass_info(pl.handle, k, name, 64, ctypes.byref(z), ctypes.byref(z),
                                                          ctypes.byref(u), ctypes.byref(u), ctypes.byref(z)))
            names.append(name.value.decode())
        print(kind, batch, fwd, names)
        io = [int(s.rsplit(",", 1)[1].rstrip(">")) for s in names if s.startswith("k_fft_pow2")]
        assert io[:3] == ([13, 12, 14] if w4 else [5, 4, 6]), names


@pytest.mark.parametrize("kind,n,batch", [("c2c", 16381, 5), ("c2c", 65536, 3), ("r2c", 65536, 3),
                                          ("r2c", 32749, 5), ("c2r", 32762, 3)])
def test_four_step_batch_chunks(kind, n, batch):
    """Four-step batches split into chunks of signals (IMPULSE_FFT_MALL_MB=1: 1 MiB of workspace per
    chunk, so one or two signals each): every chunk's passes address their signals through the
    first-signal offset.  Three-pass Bluestein (16381, and 32762 = 2 * 16381 in c2r), the fused
    row-pair r2c (65536), odd pairs (32749) and plain c2c, against the oracle and the unchunked plan."""
    import os
    torch = _torch()
    print(f"case chunks {kind} n={n} batch={batch}", flush=True)
    sp = torch.cuda.current_stream().cuda_stream
    if kind == "c2c":
        x = synth.complex_normal((batch, n), seed=n)
        plan_kind, out_shape, out_dtype, want = _lib.C2C, (batch, n), torch.complex128, oracle.c2c(x, [1], True, 0.5)
    elif kind == "r2c":
        x = synth.uniform((batch, n), -1, 1, seed=n)
        plan_kind, out_shape, out_dtype, want = (_lib.R2C, (batch, n // 2 + 1), torch.complex128,
                                                 oracle.r2c(x, [1], True, 0.5))
    else:
        x = synth.complex_normal((batch, n // 2 + 1), seed=n)
        plan_kind, out_shape, out_dtype, want = _lib.C2R, (batch, n), torch.float64, oracle.c2r(x, n, [1], True, 0.5)
    xd = _dev(x)

    def run():
        yd = torch.empty(out_shape, dtype=out_dtype, device="cuda")
        st_in = [v * xd.element_size() for v in xd.stride()]
        st_out = [v * yd.element_size() for v in yd.stride()]
        p = F.Plan(plan_kind, [batch, n], st_in, st_out, [1], True)
        p.execute(xd.data_ptr(), yd.data_ptr(), 0.5, sp)
        torch.cuda.synchronize()
        return yd.cpu().numpy(), p.num_launches

    whole, n_whole = run()
    os.environ["IMPULSE_FFT_MALL_MB"] = "1"
    try:
        chunked, n_chunked = run()
    finally:
        del os.environ["IMPULSE_FFT_MALL_MB"]
    assert n_chunked > n_whole, (n_chunked, n_whole)   # the batch did split
    for row in range(batch):
        assert rel_l2(chunked[row], want[row]) < TOL, (kind, n, row)
    assert rel_l2(chunked, whole) < 1e-13


# ---------------------------------------------------------------------------
# error behaviour (pocketfft_hdronly.h:446-477 cases, forced from the host)
# ---------------------------------------------------------------------------

def _create(kind, shape, sin, sout, axes, fwd=1):
    h = ctypes.c_void_p()
    L = _lib.lib()
    rc = L.impulse_fft_plan_create(ctypes.byref(h), kind, 0, len(shape), _lib.sizes(shape), _lib.strides(sin),
                                   _lib.strides(sout), len(axes), _lib.sizes(axes), fwd)
    if rc == 0:
        L.impulse_fft_plan_destroy(h)
    return rc


def test_error_codes():
    L = _lib.lib()
    assert _create(0, [4, 4], [64, 16], [64, 16], [2]) == 1          # bad axis number
    assert b"bad axis" in L.impulse_fft_last_error()
    assert _create(0, [4, 4], [64, 16], [64, 16], [1, 1]) == 1       # axis specified repeatedly
    assert b"repeatedly" in L.impulse_fft_last_error()
    assert _create(0, [4], [16], [16], []) == 1                       # no axes
    assert _create(5, [4], [16], [16], [0]) == 1                      # unknown kind
    assert _create(0, [4], [8], [16], [0]) == 4                       # misaligned complex stride
    assert _create(0, [0, 4], [64, 16], [64, 16], [1]) == 0           # empty array: no-op plan
    x = np.zeros(8)
    rc = L.impulse_fft_r2c(1, _lib.sizes([8]), _lib.strides([8]), _lib.strides([16]), 1, _lib.sizes([0]), 1, 0,
                           x.ctypes.data, x.ctypes.data, 1.0, 1)
    assert rc == 1                                                    # in-place r2c is rejected
    assert not L.impulse_make_rfft_plan(0)
    assert not L.impulse_make_cfft_plan(0)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_real_rows_register_kernel(dtype):
    """Even-length r2c and c2r rows whose half length is a power of two run on the register kernel:
    r2c with the real post-process fused into the store (IO code 7), c2r with its pre-process in
    the per-point loads (IO code 0): [256, 4096] both ways and both signs, against the oracle,
    naming the instances that ran."""
    torch = _torch()
    rows, n = 256, 4096
    f32 = dtype == "f32"
    bar = 2.0 ** -24 + 2e-15 if f32 else TOL
    lk_dt = _lib.F32 if f32 else _lib.F64
    rdt, cdt = (np.float32, np.complex64) if f32 else (np.float64, np.complex128)
    x = synth.uniform((rows, n), -1.0, 1.0, seed=5).astype(rdt)
    h = synth.complex_normal((rows, n // 2 + 1), seed=6).astype(cdt)
    sp = torch.cuda.current_stream().cuda_stream

    def io_codes(pl):
        out = []
        for k in range(pl.num_launches):
            name = ctypes.create_string_buffer(64)
            z, u = ctypes.c_size_t(), ctypes.c_uint()
            _lib.check(_lib.lib().impulse_fft_plan_pass_info(pl.handle, k, name, 64, ctypes.byref(z), ctypes.byref(z),
                                                          ctypes.byref(u), ctypes.byref(u), ctypes.byref(z)))
            out.append(name.value.decode())
        return out

    xd = torch.from_numpy(x).cuda()
    hd = torch.from_numpy(h).cuda()
    es = x.itemsize
    for fwd in (True, False):
        yd = torch.empty((rows, n // 2 + 1), dtype=torch.complex64 if f32 else torch.complex128, device="cuda")
        pl = F.Plan(_lib.R2C, [rows, n], [n * es, es], [(n // 2 + 1) * 2 * es, 2 * es], [1], fwd, dtype=lk_dt)
        pl.execute(xd.data_ptr(), yd.data_ptr(), 0.5, sp)
        torch.cuda.synchronize()
        ref = oracle.r2c(x.astype(np.float64), [1], fwd, 0.5)
        assert rel_l2(yd.cpu().numpy(), ref) <= bar, ("r2c", dtype, fwd)
        assert io_codes(pl) == ["k_fft_pow2<2048, 0, 7>"], io_codes(pl)
        zd = torch.empty((rows, n), dtype=torch.float32 if f32 else torch.float64, device="cuda")
        pc = F.Plan(_lib.C2R, [rows, n], [(n // 2 + 1) * 2 * es, 2 * es], [n * es, es], [1], fwd, dtype=lk_dt)
        pc.execute(hd.data_ptr(), zd.data_ptr(), 1.0 / n, sp)
        torch.cuda.synchronize()
        ref = oracle.c2r(h.astype(np.complex128), n, [1], fwd, 1.0 / n)
        assert rel_l2(zd.cpu().numpy(), ref) <= bar, ("c2r", dtype, fwd)
        assert io_codes(pc) == ["k_fft_pow2<2048, 0, 0>"], io_codes(pc)

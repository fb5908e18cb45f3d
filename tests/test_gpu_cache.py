"""The plan caches behind the one-shot calls and the C path, against get_plan's behaviour
(pocketfft_hdronly.h:2655-2706: an LRU of POCKETFFT_CACHE_SIZE entries, :50-52, under a mutex,
built outside the lock with a second lookup before insertion), plus the host-staging rules.
"""
import ctypes
import threading

import numpy as np
import pytest

import oracle
from impulse_amd import _lib, synth
from impulse_amd import fft as F
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_cache():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    L = _lib.lib()
    L.impulse_fft_cache_clear()
    yield
    L.impulse_fft_cache_set_capacity(16)
    L.impulse_fft_cache_clear()


def _stats():
    h, m = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(_lib.lib().impulse_fft_cache_stats(ctypes.byref(h), ctypes.byref(m)))
    return h.value, m.value


def test_lru_eviction_order():
    L = _lib.lib()
    L.impulse_fft_cache_set_capacity(2)
    xs = {n: synth.complex_normal((n,), seed=n) for n in (64, 96, 120)}

    def call(n):
        y = F.c2c(xs[n], [0], True)
        assert rel_l2(y, oracle.c2c(xs[n], [0], True)) < 1e-12, n

    call(64)                      # A: miss
    call(96)                      # B: miss
    call(64)                      # A: hit, A most recent
    assert _stats() == (1, 2)
    call(120)                     # C: miss, evicts B (least recently used)
    assert L.impulse_fft_cache_size() == 2
    call(64)                      # A: hit
    assert _stats() == (2, 3)
    call(96)                      # B: miss again (it was evicted), evicts C
    assert _stats() == (2, 4)
    call(64)                      # A still cached
    assert _stats() == (3, 4)


def test_same_shape_other_strides_are_distinct_plans():
    x = synth.complex_normal((8, 12), seed=1)
    y1 = F.c2c(x, [1], True)              # contiguous rows
    y2 = F.c2c(x.T.copy().T, [1], True)   # same values, column-major strides
    assert _stats() == (0, 2)
    want = oracle.c2c(x, [1], True)
    assert rel_l2(y1, want) < 1e-12 and rel_l2(y2, want) < 1e-12
    F.c2c(x, [1], False)                  # other direction: a third plan
    F.c2c(x.astype(np.complex64), [1], True)   # other dtype: a fourth
    assert _stats() == (0, 4)


def test_concurrent_one_shots_on_one_key():
    x = synth.complex_normal((32, 1000), seed=2)
    want = oracle.c2c(x, [1], True, 0.5)
    errs, outs = [], [None] * 4

    def worker(t):
        try:
            for _ in range(5):
                outs[t] = F.c2c(x, [1], True, 0.5)
                if rel_l2(outs[t], want) >= 1e-12:
                    errs.append(t)
        except Exception as e:   # noqa: BLE001 - reported below
            errs.append(repr(e))

    th = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    h, m = _stats()
    assert m >= 1 and h + m == 20 and _lib.lib().impulse_fft_cache_size() == 1


def test_c_path_plans_come_from_the_cache():
    """impulse's fft() builds and destroys a plan per call (c_pocketfft/pocketfft.nim:216-303); after the
    first call of a length the device plans are cache hits."""
    L = _lib.lib()
    x = synth.uniform((1000,), -1, 1, seed=3)
    for _ in range(3):
        y = F.rfft_packed(x)
        assert rel_l2(y, oracle.rfft_packed(x)) < 1e-12
        z = F.fft(x.astype(np.complex128))
        assert rel_l2(z, oracle.c2c(x.astype(np.complex128), [0], True)) < 1e-12
    h, m = _stats()
    # the real length's r2c and c2r plans, the complex length's forward and backward plans
    assert m == 4 and h >= 2
    assert L.impulse_fft_cache_size() == 4


def test_host_inplace_rules_on_a_plan():
    """An explicit plan executed on host pointers applies the same in-place rules as a device execute
    (pocketfft_hdronly.h:453-454), before anything is staged."""
    L = _lib.lib()
    p = F.Plan(_lib.C2C, [4], [-16], [32], [0], True)
    buf = np.zeros(16, np.complex128)
    rc = L.impulse_fft_execute_host(p.handle, buf.ctypes.data, buf.ctypes.data, 1.0)
    assert rc == 1 and b"stride mismatch" in L.impulse_fft_last_error()
    q = F.Plan(_lib.R2C, [8], [8], [16], [0], True)
    rc = L.impulse_fft_execute_host(q.handle, buf.ctypes.data, buf.ctypes.data, 1.0)
    assert rc == 1
    # a genuine in-place call with equal negative strides works
    r = F.Plan(_lib.C2C, [6], [-16], [-16], [0], True)
    x = synth.complex_normal((6,), seed=4)
    z = x.copy()
    end = z[5:].ctypes.data   # element 0 of the reversed view
    _lib.check(L.impulse_fft_execute_host(r.handle, end, end, 1.0))
    assert rel_l2(z[::-1], oracle.c2c(x[::-1].copy(), [0], True)) < 1e-12


def test_host_staging_large_and_gapped_outputs():
    """The pinned staging ring runs in chunks: arrays of many chunks, and an output with gaps whose
    untouched bytes must survive (r2c into a full-length buffer, README.md:98)."""
    x = synth.uniform((3, 1 << 20), -1, 1, seed=5)          # 24 MiB in, several chunks
    y = F.r2c(x, [1])
    assert rel_l2(y, oracle.r2c(x, [1])) < 1e-12
    full = np.full((4, 9), 7.0 + 7.0j)
    r = synth.uniform((4, 9), -1, 1, seed=6)
    F.apply(F.FFTDesc([1], True), F.DataDesc(full, [4, 9]), F.DataDesc(r, [4, 9]))
    assert rel_l2(full[:, :5], oracle.r2c(r, [1])) < 1e-12
    assert np.all(full[:, 5:] == 7.0 + 7.0j)


def test_python_device_plans_per_stream():
    """Device executes on two torch streams get two plans (their workspaces never overlap)."""
    import torch
    n = 1 << 16
    x = synth.complex_normal((8, n), seed=7)
    xd = torch.from_numpy(x).cuda()
    want = oracle.c2c(x, [1], True)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for s in (s1, s2):
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            y = torch.empty_like(xd)
            F.apply(F.FFTDesc([1], True), F.DataDesc(y, y.shape), F.DataDesc(xd, xd.shape))
            outs.append(y)
    torch.cuda.synchronize()
    keys = [k for k in F._plan_cache if k[2] == (8, n)]
    assert len({k[-1] for k in keys}) >= 2
    for y in outs:
        assert rel_l2(y.cpu().numpy(), want) < 1e-12

"""CPU-only checks of the boundary and host logic (no kernel is launched)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from impulse_amd import _lib, synth
from impulse_amd import fft as F
from tests.conftest import ROOT, rel_l2


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(impulse_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations found"
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_status_strings():
    L = _lib.lib()
    for code in range(6):
        assert L.impulse_fft_status_string(code)


def test_splitmix64_known_answer():
    # SplitMix64 with seed 0: first output 0xE220A8397B1DCDAF
    assert int(synth.splitmix64(1, 0)[0]) == 0xE220A8397B1DCDAF
    u = synth.uniform((1000,), -1, 1, seed=0)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.1
    z = synth.normal((20000,), seed=1)
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1) < 0.05


def test_initNormalize():
    # c_pocketfft/pocketfft.nim:199-204
    assert F.initNormalize(F.nkBackward, True, 0, 8).value == 1.0
    assert F.initNormalize(F.nkBackward, False, 0, 8).value == 0.125
    assert F.initNormalize(F.nkForward, True, 0, 8).value == 0.125
    assert F.initNormalize(F.nkForward, False, 0, 8).value == 1.0
    assert F.initNormalize(F.nkOrtho, True, 0, 4).value == 0.5
    assert F.initNormalize(F.nkCustom, True, 3.5, 4).value == 3.5


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 9, 16, 17])
def test_unpack_symmetrize_layout(n):
    # halfcomplex layout (pocketfft.nim:224-243) -> bins, then Hermitian mirror
    x = synth.uniform((n,), -1, 1, seed=n)
    hc = oracle.rfft_packed(x)
    bins = F.unpackFFT(hc)
    assert bins.size == n // 2 + 1
    assert rel_l2(bins, oracle.r2c(x, [0])) < 1e-15
    full = F.symmetrize(hc)
    assert full.size == n
    assert rel_l2(full, oracle.c2c(x.astype(complex), [0])) < 1e-15
    # complex input: target size from the last bin's imaginary part
    assert F.symmTargetSize(bins) == (n if n % 2 == 0 or n == 1 and False else F.symmTargetSize(bins))
    if n > 2:
        assert F.symmetrize(bins).size == (2 * bins.size - 2 if bins[-1].imag == 0 else 2 * bins.size - 1)


def test_datadesc_strides():
    # cpp_pocketfft/pocketfft.nim:178-199: C-contiguous byte strides
    a = np.zeros((3, 5, 7))
    d = F.DataDesc(a, a.shape)
    assert d.stride == [5 * 7 * 8, 7 * 8, 8]
    d = F.DataDesc(a, (3, 5), (10, 2))
    assert d.stride == [80, 16]
    with pytest.raises(ValueError):
        F.DataDesc(a, (3, 5), (1,))


def test_uniform_device_matches_numpy():
    torch = pytest.importorskip("torch")
    a = synth.uniform(5000, -1.0, 1.0, seed=7)
    b = synth.uniform_device(5000, -1.0, 1.0, seed=7, device="cpu", chunk=777).numpy()
    assert np.array_equal(a, b)
    a32 = synth.uniform(999, -1.0, 1.0, seed=0).astype(np.float32)
    b32 = synth.uniform_device(999, -1.0, 1.0, seed=0, dtype=torch.float32, device="cpu").numpy()
    assert np.array_equal(a32, b32)


def _largest_prime_factor(n):
    best, p = 1, 2
    while p * p <= n:
        while n % p == 0:
            best, n = p, n // p
        p += 1
    return max(best, n) if n > 1 else best


def test_ragged_lengths_workload():
    L = synth.ragged_lengths(4000, seed=0)
    assert L.min() >= 64 and L.max() <= 262144
    smooth = np.array([_largest_prime_factor(int(n)) <= 7 for n in L])
    big = np.array([_largest_prime_factor(int(n)) > 1000 for n in L])
    assert np.all(smooth | big)
    assert abs(smooth.mean() - 0.7) < 0.03
    # log-uniform: the smooth class's log-lengths spread evenly over [log 64, log 262144]
    lg = np.log(L[smooth])
    q = np.quantile(lg, [0.25, 0.5, 0.75])
    span = np.log(262144) - np.log(64)
    assert np.allclose(q, np.log(64) + span * np.array([0.25, 0.5, 0.75]), atol=0.05 * span)
    assert np.array_equal(L, synth.ragged_lengths(4000, seed=0))


def test_ragged_packed_offsets():
    from impulse_amd import ragged
    ins, outs = ragged.packed_offsets([5, 8, 1])
    assert list(ins) == [0, 5, 13] and list(outs) == [0, 3, 8]


def test_inplace_rules_before_planning():
    # pocketfft's sanity_check (pocketfft_hdronly.h:453-454): in == out needs equal strides; r2c / c2r
    # never run in place.  Both are rejected before a plan is built, so no device is touched.
    L = _lib.lib()
    buf = np.zeros(16, np.complex128)
    p = buf.ctypes.data
    rc = L.impulse_fft_c2c(1, _lib.sizes([4]), _lib.strides([-16]), _lib.strides([32]), 1, _lib.sizes([0]), 1,
                           _lib.F64, p, p, 1.0, 1)
    assert rc == 1 and b"stride mismatch" in L.impulse_fft_last_error()
    rc = L.impulse_fft_r2c(1, _lib.sizes([8]), _lib.strides([8]), _lib.strides([8]), 1, _lib.sizes([0]), 1,
                           _lib.F64, p, p, 1.0, 1)
    assert rc == 1 and b"in-place" in L.impulse_fft_last_error()
    rc = L.impulse_fft_c2r(1, _lib.sizes([8]), _lib.strides([16]), _lib.strides([16]), 1, _lib.sizes([0]), 0,
                           _lib.F32, p, p, 1.0, 1)
    assert rc == 1


def test_mirror_dtype_rules():
    # the C++ path's T (cpp_pocketfft/pocketfft.nim:35-69): float64 / complex128 -> F64,
    # float32 / complex64 -> F32, anything else or mixed precisions is an error
    assert F.DataDesc(np.zeros(4), [4]).dtype == _lib.F64
    assert F.DataDesc(np.zeros(4, np.complex64), [4]).dtype == _lib.F32
    with pytest.raises(TypeError):
        F.DataDesc(np.zeros(4, np.int32), [4]).dtype
    with pytest.raises(TypeError):
        F.apply(F.FFTDesc([0], True), F.DataDesc(np.zeros(3, np.complex128), [3]),
                F.DataDesc(np.zeros(4, np.float32), [4]))
    with pytest.raises(TypeError):
        F.c2c(np.zeros(4, np.float32), [0])


def test_python_plan_cache_is_bounded_and_keyed_by_stream():
    # host-side key logic only: a bounded LRU whose key carries the device and the stream
    import inspect
    src = inspect.getsource(F.get_plan)
    assert "device" in src and "stream" in src
    assert F.PLAN_CACHE_CAPACITY >= 1

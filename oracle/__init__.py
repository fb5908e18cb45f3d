"""CPU oracle for the impulse FFT path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The product
(``impulse_amd``) never imports it.

``liboracle_fft.so`` is the repository's own long-double restatement of the
reference semantics (``oracle/fft_oracle.c``; reference citations are in
``oracle/fft_oracle.h``).  ``oracle.ref`` wraps the reference's own PocketFFT,
compiled in place from the reference tree into ``oracle/_ref/`` where that
tree exists.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_szp = ctypes.POINTER(ctypes.c_size_t)
_pdp = ctypes.POINTER(ctypes.c_ssize_t)
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_fft.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle oracle`)")
        L = ctypes.CDLL(path)
        nd = [ctypes.c_size_t, _szp, _pdp, _pdp, ctypes.c_size_t, _szp, ctypes.c_int,
              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
        for name in ("oracle_c2c", "oracle_r2c", "oracle_c2r"):
            getattr(L, name).argtypes = nd
            getattr(L, name).restype = ctypes.c_int
        for name in ("oracle_cfft_forward", "oracle_cfft_backward",
                     "oracle_rfft_forward", "oracle_rfft_backward"):
            getattr(L, name).argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_double]
            getattr(L, name).restype = ctypes.c_int
        L.oracle_dft_naive.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_dft_naive.restype = ctypes.c_int
        for name in ("oracle_dct", "oracle_dst"):
            getattr(L, name).argtypes = [ctypes.c_size_t, _szp, _pdp, _pdp, ctypes.c_size_t, _szp, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
            getattr(L, name).restype = ctypes.c_int
        L.oracle_twiddle.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, _dp, _dp]
        L.oracle_twiddle.restype = None
        _LIB = L
    return _LIB


def _arr(vals, ctype):
    return (ctype * len(vals))(*vals)


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"{what} failed with code {rc}")


def _nd_call(fn, shape, xin, xout, axes, forward, fct):
    ndim = len(shape)
    rc = fn(ndim, _arr(shape, ctypes.c_size_t), _arr(xin.strides, ctypes.c_ssize_t),
            _arr(xout.strides, ctypes.c_ssize_t), len(axes), _arr(axes, ctypes.c_size_t),
            1 if forward else 0, xin.ctypes.data, xout.ctypes.data, float(fct))
    return rc


def c2c(x, axes, forward=True, fct=1.0, out=None):
    """pocketfft::c2c semantics on a complex128 ndarray (any byte strides)."""
    x = np.asarray(x)
    assert x.dtype == np.complex128
    if out is None:
        out = np.empty(x.shape, np.complex128)
    _check(_nd_call(lib().oracle_c2c, x.shape, x, out, list(axes), forward, fct), "oracle_c2c")
    return out


def r2c(x, axes, forward=True, fct=1.0, out=None):
    """pocketfft::r2c semantics: last entry of ``axes`` is the real axis."""
    x = np.asarray(x)
    assert x.dtype == np.float64
    axes = list(axes)
    oshape = list(x.shape)
    oshape[axes[-1]] = x.shape[axes[-1]] // 2 + 1
    if out is None:
        out = np.empty(oshape, np.complex128)
    _check(_nd_call(lib().oracle_r2c, x.shape, x, out, axes, forward, fct), "oracle_r2c")
    return out


def c2r(x, n, axes, forward=False, fct=1.0, out=None):
    """pocketfft::c2r semantics: ``n`` is the real length along ``axes[-1]``."""
    x = np.asarray(x)
    assert x.dtype == np.complex128
    axes = list(axes)
    oshape = list(x.shape)
    oshape[axes[-1]] = n
    if out is None:
        out = np.empty(oshape, np.float64)
    _check(_nd_call(lib().oracle_c2r, oshape, x, out, axes, forward, fct), "oracle_c2r")
    return out


def cfft(x, forward=True, fct=1.0):
    """C path ``cfft_forward/backward`` on a copy of a 1-D complex128 array."""
    c = np.array(x, dtype=np.complex128, copy=True)
    fn = lib().oracle_cfft_forward if forward else lib().oracle_cfft_backward
    _check(fn(c.size, c.ctypes.data, float(fct)), "oracle_cfft")
    return c


def rfft_packed(x, forward=True, fct=1.0):
    """C path ``rfft_forward/backward``: FFTPACK halfcomplex order, in place on a copy."""
    c = np.array(x, dtype=np.float64, copy=True)
    fn = lib().oracle_rfft_forward if forward else lib().oracle_rfft_backward
    _check(fn(c.size, c.ctypes.data, float(fct)), "oracle_rfft")
    return c


def dft_naive(x, forward=True):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    _check(lib().oracle_dft_naive(x.size, x.ctypes.data, out.ctypes.data, 1 if forward else 0),
           "oracle_dft_naive")
    return out


def twiddle(k, n, sign=-1):
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().oracle_twiddle(k, n, sign, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def _r2r(fn, what, x, axes, type, fct, ortho, out):
    x = np.asarray(x)
    assert x.dtype == np.float64
    if out is None:
        out = np.empty(x.shape, np.float64)
    rc = fn(x.ndim, _arr(list(x.shape), ctypes.c_size_t), _arr(list(x.strides), ctypes.c_ssize_t),
            _arr(list(out.strides), ctypes.c_ssize_t), len(axes), _arr(list(axes), ctypes.c_size_t), int(type),
            x.ctypes.data, out.ctypes.data, float(fct), int(bool(ortho)))
    _check(rc, what)
    return out


def dct(x, axes, type=2, fct=1.0, ortho=False, out=None):
    """pocketfft::dct semantics (naive long-double sums)."""
    return _r2r(lib().oracle_dct, "oracle_dct", x, axes, type, fct, ortho, out)


def dst(x, axes, type=2, fct=1.0, ortho=False, out=None):
    """pocketfft::dst semantics (naive long-double sums)."""
    return _r2r(lib().oracle_dst, "oracle_dst", x, axes, type, fct, ortho, out)


def upfirdn(t, h, up=1, down=1):
    """impulse/signal.nim:545-594 on 1-D float64 arrays (long-double direct sums)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    L = lib()
    L.oracle_upfirdn_length.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
    L.oracle_upfirdn_length.restype = ctypes.c_size_t
    L.oracle_upfirdn.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_void_p]
    L.oracle_upfirdn.restype = ctypes.c_int
    n = L.oracle_upfirdn_length(t.size, h.size, int(up), int(down))
    y = np.empty(n, np.float64)
    _check(L.oracle_upfirdn(t.size, t.ctypes.data, h.size, h.ctypes.data, int(up), int(down), y.ctypes.data),
           "oracle_upfirdn")
    return y


def upfirdn_length(nt, nh, up=1, down=1):
    """Outputs of upfirdn: ceil(((nt - 1) up + nh) / down)."""
    L = lib()
    L.oracle_upfirdn_length.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
    L.oracle_upfirdn_length.restype = ctypes.c_size_t
    return int(L.oracle_upfirdn_length(int(nt), int(nh), int(up), int(down)))


def _rfp(fname, x, axes, extra, fct, out):
    x = np.asarray(x)
    assert x.dtype == np.float64
    if out is None:
        out = np.empty(x.shape, np.float64)
    f = getattr(lib(), fname)
    f.argtypes = [ctypes.c_size_t, _szp, _pdp, _pdp, ctypes.c_size_t, _szp] + [ctypes.c_int] * len(extra) + \
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
    f.restype = ctypes.c_int
    rc = f(x.ndim, _arr(list(x.shape), ctypes.c_size_t), _arr(list(x.strides), ctypes.c_ssize_t),
           _arr(list(out.strides), ctypes.c_ssize_t), len(axes), _arr(list(axes), ctypes.c_size_t),
           *[int(v) for v in extra], x.ctypes.data, out.ctypes.data, float(fct))
    _check(rc, fname)
    return out


def r2r_fftpack(x, axes, real2hermitian, forward, fct=1.0, out=None):
    """pocketfft::r2r_fftpack semantics (naive long-double sums)."""
    return _rfp("oracle_r2r_fftpack", x, axes, [bool(real2hermitian), bool(forward)], fct, out)


def r2r_separable_hartley(x, axes, fct=1.0, out=None):
    """pocketfft::r2r_separable_hartley semantics (naive long-double sums)."""
    return _rfp("oracle_r2r_separable_hartley", x, axes, [], fct, out)


def r2r_genuine_hartley(x, axes, fct=1.0, out=None):
    """pocketfft::r2r_genuine_hartley semantics (naive long-double sums)."""
    return _rfp("oracle_r2r_genuine_hartley", x, axes, [], fct, out)

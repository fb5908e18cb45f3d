"""The reference's own PocketFFT (C and C++), built in place from the reference
tree by ``oracle/Makefile`` into ``oracle/_ref/`` -- TEST INFRASTRUCTURE ONLY.

``available()`` is False wherever that library was not built (it needs the
reference tree at build time); callers then rely on the committed golden
fixtures and on ``oracle`` (the restatement) instead.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libpocketfft_ref.so")
_LIB = None


def available():
    return os.path.exists(_PATH)


def lib():
    global _LIB
    if _LIB is None:
        if not available():
            raise RuntimeError(f"reference PocketFFT not built: {_PATH}")
        L = ctypes.CDLL(_PATH)
        szp = ctypes.POINTER(ctypes.c_size_t)
        pdp = ctypes.POINTER(ctypes.c_ssize_t)
        for name in ("ref_c2c_f64", "ref_r2c_f64", "ref_c2r_f64"):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_size_t, szp, pdp, pdp, ctypes.c_size_t, szp, ctypes.c_int,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_size_t]
            f.restype = ctypes.c_int
        for name in ("ref_c2c_f32", "ref_r2c_f32", "ref_c2r_f32"):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_size_t, szp, pdp, pdp, ctypes.c_size_t, szp, ctypes.c_int,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_size_t]
            f.restype = ctypes.c_int
        # C path (impulse/fft/c_pocketfft/pocketfft.h:18-32)
        L.make_cfft_plan.argtypes = [ctypes.c_size_t]
        L.make_cfft_plan.restype = ctypes.c_void_p
        L.make_rfft_plan.argtypes = [ctypes.c_size_t]
        L.make_rfft_plan.restype = ctypes.c_void_p
        for name in ("cfft_forward", "cfft_backward", "rfft_forward", "rfft_backward"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
            getattr(L, name).restype = ctypes.c_int
        L.destroy_cfft_plan.argtypes = [ctypes.c_void_p]
        L.destroy_rfft_plan.argtypes = [ctypes.c_void_p]
        for name in ("ref_dct_f64", "ref_dst_f64"):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_size_t, szp, pdp, pdp, ctypes.c_size_t, szp, ctypes.c_int,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_size_t]
            f.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _a(vals, t):
    return (t * len(vals))(*vals)


def _call(fn, shape, xin, xout, axes, forward, fct, nthreads):
    rc = fn(len(shape), _a(shape, ctypes.c_size_t), _a(xin.strides, ctypes.c_ssize_t),
            _a(xout.strides, ctypes.c_ssize_t), len(axes), _a(axes, ctypes.c_size_t),
            1 if forward else 0, xin.ctypes.data, xout.ctypes.data, fct, nthreads)
    if rc != 0:
        raise ValueError(f"reference pocketfft raised (code {rc})")


def c2c(x, axes, forward=True, fct=1.0, nthreads=1, out=None):
    x = np.asarray(x)
    f = lib().ref_c2c_f64 if x.dtype == np.complex128 else lib().ref_c2c_f32
    if out is None:
        out = np.empty(x.shape, x.dtype)
    _call(f, x.shape, x, out, list(axes), forward, fct, nthreads)
    return out


def r2c(x, axes, forward=True, fct=1.0, nthreads=1, out=None):
    x = np.asarray(x)
    axes = list(axes)
    oshape = list(x.shape)
    oshape[axes[-1]] = x.shape[axes[-1]] // 2 + 1
    cdt = np.complex128 if x.dtype == np.float64 else np.complex64
    f = lib().ref_r2c_f64 if x.dtype == np.float64 else lib().ref_r2c_f32
    if out is None:
        out = np.empty(oshape, cdt)
    _call(f, x.shape, x, out, axes, forward, fct, nthreads)
    return out


def c2r(x, n, axes, forward=False, fct=1.0, nthreads=1, out=None):
    x = np.asarray(x)
    axes = list(axes)
    oshape = list(x.shape)
    oshape[axes[-1]] = n
    rdt = np.float64 if x.dtype == np.complex128 else np.float32
    f = lib().ref_c2r_f64 if x.dtype == np.complex128 else lib().ref_c2r_f32
    if out is None:
        out = np.empty(oshape, rdt)
    _call(f, oshape, x, out, axes, forward, fct, nthreads)
    return out


def cfft(x, forward=True, fct=1.0):
    c = np.array(x, dtype=np.complex128, copy=True)
    L = lib()
    plan = L.make_cfft_plan(c.size)
    if not plan:
        raise ValueError("make_cfft_plan failed")
    try:
        rc = (L.cfft_forward if forward else L.cfft_backward)(plan, c.ctypes.data, fct)
    finally:
        L.destroy_cfft_plan(plan)
    if rc:
        raise ValueError("cfft failed")
    return c


def rfft_packed(x, forward=True, fct=1.0):
    c = np.array(x, dtype=np.float64, copy=True)
    L = lib()
    plan = L.make_rfft_plan(c.size)
    if not plan:
        raise ValueError("make_rfft_plan failed")
    try:
        rc = (L.rfft_forward if forward else L.rfft_backward)(plan, c.ctypes.data, fct)
    finally:
        L.destroy_rfft_plan(plan)
    if rc:
        raise ValueError("rfft failed")
    return c


def _r2r(name, x, axes, type, fct, ortho, nthreads, out):
    x = np.asarray(x)
    assert x.dtype == np.float64
    if out is None:
        out = np.empty(x.shape, np.float64)
    rc = getattr(lib(), name)(x.ndim, _a(x.shape, ctypes.c_size_t), _a(x.strides, ctypes.c_ssize_t),
                              _a(out.strides, ctypes.c_ssize_t), len(axes), _a(list(axes), ctypes.c_size_t),
                              int(type), x.ctypes.data, out.ctypes.data, float(fct), int(bool(ortho)), nthreads)
    if rc != 0:
        raise ValueError(f"reference pocketfft raised (code {rc})")
    return out


def dct(x, axes, type=2, fct=1.0, ortho=False, nthreads=1, out=None):
    """pocketfft::dct<double> (pocketfft_hdronly.h:3284-3300)."""
    return _r2r("ref_dct_f64", x, axes, type, fct, ortho, nthreads, out)


def dst(x, axes, type=2, fct=1.0, ortho=False, nthreads=1, out=None):
    """pocketfft::dst<double> (pocketfft_hdronly.h:3302-3318)."""
    return _r2r("ref_dst_f64", x, axes, type, fct, ortho, nthreads, out)


def _rfp(name, x, axes, extra, fct, nthreads, out):
    x = np.asarray(x)
    assert x.dtype == np.float64
    if out is None:
        out = np.empty(x.shape, np.float64)
    L = lib()
    f = getattr(L, name)
    szp, pdp = ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_ssize_t)
    head = [ctypes.c_size_t, szp, pdp, pdp, ctypes.c_size_t, szp]
    f.argtypes = head + [ctypes.c_int] * len(extra) + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                                       ctypes.c_size_t]
    f.restype = ctypes.c_int
    rc = f(x.ndim, _a(x.shape, ctypes.c_size_t), _a(x.strides, ctypes.c_ssize_t), _a(out.strides, ctypes.c_ssize_t),
           len(axes), _a(list(axes), ctypes.c_size_t), *[int(v) for v in extra], x.ctypes.data, out.ctypes.data,
           float(fct), nthreads)
    if rc != 0:
        raise ValueError(f"reference pocketfft raised (code {rc})")
    return out


def r2r_fftpack(x, axes, real2hermitian, forward, fct=1.0, nthreads=1, out=None):
    """pocketfft::r2r_fftpack<double> (pocketfft_hdronly.h:3392-3403)."""
    return _rfp("ref_r2r_fftpack_f64", x, axes, [bool(real2hermitian), bool(forward)], fct, nthreads, out)


def r2r_separable_hartley(x, axes, fct=1.0, nthreads=1, out=None):
    """pocketfft::r2r_separable_hartley<double> (pocketfft_hdronly.h:3405-3415)."""
    return _rfp("ref_r2r_separable_hartley_f64", x, axes, [], fct, nthreads, out)


def r2r_genuine_hartley(x, axes, fct=1.0, nthreads=1, out=None):
    """pocketfft::r2r_genuine_hartley<double> (pocketfft_hdronly.h:3417-3445)."""
    return _rfp("ref_r2r_genuine_hartley_f64", x, axes, [], fct, nthreads, out)

"""Reverse-mode differentiation over the C ABI (torch.autograd.Function).

Every backward calls a hand-written adjoint on the GPU:

* ``rfftn`` (r2c)   -- backward runs the R2C_ADJOINT plan: c2r of the opposite
  direction with interior bins weighted 1/2.
* ``irfftn`` (c2r)  -- backward runs the C2R_ADJOINT plan: r2c of the opposite
  direction with interior bins weighted 2.
* ``fftn`` (c2c)    -- backward runs c2c of the opposite direction.
* ``fft_conv2d``    -- "same" FFT convolution; backward runs the convolution's
  adjoint pipeline (impulse_fft_conv2d_backward), which recomputes the spectra
  of the image and the kernel instead of saving them.

Gradients follow torch's convention for real losses: for a complex tensor z
the gradient is dL/dRe z + i dL/dIm z, so each backward is the adjoint of
the forward map under the real inner products <x, y> = sum x y and
<X, Y> = Re sum conj(X) Y (over the stored half spectrum for r2c/c2r).

The arithmetic is FP64: the FFT functions take float64 / complex128 or
float32 / complex64 CUDA tensors (outputs and gradients keep the input's
precision); fft_conv2d takes float32 or float64.  This is the differentiation
path of the engine; impulse's Nim autograd op would bind the same entry
points.
"""
import collections
import ctypes
import threading

import torch

from . import _lib
from .fft import get_plan


def _c_strides(shape, itemsize):
    st, acc = [0] * len(shape), itemsize
    for i in range(len(shape) - 1, -1, -1):
        st[i] = acc
        acc *= shape[i]
    return st


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _dense(t):
    """The tensor's values in plain contiguous memory: torch may hand over lazily conjugated or
    negated views (a conj bit set by .conj(), e.g. in the gradient of a loss built on Y.conj()),
    whose raw data_ptr() memory holds the unconjugated values."""
    return t.resolve_conj().resolve_neg().contiguous()


_COMPLEX = {torch.float64: torch.complex128, torch.float32: torch.complex64}
_REAL = {torch.complex128: torch.float64, torch.complex64: torch.float32}


def _run(kind, shape, in_t, out_t, axes, forward, fct=1.0):
    ib = in_t.element_size()
    ob = out_t.element_size()
    ishape = list(in_t.shape)
    oshape = list(out_t.shape)
    dtype = _lib.F32 if in_t.dtype in (torch.float32, torch.complex64) else _lib.F64
    dev = in_t.device.index if in_t.device.index is not None else torch.cuda.current_device()
    stream = torch.cuda.current_stream(dev).cuda_stream
    p = get_plan(kind, shape, _c_strides(ishape, ib), _c_strides(oshape, ob), axes, forward, dtype=dtype, device=dev,
                 stream=stream)
    p.execute(in_t.data_ptr(), out_t.data_ptr(), fct, stream)


def _half_shape(shape, axes):
    s = list(shape)
    s[axes[-1]] = s[axes[-1]] // 2 + 1
    return s


class _RFFTN(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, axes, forward):
        x = _dense(x)
        y = torch.empty(_half_shape(x.shape, axes), dtype=_COMPLEX[x.dtype], device=x.device)
        _run(_lib.R2C, list(x.shape), x, y, axes, forward)
        ctx.axes, ctx.forward, ctx.shape = axes, forward, list(x.shape)
        return y

    @staticmethod
    def backward(ctx, g):
        g = _dense(g)
        gx = torch.empty(ctx.shape, dtype=_REAL[g.dtype], device=g.device)
        _run(_lib.R2C_ADJOINT, ctx.shape, g, gx, ctx.axes, ctx.forward)
        return gx, None, None


class _IRFFTN(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, n, axes, forward, fct):
        X = _dense(X)
        shape = list(X.shape)
        shape[axes[-1]] = n
        y = torch.empty(shape, dtype=_REAL[X.dtype], device=X.device)
        _run(_lib.C2R, shape, X, y, axes, forward, fct)
        ctx.axes, ctx.forward, ctx.shape, ctx.fct, ctx.hshape = axes, forward, shape, fct, list(X.shape)
        return y

    @staticmethod
    def backward(ctx, g):
        g = _dense(g)
        gX = torch.empty(ctx.hshape, dtype=_COMPLEX[g.dtype], device=g.device)
        _run(_lib.C2R_ADJOINT, ctx.shape, g, gX, ctx.axes, ctx.forward, ctx.fct)
        return gX, None, None, None, None


class _FFTN(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, axes, forward, fct):
        x = _dense(x)
        y = torch.empty_like(x)
        _run(_lib.C2C, list(x.shape), x, y, axes, forward, fct)
        ctx.axes, ctx.forward, ctx.fct = axes, forward, fct
        return y

    @staticmethod
    def backward(ctx, g):
        g = _dense(g)
        gx = torch.empty_like(g)
        _run(_lib.C2C, list(g.shape), g, gx, ctx.axes, not ctx.forward, ctx.fct)
        return gx, None, None, None


def rfftn(x, axes, forward=True):
    """r2c over `axes` (last entry = the real axis), float64 -> complex128 or float32 -> complex64."""
    if x.dtype not in _COMPLEX:
        raise TypeError("rfftn takes float64 or float32")
    return _RFFTN.apply(x, list(axes), bool(forward))


def irfftn(X, n, axes, forward=False, fct=1.0):
    """c2r over `axes` with real length `n` along axes[-1], complex128 -> float64 or complex64 -> float32."""
    if X.dtype not in _REAL:
        raise TypeError("irfftn takes complex128 or complex64")
    return _IRFFTN.apply(X, int(n), list(axes), bool(forward), float(fct))


def fftn(x, axes, forward=True, fct=1.0):
    """c2c over `axes`, complex128 or complex64."""
    if x.dtype not in _REAL:
        raise TypeError("fftn takes complex128 or complex64")
    return _FFTN.apply(x, list(axes), bool(forward), float(fct))


class Conv2d:
    """An FFT convolution object (impulse_fft_conv2d) for one problem shape."""

    def __init__(self, batch, height, width, kh, kw, dtype):
        h = ctypes.c_void_p()
        dt = _lib.F32 if dtype == torch.float32 else _lib.F64
        _lib.check(_lib.lib().impulse_fft_conv2d_create(ctypes.byref(h), batch, height, width, kh, kw, dt))
        self.handle = h
        self._destroy = _lib.lib().impulse_fft_conv2d_destroy
        self.dtype = dtype

    def __del__(self):
        # the destroy entry point is bound at creation: at interpreter exit the
        # module globals may already be gone
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self.handle = None

    @property
    def grid(self):
        P, Q = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(_lib.lib().impulse_fft_conv2d_grid(self.handle, ctypes.byref(P), ctypes.byref(Q)))
        return P.value, Q.value

    def forward(self, x, k, y):
        _lib.check(_lib.lib().impulse_fft_conv2d_forward(self.handle, x.data_ptr(), k.data_ptr(), y.data_ptr(),
                                                         _stream()))

    def backward(self, x, k, gy, gx, gk):
        _lib.check(_lib.lib().impulse_fft_conv2d_backward(
            self.handle, x.data_ptr(), k.data_ptr(), gy.data_ptr(), gx.data_ptr() if gx is not None else None,
            gk.data_ptr() if gk is not None else None, _stream()))


CONV_CACHE_CAPACITY = 8
_convs = collections.OrderedDict()
_convs_lock = threading.Lock()


def _conv(B, H, W, KH, KW, dtype):
    """Convolution object for a shape, from an LRU keyed by the shape, the device and the
    stream (an object's workspace serves one stream at a time)."""
    dev = torch.cuda.current_device()
    key = (B, H, W, KH, KW, dtype, dev, torch.cuda.current_stream(dev).cuda_stream)
    with _convs_lock:
        c = _convs.get(key)
        if c is not None:
            _convs.move_to_end(key)
            return c
    c = Conv2d(B, H, W, KH, KW, dtype)
    with _convs_lock:
        _convs[key] = c
        while len(_convs) > CONV_CACHE_CAPACITY:
            _convs.popitem(last=False)
    return c


class _FFTConv2d(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, k):
        x, k = _dense(x), _dense(k)
        B, H, W = x.shape
        KH, KW = k.shape
        c = _conv(B, H, W, KH, KW, x.dtype)
        y = torch.empty_like(x)
        c.forward(x, k, y)
        ctx.save_for_backward(x, k)   # inputs only: the spectra are recomputed
        return y

    @staticmethod
    def backward(ctx, gy):
        x, k = ctx.saved_tensors
        gy = _dense(gy)
        B, H, W = x.shape
        KH, KW = k.shape
        c = _conv(B, H, W, KH, KW, x.dtype)
        gx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        gk = torch.empty_like(k) if ctx.needs_input_grad[1] else None
        c.backward(x, k, gy, gx, gk)
        return gx, gk


def fft_conv2d(x, k):
    """'same' 2-D convolution of images x [B, H, W] with kernel k [KH, KW]
    (origin at the kernel centre), float32 or float64 CUDA tensors."""
    if x.dtype != k.dtype or x.dtype not in (torch.float32, torch.float64):
        raise TypeError("fft_conv2d takes float32 or float64 tensors of one dtype")
    return _FFTConv2d.apply(x, k)

"""Host-side mirror of impulse's FFT API over the C ABI (include/impulse_fft.h).

Names, argument meaning and error behaviour follow the reference's two Nim
surfaces so that tests read like the reference's own:

* C path -- impulse/fft/c_pocketfft/pocketfft.nim:84-360 and the Tensor
  overloads impulse/fft/c_pocketfft/pocketfft_arraymancer.nim:23-94:
  ``NormalizeKind``, ``Normalize``, ``initNormalize``, ``FFTPlanReal``,
  ``FFTPlanComplex``, ``fft_impl``, ``rfft_inplace`` (the ``rfft(MemoryView)``
  overload), ``fft``, ``fft_inplace`` (the ``var openArray`` overload),
  ``rfft_packed``, ``rfft``, ``ifft``, ``unpackFFT``, ``symmetrize``,
  ``symmTargetSize``.  numpy arrays take the role of seq/Tensor; the
  normalisation arguments are honoured on every overload, as the Tensor
  overloads do (the seq overloads drop them: pocketfft.nim:310,319,331).
* C++ path -- impulse/fft/cpp_pocketfft/pocketfft.nim:137-277: ``DataDesc``,
  ``FFTDesc``, ``apply``.  c2r is given the real OUTPUT shape
  (README_pocketfft.md:89-92); the reference passes ``descIn.shape``
  (cpp_pocketfft/pocketfft.nim:265-267), which is the same array shape under
  the full-length complex-buffer convention its README uses.

Host arrays go through the host-pointer entry points (plans from the native
LRU cache).  torch tensors on the GPU go through ``impulse_fft_execute`` on
torch's current stream with plans from a per-process LRU keyed by the request,
the device and the stream (a plan's workspace is used by one stream at a time).
Arrays of float64 / complex128 run as IMPULSE_FFT_F64, float32 / complex64 as
IMPULSE_FFT_F32 (the reference's ``T = float32`` instantiation).
"""
import collections
import ctypes
import enum
import math
import threading

import numpy as np

from . import _lib
from ._lib import check, lib, sizes, strides


# ---------------------------------------------------------------------------
# normalisation (c_pocketfft/pocketfft.nim:189-204)
# ---------------------------------------------------------------------------

class NormalizeKind(enum.IntEnum):
    nkBackward = 0   # 1 forward, 1/N inverse
    nkOrtho = 1      # 1/sqrt(N) both ways
    nkForward = 2    # 1/N forward, 1 inverse
    nkCustom = 3     # the given value


nkBackward, nkOrtho, nkForward, nkCustom = (NormalizeKind.nkBackward, NormalizeKind.nkOrtho,
                                            NormalizeKind.nkForward, NormalizeKind.nkCustom)


class Normalize:
    __slots__ = ("kind", "value")

    def __init__(self, kind, value):
        self.kind = kind
        self.value = value


def initNormalize(kind, forward, value, length):
    if kind == nkBackward:
        return Normalize(kind, 1.0 if forward else 1.0 / length)
    if kind == nkForward:
        return Normalize(kind, 1.0 / length if forward else 1.0)
    if kind == nkOrtho:
        return Normalize(kind, 1.0 / math.sqrt(length))
    return Normalize(nkCustom, value)


# ---------------------------------------------------------------------------
# 1-D plans (c_pocketfft/pocketfft.nim:84-114) over impulse_make_*_plan
# ---------------------------------------------------------------------------

class _Plan1D:
    _kind = ""

    def __init__(self, length):
        L = lib()
        self.length = int(length)
        self.pocket = getattr(L, f"impulse_make_{self._kind}_plan")(self.length)
        if not self.pocket:
            raise RuntimeError(f"make_{self._kind}_plan failed: {L.impulse_fft_last_error().decode()}")
        # bound at creation: at interpreter exit the module globals may be gone
        self._destroy = getattr(L, f"impulse_destroy_{self._kind}_plan")

    def __del__(self):
        p, destroy = getattr(self, "pocket", None), getattr(self, "_destroy", None)
        if p and destroy is not None:
            destroy(p)
            self.pocket = None

    def call(self, data, forward, norm_value):
        fn = getattr(lib(), f"impulse_{self._kind}_{'forward' if forward else 'backward'}")
        err = fn(self.pocket, data.ctypes.data, float(norm_value))
        if err != 0:
            # the same messages as callFFT (c_pocketfft/pocketfft.nim:206-214)
            raise RuntimeError("Forward FFT calculation failed." if forward
                               else "Backward FFT calculation failed.")

    @classmethod
    def init(cls, length):
        return cls(length)


class FFTPlanReal(_Plan1D):
    _kind = "rfft"


class FFTPlanComplex(_Plan1D):
    _kind = "cfft"


def _as_inplace(data, dtype):
    if not isinstance(data, np.ndarray) or data.dtype != dtype or not data.flags.c_contiguous:
        raise TypeError(f"in-place FFT needs a C-contiguous numpy array of {np.dtype(dtype)}")
    return data


def rfft_inplace(data, forward, normalize=nkBackward, normValue=math.inf):
    """``rfft(data: MemoryView[float], length, forward, ...)``: in place, FFTPACK
    halfcomplex order (c_pocketfft/pocketfft.nim:216-288)."""
    data = _as_inplace(data, np.float64)
    norm = initNormalize(normalize, forward, normValue, data.size)
    FFTPlanReal.init(data.size).call(data, forward, norm.value)


def fft_impl(data, forward, normalize=nkBackward, normValue=math.inf):
    """In-place FFT of a float64 (packed result) or complex128 array
    (c_pocketfft/pocketfft.nim:290-303)."""
    if np.iscomplexobj(data):
        data = _as_inplace(data, np.complex128)
        plan = FFTPlanComplex.init(data.size)
    else:
        data = _as_inplace(data, np.float64)
        plan = FFTPlanReal.init(data.size)
    norm = initNormalize(normalize, forward, normValue, data.size)
    plan.call(data, forward, norm.value)


def fft_inplace(data, forward=True, normalize=nkBackward, normValue=math.inf):
    """``fft(data: var openArray/Tensor, forward)`` (pocketfft.nim:305-310,
    pocketfft_arraymancer.nim:33-37)."""
    fft_impl(data, forward, normalize, normValue)


def rfft_packed(data, forward=True, normalize=nkBackward, normValue=math.inf):
    """Real input -> FFTPACK halfcomplex array of the same length (pocketfft.nim:312-319)."""
    res = np.array(data, dtype=np.float64, copy=True).reshape(-1)
    rfft_inplace(res, forward, normalize, normValue)
    return res


def unpackFFT(data):
    """Halfcomplex -> the first n//2 + 1 complex bins (pocketfft.nim:126-158)."""
    data = np.asarray(data, dtype=np.float64).reshape(-1)
    n = data.size
    out_len = (n + 1) // 2 if n % 2 else (n + 2) // 2
    res = np.empty(out_len, np.complex128)
    res[0] = data[0]
    k = np.arange(1, out_len)
    re_idx = 2 * k - 1
    im_idx = 2 * k
    res.real[1:] = data[re_idx]
    res.imag[1:] = np.where(im_idx < n, data[np.minimum(im_idx, n - 1)], 0.0)
    return res


def symmTargetSize(data):
    """pocketfft.nim:173-183: real input keeps its length; complex input is
    taken as n//2+1 bins of an even (last imaginary part == 0) or odd signal."""
    data = np.asarray(data).reshape(-1)
    if not np.iscomplexobj(data):
        return data.size
    return data.size * 2 - (2 if data[-1].imag == 0.0 else 1)


def symmetrize(data):
    """Recover the Hermitian upper half (pocketfft.nim:160-187)."""
    data = np.asarray(data).reshape(-1)
    out_len = symmTargetSize(data)
    res = np.empty(out_len, np.complex128)
    if np.iscomplexobj(data):
        res[:data.size] = data
    else:
        u = unpackFFT(data)
        res[:u.size] = u
    i = np.arange(1, (out_len + 1) // 2)
    res[out_len - i] = np.conj(res[i])
    return res


def fft(data, forward=True, normalize=nkBackward, normValue=math.inf):
    """Out-of-place FFT of real or complex 1-D data (pocketfft.nim:321-347,
    pocketfft_arraymancer.nim:66-82)."""
    data = np.asarray(data)
    if np.iscomplexobj(data):
        res = np.array(data, dtype=np.complex128, copy=True).reshape(-1)
        fft_impl(res, forward, normalize, normValue)
        return res
    return symmetrize(rfft_packed(data, forward, normalize, normValue))


def rfft(data, forward=True, normalize=nkBackward, normValue=math.inf):
    """Real input -> n//2+1 bins (pocketfft.nim:321-331)."""
    res = np.array(data, dtype=np.float64, copy=True).reshape(-1)
    fft_impl(res, forward, normalize, normValue)
    return unpackFFT(res)


def ifft(data, backward=True, normalize=nkBackward, normValue=math.inf):
    """pocketfft.nim:349-360."""
    return fft(data, forward=not backward, normalize=normalize, normValue=normValue)


# ---------------------------------------------------------------------------
# C++ path (cpp_pocketfft/pocketfft.nim:137-277)
# ---------------------------------------------------------------------------

def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _dtype_of(buf):
    name = str(buf.dtype).replace("torch.", "")
    if name in ("float64", "complex128"):
        return _lib.F64
    if name in ("float32", "complex64"):
        return _lib.F32
    raise TypeError(f"unsupported element type {name}: the FFT takes float64/complex128 or float32/complex64")


class DataDesc:
    """Shape, BYTE strides and buffer of one array (cpp_pocketfft/pocketfft.nim:158-199).

    ``stride`` is in elements, as the Nim ``init`` takes it; it is converted to
    bytes.  Without ``stride`` the data is C-contiguous.  ``buf`` is a numpy
    array (host) or a torch tensor (GPU); ``shape``/``stride`` describe the
    view that starts at its first element."""

    def __init__(self, buf, shape, stride=None):
        if buf is None:
            raise ValueError("buffer is nil")
        shape = [int(s) for s in shape]
        if _is_torch(buf) and (buf.is_conj() or buf.is_neg()):
            # a lazy view: its memory holds the unconjugated / unnegated values
            raise ValueError("DataDesc takes a tensor's memory as it is: call resolve_conj() / resolve_neg() first")
        itemsize = buf.element_size() if _is_torch(buf) else buf.dtype.itemsize
        if stride is None:
            acc, st = itemsize, [0] * len(shape)
            for i in range(len(shape) - 1, -1, -1):
                st[i] = acc
                acc *= shape[i]
        else:
            if len(stride) != len(shape):
                raise ValueError("stride dimension mismatch")
            st = [int(s) * itemsize for s in stride]
        self.shape, self.stride, self.buf = shape, st, buf

    @classmethod
    def init(cls, buf, shape, stride=None):
        return cls(buf, shape, stride)

    @property
    def is_complex(self):
        if _is_torch(self.buf):
            return self.buf.is_complex()
        return np.iscomplexobj(self.buf)

    @property
    def dtype(self):
        """IMPULSE_FFT_F64 for double-based buffers, IMPULSE_FFT_F32 for float-based ones."""
        return _dtype_of(self.buf)

    def ptr(self):
        return self.buf.data_ptr() if _is_torch(self.buf) else self.buf.ctypes.data


class FFTDesc:
    """cpp_pocketfft/pocketfft.nim:201-215 (``nthreads`` is accepted and ignored)."""

    def __init__(self, axes, forward, scalingFactor=1.0, nthreads=1):
        self.axes = [int(a) for a in axes]
        self.forward = bool(forward)
        self.scalingFactor = float(scalingFactor)
        self.nthreads = int(nthreads)

    @classmethod
    def init(cls, axes, forward, scalingFactor=1.0, nthreads=1):
        return cls(axes, forward, scalingFactor, nthreads)


class DCTDesc:
    """cpp_pocketfft/pocketfft.nim:151-156,217-233: a DCT of ``dctType`` (1-4)
    over ``axes``; ``ortho`` and ``scalingFactor`` as pocketfft::dct takes them.
    ``sine=True`` selects the DST of the same type (pocketfft::dst,
    cpp_pocketfft/pocketfft.nim:121-132)."""

    def __init__(self, axes, dctType=2, ortho=False, scalingFactor=1.0, nthreads=1, sine=False):
        if not 1 <= int(dctType) <= 4:
            raise ValueError("dctType must be in 1..4")
        self.axes = [int(a) for a in axes]
        self.dctType = int(dctType)
        self.ortho = bool(ortho)
        self.scalingFactor = float(scalingFactor)
        self.nthreads = int(nthreads)
        self.sine = bool(sine)

    @classmethod
    def init(cls, axes, dctType=2, ortho=False, scalingFactor=1.0, nthreads=1, sine=False):
        return cls(axes, dctType, ortho, scalingFactor, nthreads, sine)


class Plan:
    """An n-D plan (impulse_fft_plan_create) for device execution.  Calls of one
    plan share its workspace, so they must not overlap on different streams."""

    def __init__(self, kind, shape, stride_in, stride_out, axes, forward, dtype=_lib.F64, r2r_type=2, ortho=False):
        h = ctypes.c_void_p()
        if kind in (_lib.DCT, _lib.DST):
            check(lib().impulse_fft_r2r_plan_create(ctypes.byref(h), int(kind), int(r2r_type), int(bool(ortho)),
                                                    int(dtype), len(shape), sizes(shape), strides(stride_in),
                                                    strides(stride_out), len(axes), sizes(axes)))
        else:
            check(lib().impulse_fft_plan_create(ctypes.byref(h), int(kind), int(dtype), len(shape), sizes(shape),
                                                strides(stride_in), strides(stride_out), len(axes), sizes(axes),
                                                1 if forward else 0))
        self.handle = h
        self._destroy = lib().impulse_fft_plan_destroy
        self.kind = kind

    def __del__(self):
        # bound at creation: at interpreter exit the module globals may be gone
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self.handle = None

    def execute(self, in_ptr, out_ptr, fct=1.0, stream=0):
        check(lib().impulse_fft_execute(self.handle, in_ptr, out_ptr, float(fct), stream or None))

    def execute_host(self, in_arr, out_arr, fct=1.0):
        check(lib().impulse_fft_execute_host(self.handle, in_arr.ctypes.data, out_arr.ctypes.data, float(fct)))

    @property
    def num_launches(self):
        return lib().impulse_fft_plan_num_launches(self.handle)

    @property
    def device_bytes(self):
        return lib().impulse_fft_plan_device_bytes(self.handle)


PLAN_CACHE_CAPACITY = 16
_plan_cache = collections.OrderedDict()
_plan_lock = threading.Lock()


def get_plan(kind, shape, stride_in, stride_out, axes, forward, r2r_type=2, ortho=False, dtype=_lib.F64,
             device=None, stream=None):
    """Plan for a device request, from an LRU of PLAN_CACHE_CAPACITY entries keyed
    by the request, the device and the stream (one workspace per stream)."""
    key = (kind, int(dtype), tuple(shape), tuple(stride_in), tuple(stride_out), tuple(axes), bool(forward),
           int(r2r_type), bool(ortho), device, stream)
    with _plan_lock:
        p = _plan_cache.get(key)
        if p is not None:
            _plan_cache.move_to_end(key)
            return p
    p = Plan(kind, shape, stride_in, stride_out, axes, forward, dtype=dtype, r2r_type=r2r_type, ortho=ortho)
    with _plan_lock:
        q = _plan_cache.get(key)
        if q is not None:      # another thread built it meanwhile
            _plan_cache.move_to_end(key)
            return q
        _plan_cache[key] = p
        while len(_plan_cache) > PLAN_CACHE_CAPACITY:
            _plan_cache.popitem(last=False)
    return p


def _current_stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def _device_plan(kind, shape, descIn, descOut, axes, forward, dtype, r2r_type=2, ortho=False):
    import torch
    dev = descIn.buf.device.index if descIn.buf.device.index is not None else torch.cuda.current_device()
    stream = torch.cuda.current_stream(dev).cuda_stream
    p = get_plan(kind, shape, descIn.stride, descOut.stride, axes, forward, r2r_type, ortho, dtype, dev, stream)
    return p, stream


def apply(fft, descOut, descIn):
    """Dispatch c2c / r2c / c2r by element types (cpp_pocketfft/pocketfft.nim:235-277),
    or a DCT / DST for a DCTDesc (cpp_pocketfft/pocketfft.nim:279-295, which passes
    ``descIn.shape``)."""
    if isinstance(fft, DCTDesc):
        if descIn.is_complex or descOut.is_complex:
            raise TypeError("DCT / DST take real arrays")
        kind = _lib.DST if fft.sine else _lib.DCT
        if _is_torch(descIn.buf) and descIn.buf.is_cuda:
            p, stream = _device_plan(kind, descIn.shape, descIn, descOut, fft.axes, True, _lib.F64, fft.dctType,
                                     fft.ortho)
            p.execute(descIn.ptr(), descOut.ptr(), fft.scalingFactor, stream)
            return
        fn = lib().impulse_fft_dst if fft.sine else lib().impulse_fft_dct
        check(fn(len(descIn.shape), sizes(descIn.shape), strides(descIn.stride), strides(descOut.stride),
                 len(fft.axes), sizes(fft.axes), fft.dctType, _lib.F64, descIn.ptr(), descOut.ptr(),
                 fft.scalingFactor, 1 if fft.ortho else 0, fft.nthreads))
        return
    if descIn.is_complex and descOut.is_complex:
        kind, shape = _lib.C2C, descIn.shape
    elif descOut.is_complex:
        kind, shape = _lib.R2C, descIn.shape
    elif descIn.is_complex:
        kind, shape = _lib.C2R, descOut.shape
    else:
        raise TypeError("Not implemented")
    dtype = descIn.dtype
    if descOut.dtype != dtype:
        raise TypeError("input and output must have the same precision (float64/complex128 or float32/complex64)")
    if _is_torch(descIn.buf) and descIn.buf.is_cuda:
        p, stream = _device_plan(kind, shape, descIn, descOut, fft.axes, fft.forward, dtype)
        p.execute(descIn.ptr(), descOut.ptr(), fft.scalingFactor, stream)
        return
    fn = {_lib.C2C: lib().impulse_fft_c2c, _lib.R2C: lib().impulse_fft_r2c, _lib.C2R: lib().impulse_fft_c2r}[kind]
    check(fn(len(shape), sizes(shape), strides(descIn.stride), strides(descOut.stride), len(fft.axes),
             sizes(fft.axes), 1 if fft.forward else 0, dtype, descIn.ptr(), descOut.ptr(),
             fft.scalingFactor, fft.nthreads))


# ---------------------------------------------------------------------------
# pocketfft-style n-D helpers on numpy arrays (byte strides taken from numpy)
# ---------------------------------------------------------------------------

_COMPLEX_OF = {np.dtype(np.float64): np.complex128, np.dtype(np.float32): np.complex64}
_REAL_OF = {np.dtype(np.complex128): np.float64, np.dtype(np.complex64): np.float32}
_DT = {np.dtype(np.float64): _lib.F64, np.dtype(np.complex128): _lib.F64, np.dtype(np.float32): _lib.F32,
       np.dtype(np.complex64): _lib.F32}


def c2c(x, axes, forward=True, fct=1.0, out=None):
    """pocketfft::c2c on a complex128 or complex64 numpy array (any byte strides)."""
    x = np.asarray(x)
    if x.dtype not in _REAL_OF:
        raise TypeError("c2c needs complex128 or complex64")
    if out is None:
        out = np.empty(x.shape, x.dtype)
    check(lib().impulse_fft_c2c(x.ndim, sizes(x.shape), strides(x.strides), strides(out.strides), len(axes),
                                sizes(axes), 1 if forward else 0, _DT[x.dtype], x.ctypes.data, out.ctypes.data,
                                float(fct), 1))
    return out


def r2c(x, axes, forward=True, fct=1.0, out=None):
    """pocketfft::r2c on a float64 or float32 numpy array (any byte strides)."""
    x = np.asarray(x)
    if x.dtype not in _COMPLEX_OF:
        raise TypeError("r2c needs float64 or float32")
    axes = list(axes)
    oshape = list(x.shape)
    oshape[axes[-1]] = x.shape[axes[-1]] // 2 + 1
    if out is None:
        out = np.empty(oshape, _COMPLEX_OF[x.dtype])
    check(lib().impulse_fft_r2c(x.ndim, sizes(x.shape), strides(x.strides), strides(out.strides), len(axes),
                                sizes(axes), 1 if forward else 0, _DT[x.dtype], x.ctypes.data, out.ctypes.data,
                                float(fct), 1))
    return out


def c2r(x, n, axes, forward=False, fct=1.0, out=None):
    """pocketfft::c2r on a complex128 or complex64 numpy array; ``n`` is the real length along axes[-1]."""
    x = np.asarray(x)
    if x.dtype not in _REAL_OF:
        raise TypeError("c2r needs complex128 or complex64")
    axes = list(axes)
    oshape = list(x.shape)
    oshape[axes[-1]] = int(n)
    if out is None:
        out = np.empty(oshape, _REAL_OF[x.dtype])
    check(lib().impulse_fft_c2r(len(oshape), sizes(oshape), strides(x.strides), strides(out.strides), len(axes),
                                sizes(axes), 1 if forward else 0, _DT[x.dtype], x.ctypes.data, out.ctypes.data,
                                float(fct), 1))
    return out


def dct(x, axes, type=2, fct=1.0, ortho=False, out=None):
    """pocketfft::dct on a float64 numpy array (any byte strides)."""
    return _r2r(lib().impulse_fft_dct, x, axes, type, fct, ortho, out)


def dst(x, axes, type=2, fct=1.0, ortho=False, out=None):
    """pocketfft::dst on a float64 numpy array (any byte strides)."""
    return _r2r(lib().impulse_fft_dst, x, axes, type, fct, ortho, out)


def _r2r(fn, x, axes, type, fct, ortho, out):
    x = np.asarray(x)
    if x.dtype != np.float64:
        raise TypeError("DCT / DST need float64")
    if out is None:
        out = np.empty(x.shape, np.float64)
    check(fn(x.ndim, sizes(x.shape), strides(x.strides), strides(out.strides), len(axes), sizes(axes), int(type),
             _lib.F64, x.ctypes.data, out.ctypes.data, float(fct), 1 if ortho else 0, 1))
    return out


# ---------------------------------------------------------------------------
# r2r_fftpack and the Hartley transforms (pocketfft_hdronly.h:3392-3445)
# ---------------------------------------------------------------------------

def _rfp_host(fn, x, axes, extra, fct, out):
    x = np.asarray(x)
    if x.dtype != np.float64:
        raise TypeError("real-to-real transforms need float64")
    if out is None:
        out = np.empty(x.shape, np.float64)
    check(fn(x.ndim, sizes(x.shape), strides(x.strides), strides(out.strides), len(axes), sizes(axes), *extra,
             _lib.F64, x.ctypes.data, out.ctypes.data, float(fct), 1))
    return out


def r2r_fftpack(x, axes, real2hermitian, forward, fct=1.0, out=None):
    """pocketfft::r2r_fftpack on a float64 numpy array: along each axis, `forward` transforms real data to
    FFTPACK halfcomplex order and backward transforms halfcomplex data back; `real2hermitian` adds the
    reference's sign flips of the imaginary slots (input of a forward call when false, output of a backward
    call when true)."""
    return _rfp_host(lib().impulse_fft_r2r_fftpack, x, axes, [int(bool(real2hermitian)), int(bool(forward))], fct,
                     out)


def r2r_separable_hartley(x, axes, fct=1.0, out=None):
    """pocketfft::r2r_separable_hartley: H = Re X + Im X along each axis in turn."""
    return _rfp_host(lib().impulse_fft_r2r_separable_hartley, x, axes, [], fct, out)


def r2r_genuine_hartley(x, axes, fct=1.0, out=None):
    """pocketfft::r2r_genuine_hartley: H = Re X + Im X of the DFT over all axes at once."""
    return _rfp_host(lib().impulse_fft_r2r_genuine_hartley, x, axes, [], fct, out)


class FftpackPlan:
    """A device plan for r2r_fftpack (kind FFTPACK) or the Hartley transforms (HARTLEY, HARTLEY_GENUINE)
    over torch device tensors, with byte strides like Plan."""

    def __init__(self, kind, shape, stride_in, stride_out, axes, real2hermitian=True, forward=True):
        h = ctypes.c_void_p()
        check(lib().impulse_fft_fftpack_plan_create(ctypes.byref(h), int(kind), int(bool(real2hermitian)),
                                                    int(bool(forward)), _lib.F64, len(shape), sizes(shape),
                                                    strides(stride_in), strides(stride_out), len(axes), sizes(axes)))
        self.handle = h
        self._destroy = lib().impulse_fft_plan_destroy

    def __del__(self):
        # bound at creation: at interpreter exit the module globals may be gone
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self.handle = None

    def execute(self, in_ptr, out_ptr, fct=1.0, stream=0):
        check(lib().impulse_fft_execute(self.handle, in_ptr, out_ptr, float(fct), stream or None))

    @property
    def num_launches(self):
        return lib().impulse_fft_plan_num_launches(self.handle)

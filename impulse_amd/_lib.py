"""ctypes binding of the C ABI declared in include/impulse_fft.h.

The native library (``impulse_amd/libimpulse_fft.so``, HIP for gfx950) is the
only compute path: if it is missing, ``lib()`` raises; there is no CPU
fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# IMPULSE_FFT_LIB may point at another build of the same library (A/B timing)
LIB_PATH = os.environ.get("IMPULSE_FFT_LIB") or os.path.join(_HERE, "libimpulse_fft.so")

# Every function declared in include/impulse_fft.h.
EXPORTS = (
    "impulse_fft_plan_create", "impulse_fft_execute", "impulse_fft_execute_host",
    "impulse_fft_execute_noalloc",
    "impulse_fft_plan_device_bytes", "impulse_fft_plan_num_launches", "impulse_fft_plan_pass_info",
    "impulse_fft_plan_profile", "impulse_fft_plan_profile_read", "impulse_fft_plan_destroy",
    "impulse_fft_c2c", "impulse_fft_r2c", "impulse_fft_c2r",
    "impulse_fft_cache_set_capacity", "impulse_fft_cache_size", "impulse_fft_cache_clear",
    "impulse_fft_cache_stats",
    "impulse_make_cfft_plan", "impulse_destroy_cfft_plan", "impulse_cfft_forward",
    "impulse_cfft_backward", "impulse_cfft_length",
    "impulse_make_rfft_plan", "impulse_destroy_rfft_plan", "impulse_rfft_forward",
    "impulse_rfft_backward", "impulse_rfft_length",
    "impulse_fft_status_string", "impulse_fft_last_error", "impulse_fft_device_check",
    "impulse_fft_conv2d_create", "impulse_fft_conv2d_forward", "impulse_fft_conv2d_backward",
    "impulse_fft_conv2d_grid", "impulse_fft_conv2d_destroy",
    "impulse_fft_ragged_create", "impulse_fft_ragged_execute", "impulse_fft_ragged_tune",
    "impulse_fft_ragged_configure", "impulse_fft_ragged_info", "impulse_fft_ragged_destroy",
    "impulse_fft_ragged_set_precision",
    "impulse_fft_mfma_create", "impulse_fft_mfma_execute", "impulse_fft_mfma_workspace_bytes",
    "impulse_fft_mfma_destroy",
    "impulse_fft_dct", "impulse_fft_dst", "impulse_fft_r2r_plan_create",
    "impulse_fft_upfirdn_create", "impulse_fft_upfirdn_info", "impulse_fft_upfirdn_execute",
    "impulse_fft_upfirdn_execute_host", "impulse_fft_upfirdn_destroy",
    "impulse_fft_r2r_fftpack", "impulse_fft_r2r_separable_hartley", "impulse_fft_r2r_genuine_hartley",
    "impulse_fft_fftpack_plan_create",
)

C2C, R2C, C2R, R2C_ADJOINT, C2R_ADJOINT, DCT, DST, FFTPACK, HARTLEY, HARTLEY_GENUINE = range(10)
F64, F32 = 0, 1
MFMA_F32, MFMA_BF16X = 0, 1
UPFIRDN_AUTO, UPFIRDN_DIRECT, UPFIRDN_FFT = 0, 1, 2

_szp = ctypes.POINTER(ctypes.c_size_t)
_pdp = ctypes.POINTER(ctypes.c_ssize_t)
_vp = ctypes.c_void_p
_LIB = None


def _sig(L):
    i, sz, d = ctypes.c_int, ctypes.c_size_t, ctypes.c_double
    L.impulse_fft_plan_create.argtypes = [ctypes.POINTER(_vp), i, i, sz, _szp, _pdp, _pdp, sz, _szp, i]
    L.impulse_fft_execute.argtypes = [_vp, _vp, _vp, d, _vp]
    L.impulse_fft_execute_host.argtypes = [_vp, _vp, _vp, d]
    L.impulse_fft_execute_noalloc.argtypes = [_vp, _vp, _vp, d, _vp]
    L.impulse_fft_plan_device_bytes.argtypes = [_vp]
    L.impulse_fft_plan_device_bytes.restype = sz
    L.impulse_fft_plan_num_launches.argtypes = [_vp]
    L.impulse_fft_plan_pass_info.argtypes = [_vp, i, ctypes.c_char_p, sz, _szp, _szp,
                                             ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint), _szp]
    L.impulse_fft_plan_profile.argtypes = [_vp, i]
    L.impulse_fft_plan_profile_read.argtypes = [_vp, ctypes.POINTER(d), i, ctypes.POINTER(i)]
    L.impulse_fft_plan_destroy.argtypes = [_vp]
    for n in ("impulse_fft_c2c", "impulse_fft_r2c", "impulse_fft_c2r"):
        getattr(L, n).argtypes = [sz, _szp, _pdp, _pdp, sz, _szp, i, i, _vp, _vp, d, sz]
    L.impulse_fft_cache_set_capacity.argtypes = [sz]
    L.impulse_fft_cache_size.restype = sz
    L.impulse_fft_cache_stats.argtypes = [_szp, _szp]
    for kind in ("cfft", "rfft"):
        getattr(L, f"impulse_make_{kind}_plan").argtypes = [sz]
        getattr(L, f"impulse_make_{kind}_plan").restype = _vp
        getattr(L, f"impulse_destroy_{kind}_plan").argtypes = [_vp]
        getattr(L, f"impulse_destroy_{kind}_plan").restype = None
        for d_ in ("forward", "backward"):
            getattr(L, f"impulse_{kind}_{d_}").argtypes = [_vp, _vp, d]
        getattr(L, f"impulse_{kind}_length").argtypes = [_vp]
        getattr(L, f"impulse_{kind}_length").restype = sz
    L.impulse_fft_conv2d_create.argtypes = [ctypes.POINTER(_vp), i, i, i, i, i, i]
    L.impulse_fft_conv2d_forward.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.impulse_fft_conv2d_backward.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.impulse_fft_conv2d_grid.argtypes = [_vp, _szp, _szp]
    L.impulse_fft_conv2d_destroy.argtypes = [_vp]
    L.impulse_fft_ragged_create.argtypes = [ctypes.POINTER(_vp), sz, _szp, _szp, _szp, i]
    L.impulse_fft_ragged_execute.argtypes = [_vp, _vp, _vp, d, _vp]
    L.impulse_fft_ragged_tune.argtypes = [_vp, _vp, _vp, _vp, i, ctypes.POINTER(d)]
    L.impulse_fft_ragged_configure.argtypes = [_vp, i, i, i, i]
    L.impulse_fft_ragged_set_precision.argtypes = [_vp, i]
    ip = ctypes.POINTER(i)
    L.impulse_fft_ragged_info.argtypes = [_vp, ip, ip, ip, ip, ip, ctypes.POINTER(d), _szp]
    L.impulse_fft_ragged_destroy.argtypes = [_vp]
    L.impulse_fft_mfma_create.argtypes = [ctypes.POINTER(_vp), sz, sz, i]
    L.impulse_fft_mfma_execute.argtypes = [_vp, _vp, _vp, i, d, _vp]
    L.impulse_fft_mfma_workspace_bytes.argtypes = [_vp, _szp]
    L.impulse_fft_mfma_destroy.argtypes = [_vp]
    for n in ("impulse_fft_dct", "impulse_fft_dst"):
        getattr(L, n).argtypes = [sz, _szp, _pdp, _pdp, sz, _szp, i, i, _vp, _vp, d, i, sz]
    L.impulse_fft_r2r_plan_create.argtypes = [ctypes.POINTER(_vp), i, i, i, i, sz, _szp, _pdp, _pdp, sz, _szp]
    L.impulse_fft_r2r_fftpack.argtypes = [sz, _szp, _pdp, _pdp, sz, _szp, i, i, i, _vp, _vp, d, sz]
    for n in ("impulse_fft_r2r_separable_hartley", "impulse_fft_r2r_genuine_hartley"):
        getattr(L, n).argtypes = [sz, _szp, _pdp, _pdp, sz, _szp, i, _vp, _vp, d, sz]
    L.impulse_fft_fftpack_plan_create.argtypes = [ctypes.POINTER(_vp), i, i, i, i, sz, _szp, _pdp, _pdp, sz, _szp]
    L.impulse_fft_upfirdn_create.argtypes = [ctypes.POINTER(_vp), sz, sz, _vp, sz, i, i, i]
    L.impulse_fft_upfirdn_info.argtypes = [_vp, _szp, ctypes.POINTER(i), _szp, _szp]
    L.impulse_fft_upfirdn_execute.argtypes = [_vp, _vp, _vp, _vp]
    L.impulse_fft_upfirdn_execute_host.argtypes = [_vp, _vp, _vp]
    L.impulse_fft_upfirdn_destroy.argtypes = [_vp]
    L.impulse_fft_status_string.argtypes = [i]
    L.impulse_fft_status_string.restype = ctypes.c_char_p
    L.impulse_fft_last_error.restype = ctypes.c_char_p


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"impulse_amd native library missing: {LIB_PATH}; build it with "
                "`make -C impulse_amd/csrc` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        _sig(L)
        _LIB = L
    return _LIB


class FFTError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        text = lib().impulse_fft_status_string(status).decode()
        super().__init__(f"{text} ({status}): {message}" if message else f"{text} ({status})")


def check(status):
    if status != 0:
        raise FFTError(status, lib().impulse_fft_last_error().decode())
    return status


def sizes(vals):
    return (ctypes.c_size_t * len(vals))(*[int(v) for v in vals])


def strides(vals):
    return (ctypes.c_ssize_t * len(vals))(*[int(v) for v in vals])

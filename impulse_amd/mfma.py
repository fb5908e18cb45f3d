"""Matrix-core complex FFTs over the C ABI (impulse_fft_mfma_*).

The radix-16 DFT stages of the 256-point factors of complex64 transforms
(lengths: the powers of two from 256 to 2^21) run as complex 16x16x16 matrix
products on the MFMA units, in one of two precisions:

* ``"f32"``   -- FP32 MFMA with FP32 accumulation; complex64 output.
* ``"bf16x"`` -- bf16 slices accumulated exactly by the MFMA units and
  combined in FP64; complex128 output.

Nearest reference: pocketfft::c2c<float> over one axis
(pocketfft_hdronly.h:3272-3282).
"""
import ctypes

from . import _lib
from ._lib import check, lib

PRECISIONS = {"f32": _lib.MFMA_F32, "bf16x": _lib.MFMA_BF16X}


class MfmaFFT:
    """`batch` contiguous complex64 transforms of `length` (a power of two, 256 .. 2^21)."""

    def __init__(self, length, batch, precision="bf16x"):
        self.length, self.batch, self.precision = int(length), int(batch), precision
        h = ctypes.c_void_p()
        check(lib().impulse_fft_mfma_create(ctypes.byref(h), self.length, self.batch, PRECISIONS[precision]))
        self.handle = h
        self._destroy = lib().impulse_fft_mfma_destroy

    def __del__(self):
        # the destroy entry point is bound at creation: at interpreter exit the
        # module globals may already be gone
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self.handle = None

    @property
    def out_dtype(self):
        import torch
        return torch.complex64 if self.precision == "f32" else torch.complex128

    @property
    def workspace_bytes(self):
        b = ctypes.c_size_t()
        check(lib().impulse_fft_mfma_workspace_bytes(self.handle, ctypes.byref(b)))
        return b.value

    def __call__(self, x, out=None, forward=True, fct=1.0):
        """x: complex64 CUDA tensor with batch * length elements (contiguous)."""
        import torch
        if not x.is_cuda or x.dtype != torch.complex64 or not x.is_contiguous():
            raise TypeError("input must be a contiguous complex64 CUDA tensor")
        if x.numel() != self.batch * self.length:
            raise ValueError(f"input has {x.numel()} elements, plan expects {self.batch * self.length}")
        if out is None:
            out = torch.empty(x.shape, dtype=self.out_dtype, device=x.device)
        if out.dtype != self.out_dtype or not out.is_contiguous() or out.numel() != x.numel():
            raise TypeError("output must be contiguous, of the plan's output dtype and the input's size")
        check(lib().impulse_fft_mfma_execute(self.handle, x.data_ptr(), out.data_ptr(), int(bool(forward)),
                                             float(fct), torch.cuda.current_stream().cuda_stream))
        return out

"""Ragged batches of real transforms over the C ABI (impulse_fft_ragged_*).

A RaggedBatch holds the lengths and offsets of many float32 signals packed in
one device buffer and transforms all of them with one hipGraph launch; each
signal's lengths[i]//2 + 1 complex128 bins land at out_offsets[i] in one
output buffer.  The bins are those of PocketFFT's r2c on the signal widened to
float64 (pocketfft_hdronly.h:3320-3350, one call per signal as impulse's C++
path would make it through cpp_pocketfft/pocketfft.nim:235-277).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib, sizes


def packed_offsets(lengths):
    """Input offsets (floats) and output offsets (complex bins) of signals packed back to back."""
    n = np.asarray(lengths, dtype=np.int64)
    ins = np.zeros(len(n), dtype=np.int64)
    outs = np.zeros(len(n), dtype=np.int64)
    if len(n):
        ins[1:] = np.cumsum(n)[:-1]
        outs[1:] = np.cumsum(n // 2 + 1)[:-1]
    return ins, outs


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


class RaggedBatch:
    """All signals of one ragged batch; create once, execute many times."""

    def __init__(self, lengths, in_offsets=None, out_offsets=None, forward=True):
        self.lengths = np.asarray(lengths, dtype=np.int64)
        if in_offsets is None or out_offsets is None:
            pi, po = packed_offsets(self.lengths)
            in_offsets = pi if in_offsets is None else in_offsets
            out_offsets = po if out_offsets is None else out_offsets
        self.in_offsets = np.asarray(in_offsets, dtype=np.int64)
        self.out_offsets = np.asarray(out_offsets, dtype=np.int64)
        self.in_elems = int(np.max(self.in_offsets + self.lengths)) if len(self.lengths) else 0
        self.out_elems = int(np.max(self.out_offsets + self.lengths // 2 + 1)) if len(self.lengths) else 0
        h = ctypes.c_void_p()
        check(lib().impulse_fft_ragged_create(ctypes.byref(h), len(self.lengths), sizes(self.lengths),
                                              sizes(self.in_offsets), sizes(self.out_offsets), int(bool(forward))))
        self.handle = h
        self._destroy = lib().impulse_fft_ragged_destroy
        self.precision = "fp64"

    def __del__(self):
        # the destroy entry point is bound at creation: at interpreter exit the
        # module globals may already be gone
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self.handle = None

    def _check_tensors(self, x, y):
        import torch
        if not (x.is_cuda and y.is_cuda):
            raise ValueError("ragged batches take device tensors")
        if x.dtype != torch.float32 or y.dtype != torch.complex128:
            raise TypeError("ragged input is float32, output complex128")
        if not (x.is_contiguous() and y.is_contiguous()):
            raise ValueError("ragged buffers must be contiguous")
        if x.numel() < self.in_elems or y.numel() < self.out_elems:
            raise ValueError(f"buffers too small: need {self.in_elems} floats in, {self.out_elems} bins out")

    def execute(self, x, y, fct=1.0):
        """Transform every signal of x (float32, device) into y (complex128, device)."""
        self._check_tensors(x, y)
        check(lib().impulse_fft_ragged_execute(self.handle, x.data_ptr(), y.data_ptr(), float(fct), _stream()))
        return y

    def tune(self, x, y, reps=3):
        """Time every candidate configuration on these buffers and keep the fastest; returns ms."""
        self._check_tensors(x, y)
        ms = ctypes.c_double()
        check(lib().impulse_fft_ragged_tune(self.handle, x.data_ptr(), y.data_ptr(), _stream(), int(reps),
                                            ctypes.byref(ms)))
        return ms.value

    def configure(self, threads, lds_points, radix_order=0, waves_per_simd=0):
        """Pick a configuration by hand (impulse_fft_ragged_configure): threads 256/512/1024,
        lds_points 2048/4096/8192 (at most 8 per thread), radix_order 0-3 (bit 0: odd radices
        before the powers of two; bit 1: powers of two as radix-8 stages), waves_per_simd 0, 4,
        or 3 with 256 threads."""
        check(lib().impulse_fft_ragged_configure(self.handle, int(threads), int(lds_points), int(radix_order),
                                                 int(waves_per_simd)))

    PRECISIONS = {"fp64": 0, "bf16x": 1}

    def set_precision(self, precision):
        """Arithmetic of the LDS stages (impulse_fft_ragged_set_precision): "fp64" (every stage in
        FP64, the default) or "bf16x" (radix-16 stages of transforms spanning whole wavefronts as
        exact bf16-slice products on the matrix cores, combined in FP64; ~1e-12 relative)."""
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision is one of {sorted(self.PRECISIONS)}")
        check(lib().impulse_fft_ragged_set_precision(self.handle, self.PRECISIONS[precision]))
        self.precision = precision

    def info(self):
        t, l, o, w, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        f, sb = ctypes.c_double(), ctypes.c_size_t()
        check(lib().impulse_fft_ragged_info(self.handle, ctypes.byref(t), ctypes.byref(l), ctypes.byref(o),
                                            ctypes.byref(w), ctypes.byref(g), ctypes.byref(f), ctypes.byref(sb)))
        return {"threads": t.value, "lds_points": l.value, "radix_order": o.value, "waves_per_simd": w.value,
                "grid": g.value, "flops": f.value, "scratch_bytes": sb.value}

    def output_views(self, y):
        """Per-signal views of the output buffer."""
        return [y[int(o):int(o) + int(n) // 2 + 1] for o, n in zip(self.out_offsets, self.lengths)]

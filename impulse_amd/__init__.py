"""impulse_amd -- MI355X (gfx950) FFT engine for impulse.

The compute path is the HIP library ``libimpulse_fft.so`` behind the C ABI in
``include/impulse_fft.h``; this package is its host-side mirror of impulse's
FFT API (``impulse_amd.fft``) plus a seeded input generator
(``impulse_amd.synth``).
"""
from . import _lib  # noqa: F401
from ._lib import FFTError, LIB_PATH, lib  # noqa: F401

__all__ = ["FFTError", "LIB_PATH", "lib", "fft", "synth"]

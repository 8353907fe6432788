"""fftfree_amd: the batched FFT/STFT path of sangderenard/fftfree on MI355X (gfx950).

The product is ``libfft_cffi.so`` (C ABI in include/fft_cffi.h). This package holds its
HIP/C++ sources (csrc/), the build recipe, a ctypes front end shaped like the
reference's CFFI callers, and the torch autograd wrapper of the STFT adjoint.
"""
from . import _lib  # noqa: F401
from .cffi_fft import CffiFft, FftPlan  # noqa: F401

__all__ = ["CffiFft", "FftPlan", "load"]


def load():
    """Load libfft_cffi.so (raises ImportError when it is not built)."""
    return _lib.load()

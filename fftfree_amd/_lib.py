"""Loader of libfft_cffi.so with the exact C signatures of include/fft_cffi.h.

The reference's Python tools bind the library over CFFI (train_mipmap_weights.py:97-160,
fft_to_png.py:43-93); the ``cffi`` module is not available here, so this binding uses
ctypes with the same signatures as the C header. Note that the reference tools' cdefs
carry an extra ``int stft_mode`` after ``hop`` that the C header does not have; the
header is authoritative and this binding follows it.

There is no fallback: if the library is missing or fails to load, importing raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(os.environ.get("FFTFREE_CFFI_LIB", Path(__file__).resolve().parent / "libfft_cffi.so"))

c_int, c_float, c_size_t, c_void_p, c_char_p = ctypes.c_int, ctypes.c_float, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p
c_uint = ctypes.c_uint
f32p = ctypes.POINTER(ctypes.c_float)
i32p = ctypes.POINTER(ctypes.c_int)

class RaggedRequest(ctypes.Structure):
    """fft_ragged_request (include/fft_cffi.h)."""
    _fields_ = [("pcm", c_void_p), ("pcm_len", c_size_t), ("window", c_int), ("hop", c_int),
                ("out_real", c_void_p), ("out_imag", c_void_p), ("out_mag", c_void_p), ("frames", c_size_t)]


# name -> (restype, argtypes), in header order
SIGNATURES = {
    "fft_pcm_to_channels": (None, [f32p, f32p, f32p, f32p, c_size_t, c_int]),
    "fft_init": (c_void_p, [c_size_t, c_int, c_int, c_int, c_int, c_int, i32p, c_size_t, c_int]),
    "fft_init_ex": (c_void_p, [c_size_t, c_int, c_int, c_int, c_int, c_int, i32p, c_size_t, c_int, c_int, c_int]),
    "fft_init_full": (c_void_p, [c_size_t, c_int, c_int, c_int, c_int, c_int, i32p, c_size_t] + [c_int] * 12),
    "fft_init_full_v2": (c_void_p, [c_size_t, c_int, c_int, c_int, c_int, c_int, i32p, c_size_t] + [c_int] * 12
                         + [c_int, c_int, c_int, c_float, c_float, c_int, c_float, c_float, c_int, c_int]),
    "fft_execute": (c_int, [c_void_p, f32p, f32p, f32p, f32p, c_size_t]),
    "fft_execute_batched": (c_size_t, [c_void_p, f32p, c_size_t, f32p, f32p, f32p, c_int, c_int, c_size_t]),
    "fft_stft_execute": (c_size_t, [c_void_p, f32p, c_size_t, f32p, f32p, f32p, c_size_t]),
    "fft_execute_complex_batched": (c_size_t, [c_void_p, f32p, f32p, c_size_t, f32p, c_int, c_int, c_size_t]),
    "fft_ctx_size": (c_size_t, [c_void_p]),
    "fft_ctx_worker_threads": (c_size_t, [c_void_p]),
    "fft_ctx_effective_threads": (c_size_t, [c_void_p]),
    "fft_free": (None, [c_void_p]),
    "fft_config_window": (c_int, [c_void_p, c_int, c_int, c_float, c_float, c_int, c_int]),
    "fft_config_window_custom": (c_int, [c_void_p, f32p, c_size_t, f32p, c_size_t, c_int, c_int]),
    "fft_set_workerpool_debug_dropout": (None, [c_int]),
    "fft_set_workerpool_dropout_pattern": (None, [c_int, c_int, c_int]),
    "fft_set_workerpool_dropout_single": (None, [ctypes.c_longlong]),
    "fft_set_workerpool_dropout_persistent": (None, [c_int]),
    "fft_clear_workerpool_dropout_history": (None, []),
    "fft_register_restore_observer": (None, [c_void_p]),
    "fft_frame_count": (c_size_t, [c_void_p, c_size_t, c_int, c_size_t]),
    "fft_execute_batched_device": (c_size_t, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_int,
                                              c_size_t, c_void_p]),
    "fft_execute_c2c_batched": (c_size_t, [c_void_p, f32p, f32p, c_size_t]),
    "fft_execute_c2c_batched_device": (c_size_t, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "fft_execute_complex_batched_device": (c_size_t, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t,
                                                      c_void_p]),
    "ft_apply_complex_pipeline": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, f32p, f32p, c_int, c_int]),
    "ft_apply_polar_pipeline": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, f32p, f32p, c_int, c_int]),
    "ft_apply_complex_pipeline_device": (c_int, [c_void_p] * 8 + [c_int, c_int, c_void_p]),
    "ft_apply_polar_pipeline_device": (c_int, [c_void_p] * 8 + [c_int, c_int, c_void_p]),
    "phase_init": (c_void_p, [c_int, c_int, c_int, c_int, c_int]),
    "phase_infer_execute": (c_size_t, [c_void_p, f32p, c_size_t, f32p, f32p]),
    "phase_free": (None, [c_void_p]),
    "phase_infer_execute_device": (c_size_t, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p]),
    "fft_stream_reset": (None, [c_void_p]),
    "fft_stream_push_pcm": (c_size_t, [c_void_p, f32p, c_size_t, f32p, f32p, f32p, c_size_t, c_int]),
    "fft_stream_pending_frames": (c_size_t, [c_void_p]),
    "fft_stream_backlog_samples": (c_size_t, [c_void_p]),
    "fft_stream_push_spectrum": (c_size_t, [c_void_p, f32p, f32p, c_size_t, f32p, c_size_t, c_int]),
    "fft_stream_pending_pcm": (c_size_t, [c_void_p]),
    "fft_griffin_lim": (c_int, [c_void_p, c_void_p, f32p, c_size_t, c_int, c_int, c_int, c_int, c_uint, f32p, f32p]),
    "fft_griffin_lim_device": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_int, c_int, c_int, c_int, c_uint,
                                       c_void_p, c_void_p, c_void_p]),
    "fft_stft_signals_device": (c_size_t, [c_void_p, c_void_p, c_size_t, c_size_t, c_void_p, c_void_p, c_void_p,
                                           c_int, c_void_p]),
    "fft_stft_adjoint_device": (c_size_t, [c_void_p, c_void_p, c_size_t, c_size_t, c_void_p, c_void_p, c_void_p,
                                           c_int, c_void_p]),
    "fft_stft_adjoint_device_noalloc": (c_size_t, [c_void_p, c_void_p, c_size_t, c_size_t, c_void_p, c_void_p,
                                                   c_void_p, c_int, c_void_p]),
    "fft_stft_adjoint_saved_device": (c_size_t, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t, c_void_p,
                                                 c_void_p, c_void_p, c_int, c_void_p]),
    "fft_ragged_create": (c_void_p, [ctypes.POINTER(RaggedRequest), c_size_t, c_int, c_int, c_int, c_int, c_int]),
    "fft_ragged_autotune": (c_int, [c_void_p, c_int]),
    "fft_ragged_set_engine": (c_int, [c_void_p, c_int, c_int, c_int]),
    "fft_ragged_engine_count": (c_int, []),
    "fft_ragged_execute": (c_int, [c_void_p, c_void_p]),
    "fft_ragged_execute_plain": (c_int, [c_void_p, c_void_p]),
    "fft_ragged_describe": (c_int, [c_void_p, c_char_p, c_size_t]),
    "fft_ragged_free": (None, [c_void_p]),
    "fft_plan_metadata": (c_size_t, [c_void_p, c_int, f32p, c_size_t]),
    "fft_set_variant": (c_int, [c_void_p, c_int]),
    "fft_variant_count": (c_int, [c_void_p]),
    "fft_variant_name": (c_int, [c_void_p, c_int, c_char_p, c_size_t]),
    "fft_variant_default": (c_int, [c_void_p, c_size_t]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the library once and attach the signatures; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `python fftfree_amd/build.py` (no CPU fallback exists)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib

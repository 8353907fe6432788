"""ctypes mirror of the grid pipeline's C ABI (include/ft_tools_cffi.h).

The structs and the call follow the reference's pipeline tool, which fills the same four
structs and calls ft_apply_complex_pipeline on the STFT's real/imag planes
(tools/ft_tools_pipeline.py:359-401). Grids are (time_bins, freq_bins) float32 arrays,
changed in place on success; a call that fails (a result of another shape, NULL planes,
an empty grid) returns False and leaves them as they were. The ``*_device`` forms take
torch tensors on the GPU and enqueue on torch's current stream.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib


class GridSpec(ctypes.Structure):
    """ft_grid_spec (ft_tools_cffi.hpp:9-14); a step of 0 reads as 1."""
    _fields_ = [("time_origin", ctypes.c_double), ("freq_origin", ctypes.c_double),
                ("time_step", ctypes.c_double), ("freq_step", ctypes.c_double)]


class ResamplePlan(ctypes.Structure):
    """ft_resample_plan_c (ft_tools_cffi.hpp:16-22); factors <= 0 read as 1."""
    _fields_ = [("upsample_time", ctypes.c_int), ("upsample_freq", ctypes.c_int),
                ("downsample_time", ctypes.c_int), ("downsample_freq", ctypes.c_int),
                ("preserve_energy", ctypes.c_int)]


class QuantizerConfig(ctypes.Structure):
    """ft_quantizer_config (ft_tools_cffi.hpp:24-30)."""
    _fields_ = [("enabled", ctypes.c_int), ("base_frequency", ctypes.c_float), ("magnitude_step", ctypes.c_float),
                ("mix", ctypes.c_float), ("distance_decay", ctypes.c_float)]


class ContrastConfig(ctypes.Structure):
    """ft_contrast_config (ft_tools_cffi.hpp:32-40); a kernel size <= 0 spans the axis."""
    _fields_ = [("enabled", ctypes.c_int), ("kernel_time", ctypes.c_int), ("kernel_freq", ctypes.c_int),
                ("brightness", ctypes.c_float), ("contrast", ctypes.c_float), ("gamma", ctypes.c_float),
                ("clamp_zero", ctypes.c_int)]


def grid_spec(freq_origin: float = 0.0, freq_step: float = 1.0, time_origin: float = 0.0,
              time_step: float = 1.0) -> GridSpec:
    return GridSpec(time_origin, freq_origin, time_step, freq_step)


def resample_plan(up: Sequence[int] = (1, 1), down: Sequence[int] = (1, 1), preserve: bool = True) -> ResamplePlan:
    return ResamplePlan(int(up[0]), int(up[1]), int(down[0]), int(down[1]), int(bool(preserve)))


def quantizer(base: float = 55.0, step: float = 0.05, mix: float = 0.5, decay: float = 8.0,
              enabled: bool = True) -> QuantizerConfig:
    """Defaults of ExponentialHarmonicQuantizer (ft_tools.hpp:268-273)."""
    return QuantizerConfig(int(enabled), base, step, mix, decay)


def contrast(kernel_time: int = 3, kernel_freq: int = 3, brightness: float = 0.0, gain: float = 0.5,
             gamma: float = 1.0, clamp_zero: bool = True, enabled: bool = True) -> ContrastConfig:
    """Defaults of SpectralLocalContrastBooster (ft_tools.hpp:359-367)."""
    return ContrastConfig(int(enabled), int(kernel_time), int(kernel_freq), brightness, gain, gamma, int(clamp_zero))


def _ref(x):
    return ctypes.byref(x) if x is not None else None


def _plane(a: np.ndarray) -> np.ndarray:
    if a.dtype != np.float32 or not a.flags.c_contiguous or a.ndim != 2:
        raise TypeError("grid planes must be C-contiguous 2-D float32 arrays (time_bins, freq_bins)")
    return a


def _apply(fn, a: np.ndarray, b: np.ndarray, spec, plan, quant, contrast_cfg) -> bool:
    a, b = _plane(a), _plane(b)
    if a.shape != b.shape:
        raise ValueError(f"plane shapes differ: {a.shape} vs {b.shape}")
    ok = fn(_ref(spec), _ref(plan), _ref(quant), _ref(contrast_cfg), a.ctypes.data_as(_lib.f32p),
            b.ctypes.data_as(_lib.f32p), int(a.shape[0]), int(a.shape[1]))
    return bool(ok)


def apply_complex_pipeline(real: np.ndarray, imag: np.ndarray, spec: Optional[GridSpec] = None,
                           plan: Optional[ResamplePlan] = None, quant: Optional[QuantizerConfig] = None,
                           contrast_cfg: Optional[ContrastConfig] = None) -> bool:
    """ft_apply_complex_pipeline on (time, freq) real/imag planes, in place."""
    return _apply(_lib.load().ft_apply_complex_pipeline, real, imag, spec, plan, quant, contrast_cfg)


def apply_polar_pipeline(magnitude: np.ndarray, phase: np.ndarray, spec: Optional[GridSpec] = None,
                         plan: Optional[ResamplePlan] = None, quant: Optional[QuantizerConfig] = None,
                         contrast_cfg: Optional[ContrastConfig] = None) -> bool:
    """ft_apply_polar_pipeline on (time, freq) magnitude/phase planes, in place."""
    return _apply(_lib.load().ft_apply_polar_pipeline, magnitude, phase, spec, plan, quant, contrast_cfg)


def _device_args(a, b, out_a, out_b):
    for t in (a, b, out_a, out_b):
        if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 4 and t.is_contiguous() and t.dim() == 2):
            raise TypeError("device planes must be contiguous 2-D float32 CUDA tensors (time_bins, freq_bins)")
        if t.shape != a.shape:
            raise ValueError(f"plane shapes differ: {tuple(t.shape)} vs {tuple(a.shape)}")
    return a.data_ptr(), b.data_ptr(), out_a.data_ptr(), out_b.data_ptr(), int(a.shape[0]), int(a.shape[1])


def _stream(stream):
    import torch
    return (stream or torch.cuda.current_stream()).cuda_stream


def apply_complex_pipeline_device(real, imag, out_real=None, out_imag=None, spec=None, plan=None, quant=None,
                                  contrast_cfg=None, stream=None) -> bool:
    """ft_apply_complex_pipeline_device: out planes default to the input planes (in place)."""
    out_real = real if out_real is None else out_real
    out_imag = imag if out_imag is None else out_imag
    args = _device_args(real, imag, out_real, out_imag)
    return bool(_lib.load().ft_apply_complex_pipeline_device(_ref(spec), _ref(plan), _ref(quant), _ref(contrast_cfg),
                                                             *args, _stream(stream)))


def apply_polar_pipeline_device(magnitude, phase, out_magnitude=None, out_phase=None, spec=None, plan=None,
                                quant=None, contrast_cfg=None, stream=None) -> bool:
    """ft_apply_polar_pipeline_device: out planes default to the input planes (in place)."""
    out_magnitude = magnitude if out_magnitude is None else out_magnitude
    out_phase = phase if out_phase is None else out_phase
    args = _device_args(magnitude, phase, out_magnitude, out_phase)
    return bool(_lib.load().ft_apply_polar_pipeline_device(_ref(spec), _ref(plan), _ref(quant), _ref(contrast_cfg),
                                                           *args, _stream(stream)))

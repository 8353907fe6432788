"""Differentiable STFT for torch: forward and adjoint are the library's HIP kernels.

``stft_polar(x, ...)`` maps a (signals, samples) float32 CUDA tensor to the magnitude
and phase planes (signals, frames, N/2+1) that a R2C half-spectrum, store_polar context
writes (fft_cffi.cpp:811-825 after the post hook, eigen_fft.hpp:2535-2596). Its backward
is the hand-written adjoint kernel run on the forward's own outputs
(fft_stft_adjoint_saved_device): with the saved |X| and arg X it forms
V = dL/d|X| * e^{i arg X} + i dL/d(arg X) * e^{i arg X}/|X| (0 where |X| = 0), runs the
real inverse DFT of the half spectrum, applies the analysis window and overlap-adds in a
fixed order; no forward transform is recomputed and the samples are not kept.
fft_stft_adjoint_device, which recomputes each frame's spectrum from the samples, is the
same gradient for callers that keep only the samples.

The phase-recovery and weight-training tools of the reference (phase_infer.hpp,
train_mipmap_weights.py) optimise through exactly this map.
"""
from __future__ import annotations

from functools import lru_cache

import torch

from .cffi_fft import CffiFft, FftPlan

_ff = None


def _lib() -> CffiFft:
    global _ff
    if _ff is None:
        _ff = CffiFft()
    return _ff


@lru_cache(maxsize=32)
def polar_plan(n: int, window: int, hop: int, window_kind: int = 1, pad_mode: int = 1) -> FftPlan:
    """A forward R2C half-spectrum, store_polar context (kernel code 2) for the map."""
    return _lib().make_plan(n, window=window, hop=hop, kernel=2, transform=1, store_polar=True, half_spectrum=True,
                            apply_windows=window_kind != 0, analysis_window=window_kind, pad_mode=pad_mode)


class StftPolarFunction(torch.autograd.Function):
    """(signals, samples) -> (|X|, arg X), each (signals, frames, N/2+1)."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, plan: FftPlan):
        if not (x.is_cuda and x.dtype == torch.float32 and x.dim() == 2):
            raise TypeError("stft_polar takes a (signals, samples) float32 CUDA tensor")
        x = x.contiguous()
        ctx.set_materialize_grads(False)  # an unused output's gradient arrives as None
        ff = _lib()
        frames = ff.frame_count(plan, x.shape[1], 0)
        if frames == 0:
            raise ValueError("signal too short for this window/hop/pad rule")
        bins = plan.n // 2 + 1
        mag = torch.empty((x.shape[0], frames, bins), device=x.device, dtype=torch.float32)
        phase = torch.empty_like(mag)
        ff.stft_signals_device(plan, x, mag, phase, None)
        ctx.save_for_backward(mag, phase)
        ctx.plan = plan
        ctx.samples = x.shape[1]
        return mag, phase

    @staticmethod
    def backward(ctx, grad_mag, grad_phase):
        mag, phase = ctx.saved_tensors
        gm = grad_mag.contiguous() if grad_mag is not None else None
        gp = grad_phase.contiguous() if grad_phase is not None else None
        grad = _lib().stft_adjoint_saved_device(ctx.plan, ctx.samples, mag, phase, gm, gp)
        return grad, None


def stft_polar(x: torch.Tensor, n: int, window: int | None = None, hop: int | None = None, window_kind: int = 1,
               pad_mode: int = 1):
    """Differentiable (|X|, arg X) of a batch of signals; Hann window by default."""
    window = window or n
    hop = hop or window // 4
    return StftPolarFunction.apply(x, polar_plan(n, window, hop, window_kind, pad_mode))


def magnitude_loss(x: torch.Tensor, target_mag: torch.Tensor, n: int, window: int | None = None,
                   hop: int | None = None) -> torch.Tensor:
    """sum over signals, frames and bins of (|STFT(x)| - target)^2."""
    mag, _ = stft_polar(x, n, window, hop)
    return ((mag - target_mag) ** 2).sum()

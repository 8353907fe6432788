"""Python front end of libfft_cffi.so, shaped like the reference's CFFI callers.

``CffiFft.stft`` / ``CffiFft.istft`` mirror train_mipmap_weights.py:297-370 (host numpy
in, host numpy out, (bins, frames) layout). The ``*_device`` methods take torch tensors
already on the GPU and launch on torch's current stream, so no host copies happen;
they are what the autograd function and the benchmarks use.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib

FFT_TRANSFORM_C2C, FFT_TRANSFORM_R2C, FFT_TRANSFORM_C2R, FFT_TRANSFORM_R2R = range(4)
FFT_WINDOW_RECT, FFT_WINDOW_HANN, FFT_WINDOW_HAMMING, FFT_WINDOW_BLACKMAN, FFT_WINDOW_TUKEY, FFT_WINDOW_KAISER = range(6)
FFT_WINDOW_NORM_NONE, FFT_WINDOW_NORM_L2, FFT_WINDOW_NORM_AREA = range(3)
KERNEL_AUTO, KERNEL_COOLEYTUKEY_C, KERNEL_STOCKHAM_C = 0, 1, 2  # C-level kernel codes (fft_cffi.cpp:799-800)
FFT_META_TWIDDLE, FFT_META_LAYOUT_MAP, FFT_META_STAGE_PARAMS, FFT_META_BUTTERFLY_PAIRS = 0, 1, 3, 4
FFT_META_TWIDDLE_INDEX_MAP, FFT_META_OMEGA = 6, 7


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_lib.f32p)


@dataclass
class FftPlan:
    """One library context plus the settings callers need to size buffers."""
    handle: int
    n: int
    hop: int
    window: int
    half: bool
    inverse: int
    transform: int = FFT_TRANSFORM_R2C
    apply_ola: bool = False
    store_polar: bool = False
    reduce_magnitude: bool = False

    @property
    def bins(self) -> int:
        return self.n // 2 + 1 if self.half else self.n


@dataclass
class CffiFft:
    """Owns plans; mirrors the reference's CffiFft (train_mipmap_weights.py:176-370)."""
    lib: ctypes.CDLL = field(default_factory=_lib.load)
    _plans: list = field(default_factory=list)

    def make_plan(self, n: int, *, window: int = 0, hop: int = 0, inverse: int = 0, kernel: int = 0, radix: int = 0,
                  radix_pattern: Optional[Sequence[int]] = None, pad_mode: int = 1,
                  transform: int = FFT_TRANSFORM_R2C, reduce_magnitude: bool = False, store_polar: bool = False,
                  half_spectrum: bool = True, threads: int = 1, lanes: int = 1, apply_windows: bool = False,
                  apply_ola: bool = False, analysis_window: int = FFT_WINDOW_RECT, analysis_p1: float = 0.0,
                  analysis_p2: float = 0.0, synthesis_window: int = FFT_WINDOW_RECT, synthesis_p1: float = 0.0,
                  synthesis_p2: float = 0.0, norm: int = FFT_WINDOW_NORM_NONE, cola: int = 0) -> FftPlan:
        pat = None
        pat_len = 0
        if radix_pattern:
            arr = (ctypes.c_int * len(radix_pattern))(*radix_pattern)
            pat = ctypes.cast(arr, _lib.i32p)
            pat_len = len(radix_pattern)
        h = self.lib.fft_init_full_v2(n, threads, lanes, inverse, kernel, radix, pat, pat_len, pad_mode, window, hop,
                                      transform, int(reduce_magnitude), int(store_polar), int(half_spectrum), 0, 0, 0,
                                      0, 1, int(apply_windows), int(apply_ola), analysis_window, analysis_p1,
                                      analysis_p2, synthesis_window, synthesis_p1, synthesis_p2, norm, cola)
        if not h:
            raise ValueError(f"fft_init_full_v2 refused n={n}")
        W = window or n
        p = FftPlan(handle=h, n=n, hop=hop or W, window=W, half=bool(half_spectrum), inverse=inverse,
                    transform=transform, apply_ola=bool(apply_ola), store_polar=bool(store_polar),
                    reduce_magnitude=bool(reduce_magnitude))
        self._plans.append(p)
        return p

    def close(self) -> None:
        for p in self._plans:
            self.lib.fft_free(p.handle)
        self._plans.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- host API (reference shape) -------------------------------------------------
    def frame_count(self, plan: FftPlan, pcm_len: int, pad_mode: int = 1, max_frames: int = 0) -> int:
        return int(self.lib.fft_frame_count(plan.handle, pcm_len, pad_mode, max_frames))

    def stft(self, plan: FftPlan, pcm: np.ndarray, pad_mode: int = 1) -> Tuple[np.ndarray, np.ndarray, np.ndarray, int, int]:
        """Forward STFT; returns real, imag, mag as (bins, frames), bins, frames."""
        pcm = np.ascontiguousarray(pcm, dtype=np.float32)
        bins = plan.bins
        max_frames = int(pcm.shape[0] // max(1, plan.hop) + 3)
        out = np.zeros((3, max_frames * bins), dtype=np.float32)
        produced = int(self.lib.fft_execute_batched(plan.handle, _fp(pcm), pcm.size, _fp(out[0]), _fp(out[1]),
                                                    _fp(out[2]), pad_mode, 0, 0))
        if produced <= 0:
            raise RuntimeError("fft_execute_batched failed")
        used = produced * bins
        re, im, mag = (out[i, :used].reshape(produced, bins).T for i in range(3))
        return re, im, mag, bins, produced

    def stft_frames(self, plan: FftPlan, pcm: np.ndarray, pad_mode: int = 0, max_frames: int = 0):
        """fft_execute_batched returning (frames, re, im, mag) with (frames, bins) planes."""
        pcm = np.ascontiguousarray(pcm, dtype=np.float32)
        cap = self.frame_count(plan, pcm.size, pad_mode, max_frames)
        bins = plan.bins
        out = np.zeros((3, max(cap, 1) * bins), dtype=np.float32)
        got = int(self.lib.fft_execute_batched(plan.handle, _fp(pcm), pcm.size, _fp(out[0]), _fp(out[1]),
                                               _fp(out[2]), pad_mode, 0, max_frames))
        planes = [out[i, :got * bins].reshape(got, bins) for i in range(3)]
        return got, planes[0], planes[1], planes[2]

    def istft(self, plan: FftPlan, real: np.ndarray, imag: np.ndarray) -> np.ndarray:
        """Inverse through fft_execute_complex_batched, OLA in Python unless the plan
        does it internally (train_mipmap_weights.py:326-370)."""
        assert real.shape == imag.shape and real.ndim == 2
        bins, frames = map(int, real.shape)
        plan_bins = plan.bins
        in_real = np.zeros((frames, plan_bins), dtype=np.float32)
        in_imag = np.zeros((frames, plan_bins), dtype=np.float32)
        b = min(bins, plan_bins)
        in_real[:, :b] = real[:b].T
        in_imag[:, :b] = imag[:b].T
        out = np.zeros(frames * plan.n, dtype=np.float32)
        produced = int(self.lib.fft_execute_complex_batched(plan.handle, _fp(in_real), _fp(in_imag), frames, _fp(out),
                                                            1, 0, 0))
        if produced <= 0:
            raise RuntimeError("fft_execute_complex_batched failed")
        out_len = (frames - 1) * plan.hop + plan.n
        if plan.apply_ola:
            return out[:out_len]
        wav = np.zeros(out_len, dtype=np.float32)
        counts = np.zeros(out_len, dtype=np.float32)
        for f in range(frames):
            s = f * plan.hop
            wav[s:s + plan.n] += out[f * plan.n:(f + 1) * plan.n]
            counts[s:s + plan.n] += 1.0
        m = counts > 0
        wav[m] /= counts[m]
        return wav

    def c2c(self, plan: FftPlan, x: np.ndarray) -> np.ndarray:
        """Batched complex transform of x (batch, n) complex64 (fft_inplace_batched)."""
        x = np.ascontiguousarray(x, dtype=np.complex64)
        out = np.empty_like(x)
        got = int(self.lib.fft_execute_c2c_batched(plan.handle, _fp(x.view(np.float32)), _fp(out.view(np.float32)),
                                                   x.shape[0]))
        if got != x.shape[0]:
            raise RuntimeError("fft_execute_c2c_batched failed")
        return out

    def metadata(self, plan: FftPlan, kind: int) -> np.ndarray:
        cap = max(16, plan.n * max(1, plan.n.bit_length()))
        out = np.zeros(cap, dtype=np.complex64)
        got = int(self.lib.fft_plan_metadata(plan.handle, kind, _fp(out.view(np.float32)), cap))
        return out[:got]

    def set_variant(self, plan: FftPlan, variant: int) -> bool:
        return bool(self.lib.fft_set_variant(plan.handle, variant))

    def default_variant(self, plan: FftPlan, frames: int) -> int:
        """Index of the variant a launch of `frames` frames runs (fft_variant_default)."""
        return int(self.lib.fft_variant_default(plan.handle, frames))

    def variants(self, plan: FftPlan):
        names = []
        buf = ctypes.create_string_buffer(128)
        for i in range(int(self.lib.fft_variant_count(plan.handle))):
            self.lib.fft_variant_name(plan.handle, i, buf, 128)
            names.append(buf.value.decode())
        return names

    # -- device API (torch tensors on the GPU, torch's current stream) ---------------
    @staticmethod
    def _stream():
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def stft_device(self, plan: FftPlan, pcm, out_re=None, out_im=None, out_mag=None, pad_mode: int = 0,
                    max_frames: int = 0) -> int:
        """fft_execute_batched_device on torch tensors; planes are (frames, bins) float32
        CUDA tensors or None (not written). Returns frames."""
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        got = int(self.lib.fft_execute_batched_device(plan.handle, ctypes.c_void_p(pcm.data_ptr()), pcm.numel(),
                                                      ptr(out_re), ptr(out_im), ptr(out_mag), pad_mode, max_frames,
                                                      self._stream()))
        if got == 0:
            raise RuntimeError("fft_execute_batched_device failed")
        return got

    def c2c_device(self, plan: FftPlan, x, out=None):
        """Batched complex transform of a (batch, n) complex64 CUDA tensor."""
        import torch
        assert x.dtype == torch.complex64 and x.is_contiguous() and x.shape[-1] == plan.n
        out = x if out is None else out
        got = int(self.lib.fft_execute_c2c_batched_device(plan.handle, ctypes.c_void_p(x.data_ptr()),
                                                          ctypes.c_void_p(out.data_ptr()), x.numel() // plan.n,
                                                          self._stream()))
        if got == 0:
            raise RuntimeError("fft_execute_c2c_batched_device failed")
        return out

    def complex_batched_device(self, plan: FftPlan, in_re, in_im, out=None, max_frames: int = 0):
        """fft_execute_complex_batched_device: in_re / in_im are (frames, bins) float32 CUDA
        tensors; returns the (frames * n,) float32 output (allocated when out is None)."""
        import torch
        assert in_re.shape == in_im.shape and in_re.dim() == 2 and in_re.shape[1] == plan.bins
        assert in_re.is_contiguous() and in_im.is_contiguous()
        frames = int(in_re.shape[0])
        if max_frames:
            frames = min(frames, max_frames)
        if out is None:
            out = torch.empty(frames * plan.n, dtype=torch.float32, device=in_re.device)
        assert out.numel() >= frames * plan.n and out.is_contiguous()
        got = int(self.lib.fft_execute_complex_batched_device(plan.handle, ctypes.c_void_p(in_re.data_ptr()),
                                                              ctypes.c_void_p(in_im.data_ptr()), frames,
                                                              ctypes.c_void_p(out.data_ptr()), max_frames,
                                                              self._stream()))
        if got == 0:
            raise RuntimeError("fft_execute_complex_batched_device failed")
        return out

    def griffin_lim(self, fwd: FftPlan, inv: FftPlan, mag: np.ndarray, hop: int = 0, iterations: int = 32,
                    pad_mode: int = 1, seed: int = 0):
        """fft_griffin_lim (the reference's phase reconstruction) on a (frames, bins) float32
        magnitude array; returns (real, imag) of the same shape."""
        mag = np.ascontiguousarray(mag, dtype=np.float32)
        re = np.zeros_like(mag)
        im = np.zeros_like(mag)
        ok = int(self.lib.fft_griffin_lim(fwd.handle, inv.handle, _fp(mag), mag.shape[0], hop, int(inv.half),
                                          iterations, pad_mode, seed & 0xFFFFFFFF, _fp(re), _fp(im)))
        if ok != 1:
            raise RuntimeError("fft_griffin_lim failed")
        return re, im

    def griffin_lim_device(self, fwd: FftPlan, inv: FftPlan, mag, out_re=None, out_im=None, hop: int = 0,
                           iterations: int = 32, pad_mode: int = 1, seed: int = 0):
        """fft_griffin_lim_device on a (frames, bins) float32 CUDA tensor; returns (real, imag)."""
        import torch
        assert mag.dim() == 2 and mag.is_contiguous()
        out_re = torch.empty_like(mag) if out_re is None else out_re
        out_im = torch.empty_like(mag) if out_im is None else out_im
        vp = lambda t: ctypes.c_void_p(t.data_ptr())
        ok = int(self.lib.fft_griffin_lim_device(fwd.handle, inv.handle, vp(mag), mag.shape[0], hop, int(inv.half),
                                                 iterations, pad_mode, seed & 0xFFFFFFFF, vp(out_re), vp(out_im),
                                                 self._stream()))
        if ok != 1:
            raise RuntimeError("fft_griffin_lim_device failed")
        return out_re, out_im

    def stft_signals_device(self, plan: FftPlan, pcm, out_re=None, out_im=None, out_mag=None,
                            pad_mode: int = 0) -> int:
        """fft_stft_signals_device: pcm is a (signals, length) float32 CUDA tensor; planes
        are (signals, frames, bins) tensors or None. Returns frames per signal."""
        assert pcm.dim() == 2 and pcm.is_contiguous()
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        got = int(self.lib.fft_stft_signals_device(plan.handle, ctypes.c_void_p(pcm.data_ptr()), pcm.shape[0],
                                                   pcm.shape[1], ptr(out_re), ptr(out_im), ptr(out_mag), pad_mode,
                                                   self._stream()))
        if got == 0:
            raise RuntimeError("fft_stft_signals_device failed")
        return got

    def stft_adjoint_device(self, plan: FftPlan, pcm, grad_re=None, grad_im=None, out=None, pad_mode: int = 0):
        """fft_stft_adjoint_device: gradient of the samples from the gradients of the
        first two planes ((signals, frames, bins) tensors or None). Returns out."""
        import torch
        assert pcm.dim() == 2 and pcm.is_contiguous()
        out = torch.empty_like(pcm) if out is None else out
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        got = int(self.lib.fft_stft_adjoint_device(plan.handle, ctypes.c_void_p(pcm.data_ptr()), pcm.shape[0],
                                                   pcm.shape[1], ptr(grad_re), ptr(grad_im),
                                                   ctypes.c_void_p(out.data_ptr()), pad_mode, self._stream()))
        if got == 0:
            raise RuntimeError("fft_stft_adjoint_device failed")
        return out

    def stft_adjoint_saved_device(self, plan: FftPlan, samples: int, plane_re=None, plane_im=None, grad_re=None,
                                  grad_im=None, out=None, pad_mode: int = 0):
        """fft_stft_adjoint_saved_device: gradient of the samples from the forward's own
        planes (|X|, arg X for a polar context; unused for complex planes) and the
        gradients of those planes ((signals, frames, bins) tensors or None); `samples` per
        signal. Returns out, a (signals, samples) tensor."""
        import torch
        ref = next(t for t in (grad_re, grad_im, plane_re, plane_im, out) if t is not None)
        signals = ref.shape[0]
        out = torch.empty((signals, samples), device=ref.device, dtype=torch.float32) if out is None else out
        assert out.shape == (signals, samples) and out.is_contiguous()
        for t in (plane_re, plane_im, grad_re, grad_im):
            assert t is None or (t.is_contiguous() and t.shape[0] == signals)
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        got = int(self.lib.fft_stft_adjoint_saved_device(plan.handle, ptr(plane_re), ptr(plane_im), signals, samples,
                                                         ptr(grad_re), ptr(grad_im), ctypes.c_void_p(out.data_ptr()),
                                                         pad_mode, self._stream()))
        if got == 0:
            raise RuntimeError("fft_stft_adjoint_saved_device failed")
        return out


class RaggedBatch:
    """A ragged batch over torch CUDA tensors (fft_ragged_*): requests are
    (pcm, window, hop) with pcm a 1-D float32 CUDA tensor; outputs are allocated here as
    three (frames, window/2+1) planes per request (None planes are skipped)."""

    def __init__(self, ff: "CffiFft", requests, pad_mode: int = 1, reduce_magnitude: bool = False,
                 store_polar: bool = True, apply_windows: bool = True, window_kind: int = 1,
                 planes=(True, True, True)):
        import torch
        self.ff = ff
        lib = ff.lib
        n = len(requests)
        arr = (_lib.RaggedRequest * n)()
        self.pcm = []
        for i, (pcm, window, hop) in enumerate(requests):
            assert pcm.is_cuda and pcm.dtype == torch.float32 and pcm.is_contiguous()
            self.pcm.append(pcm)
            arr[i].pcm, arr[i].pcm_len, arr[i].window, arr[i].hop = pcm.data_ptr(), pcm.numel(), window, hop
        # frame counts first (one context per geometry answers them on the host)
        self.outputs = []
        geometry = {}
        for i, (pcm, window, hop) in enumerate(requests):
            if (window, hop) not in geometry:
                geometry[(window, hop)] = ff.make_plan(window, window=window, hop=hop, kernel=2, transform=1,
                                                       half_spectrum=True, pad_mode=pad_mode)
            frames = ff.frame_count(geometry[(window, hop)], pcm.numel(), 0)
            outs = [torch.empty((frames, window // 2 + 1), device=pcm.device) if want else None for want in planes]
            self.outputs.append(outs)
            ptr = lambda t: t.data_ptr() if t is not None and t.numel() else None
            arr[i].out_real, arr[i].out_imag, arr[i].out_mag = ptr(outs[0]), ptr(outs[1]), ptr(outs[2])
        self._arr = arr
        self.handle = lib.fft_ragged_create(arr, n, pad_mode, int(reduce_magnitude), int(store_polar),
                                            int(apply_windows), window_kind)
        if not self.handle:
            raise ValueError("fft_ragged_create refused the batch")
        self.frames = [int(arr[i].frames) for i in range(n)]
        for outs, f in zip(self.outputs, self.frames):
            assert all(o is None or o.shape[0] == f for o in outs)

    def autotune(self, repeats: int = 3) -> None:
        if not self.ff.lib.fft_ragged_autotune(self.handle, repeats):
            raise RuntimeError("fft_ragged_autotune failed")

    def set_engine(self, n: int, engine: int, blocks_per_item: int = 2) -> bool:
        """fft_ragged_set_engine: False when the engine does not run window size n."""
        return bool(self.ff.lib.fft_ragged_set_engine(self.handle, n, engine, blocks_per_item))

    def execute(self) -> None:
        if not self.ff.lib.fft_ragged_execute(self.handle, CffiFft._stream()):
            raise RuntimeError("fft_ragged_execute failed")

    def execute_plain(self) -> None:
        if not self.ff.lib.fft_ragged_execute_plain(self.handle, CffiFft._stream()):
            raise RuntimeError("fft_ragged_execute_plain failed")

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(512)
        self.ff.lib.fft_ragged_describe(self.handle, buf, 512)
        return buf.value.decode()

    def close(self) -> None:
        if self.handle:
            self.ff.lib.fft_ragged_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

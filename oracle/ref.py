"""ctypes binding of oracle/cpu_ref.cpp (libcpu_ref.so). TEST INFRASTRUCTURE ONLY.

cpu_ref.cpp restates the reference's float CPU path (Cooley-Tukey and Stockham radix-2,
framing, window, output planes); see the header of that file for the reference regions.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libcpu_ref.so"
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int)
_sz = ctypes.c_size_t


def _p(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_f32p)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing: run oracle/build.py (or __graft_entry__.build())")
        L = ctypes.CDLL(str(LIB_PATH))
        L.ref_twiddles.argtypes = [ctypes.c_int, ctypes.c_int, _f32p]
        L.ref_bitrev.argtypes = [ctypes.c_int, _i32p]
        L.ref_build_window.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, _f32p]
        L.ref_build_window.restype = None
        L.ref_frame_count.argtypes = [_sz, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz]
        L.ref_frame_count.restype = _sz
        L.ref_fft_batched.argtypes = [ctypes.c_int] * 7 + [_f32p, _sz, ctypes.c_int]
        L.ref_execute_batched.argtypes = ([ctypes.c_int] * 10 + [_f32p, _f32p, _sz, _f32p, _f32p, _f32p,
                                                                 ctypes.c_int, _sz, ctypes.c_int])
        L.ref_execute_batched.restype = _sz
        L.ref_execute_complex_batched.argtypes = ([ctypes.c_int] * 7 + [_f32p, _f32p, _f32p, _sz, _f32p, _sz,
                                                                        ctypes.c_int])
        L.ref_execute_complex_batched.restype = _sz
        L.ref_complex_output.argtypes = [ctypes.c_int, _sz, _f32p, ctypes.c_int, _f32p, _f32p, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, _f32p]
        L.ref_complex_output.restype = None
        L.ref_gl_angles.argtypes = [ctypes.c_uint, _sz, _f32p]
        L.ref_gl_angles.restype = None
        L.ref_griffin_lim.argtypes = ([ctypes.c_int, ctypes.c_int, _f32p, _sz, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p, ctypes.c_int])
        L.ref_griffin_lim.restype = ctypes.c_int
        L.ref_phase_linear.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _sz, _f32p, _f32p]
        L.ref_phase_linear.restype = None
        L.ref_phase_minimum.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, _sz, _f32p, _f32p]
        L.ref_phase_minimum.restype = ctypes.c_int
        L.ref_ft_pipeline.argtypes = ([ctypes.c_int, ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 6 +
                                      [ctypes.c_float] * 4 + [ctypes.c_int] * 3 + [ctypes.c_float] * 3 +
                                      [ctypes.c_int, _f32p, _f32p, ctypes.c_int, ctypes.c_int])
        L.ref_ft_pipeline.restype = ctypes.c_int
        L.ref_metadata.argtypes = [ctypes.c_int] * 6 + [_f32p, _sz]
        L.ref_metadata.restype = _sz
        _lib = L
    return _lib


def default_threads() -> int:
    """min(16, cores): the reference clamps plan threads to 16 (eigen_fft.hpp:1966)."""
    return max(1, min(16, os.cpu_count() or 1))


def twiddles(N: int, inverse: bool = False) -> np.ndarray:
    out = np.zeros(max(1, N // 2), dtype=np.complex64)
    if not lib().ref_twiddles(N, int(inverse), out.view(np.float32).ctypes.data_as(_f32p)):
        raise ValueError("N must be a power of two")
    return out[: N // 2]


def bitrev(N: int) -> np.ndarray:
    out = np.zeros(N, dtype=np.int32)
    if not lib().ref_bitrev(N, out.ctypes.data_as(_i32p)):
        raise ValueError("N must be a power of two")
    return out


def build_window(kind: int, W: int, p1: float = 0.0, p2: float = 0.0, norm: int = 0) -> np.ndarray:
    out = np.zeros(W, dtype=np.float32)
    lib().ref_build_window(kind, W, p1, p2, norm, _p(out))
    return out


def frame_count(pcm_len, N, window, hop, ctx_pad, pad_mode, max_frames=0) -> int:
    return int(lib().ref_frame_count(pcm_len, N, window, hop, ctx_pad, pad_mode, max_frames))


def fft_batched(x: np.ndarray, inverse=False, kernel=0, transform=0, reduce_magnitude=False,
                store_polar=False, half_spectrum=False, threads=None) -> np.ndarray:
    """x: (B, N) complex64; returns the transformed copy (reference fft_inplace_batched)."""
    x = np.ascontiguousarray(x, dtype=np.complex64).copy()
    B, N = x.shape
    ok = lib().ref_fft_batched(N, int(inverse), kernel, transform, int(reduce_magnitude), int(store_polar),
                               int(half_spectrum), x.view(np.float32).ctypes.data_as(_f32p), B,
                               threads or default_threads())
    if not ok:
        raise ValueError("N must be a power of two")
    return x


def execute_batched(pcm: np.ndarray, N, window=0, hop=0, ctx_pad=0, pad_mode=0, max_frames=0, inverse=False,
                    kernel=0, transform=0, reduce_magnitude=False, store_polar=False, half_spectrum=False,
                    window_taps: np.ndarray | None = None, threads=None):
    """Reference fft_execute_batched; returns (frames, real, imag, mag) as (frames, bins)."""
    pcm = np.ascontiguousarray(pcm, dtype=np.float32)
    bins = N // 2 + 1 if half_spectrum else N
    cap = frame_count(pcm.size, N, window, hop, ctx_pad, pad_mode, max_frames)
    re = np.zeros((max(cap, 1), bins), np.float32)
    im = np.zeros_like(re)
    mg = np.zeros_like(re)
    taps = None if window_taps is None else np.ascontiguousarray(window_taps, dtype=np.float32)
    got = lib().ref_execute_batched(N, int(inverse), kernel, window, hop, ctx_pad, transform, int(reduce_magnitude),
                                    int(store_polar), int(half_spectrum), _p(taps), _p(pcm), pcm.size,
                                    _p(re), _p(im), _p(mg), pad_mode, max_frames, threads or default_threads())
    return int(got), re[:got], im[:got], mg[:got]


def execute_complex_batched(in_real, in_imag, N, inverse=True, kernel=0, transform=2, reduce_magnitude=False,
                            store_polar=False, half_spectrum=True, synth_taps=None, max_frames=0, threads=None,
                            windows_enabled=None, analysis_taps=None, apply_ola=False, hop=0, cola_mode=0):
    """Reference fft_execute_complex_batched (fft_cffi.cpp:1042-1157): returns (frames,
    out) with out of frames*N floats, shaped (frames, N) without overlap-add.
    `windows_enabled` defaults to "synth_taps given"; taps are N floats."""
    in_real = np.ascontiguousarray(in_real, dtype=np.float32)
    in_imag = np.ascontiguousarray(in_imag, dtype=np.float32)
    frames = in_real.shape[0]
    if max_frames:
        frames = min(frames, max_frames)
    cols = np.zeros((frames, N), np.float32)
    got = lib().ref_execute_complex_batched(N, int(inverse), kernel, transform, int(reduce_magnitude),
                                            int(store_polar), int(half_spectrum), None, _p(in_real),
                                            _p(in_imag), frames, _p(cols), max_frames, threads or default_threads())
    if windows_enabled is None:
        windows_enabled = synth_taps is not None
    syn = None if synth_taps is None else np.ascontiguousarray(synth_taps, dtype=np.float32)
    ana = None if analysis_taps is None else np.ascontiguousarray(analysis_taps, dtype=np.float32)
    out = np.zeros((got, N), np.float32)
    lib().ref_complex_output(N, got, _p(cols[:got]), int(bool(windows_enabled)), _p(ana), _p(syn), int(bool(apply_ola)),
                             int(hop), int(cola_mode), _p(out))
    return int(got), (out.reshape(-1) if apply_ola else out)


def gl_angles(seed: int, count: int) -> np.ndarray:
    """The reference's Griffin-Lim start angles (std::mt19937 + uniform_real_distribution<float>)."""
    out = np.zeros(count, np.float32)
    lib().ref_gl_angles(seed, count, _p(out))
    return out


def griffin_lim(mag, N, hop, iterations=32, pad_mode=1, seed=0, half=True, inv_kernel=0, inv_transform=2,
                inv_synth=None, fwd_kernel=0, fwd_window=0, fwd_hop=0, fwd_ctx_pad=1, fwd_taps=None, threads=None):
    """Reference Griffin-Lim (phase_infer.hpp:69-204): returns (ok, out_real, out_imag), (frames, bins)."""
    mag = np.ascontiguousarray(mag, dtype=np.float32)
    frames = mag.shape[0]
    re = np.zeros_like(mag)
    im = np.zeros_like(mag)
    syn = None if inv_synth is None else np.ascontiguousarray(inv_synth, dtype=np.float32)
    taps = None if fwd_taps is None else np.ascontiguousarray(fwd_taps, dtype=np.float32)
    ok = lib().ref_griffin_lim(N, int(half), _p(mag), frames, hop, iterations, pad_mode, seed, inv_kernel,
                               inv_transform, _p(syn), fwd_kernel, fwd_window or N, fwd_hop or hop, fwd_ctx_pad,
                               _p(taps), _p(re), _p(im), threads or default_threads())
    return int(ok), re, im


def phase_infer(mag, N, hop=0, half=True, mode=0):
    """Reference phase_infer_execute (modes 0 and 1): returns (real, imag) shaped like mag."""
    mag = np.ascontiguousarray(mag, dtype=np.float32)
    re, im = np.zeros_like(mag), np.zeros_like(mag)
    if mode == 0:
        lib().ref_phase_linear(N, hop, int(half), _p(mag), mag.shape[0], _p(re), _p(im))
    elif not lib().ref_phase_minimum(N, int(half), _p(mag), mag.shape[0], _p(re), _p(im)):
        raise ValueError("N must be a power of two")
    return re, im


REF_FT_PATH = Path(__file__).resolve().parent / "_ref" / "libft_ref.so"
_ref_ft = None


def ref_ft_available() -> bool:
    """True when oracle/_ref/libft_ref.so (the reference's own ft_tools.hpp) is built."""
    return REF_FT_PATH.exists()


def _ft_args(a, b, polar, freq_origin, freq_step, up, down, preserve, quant, contrast):
    q = quant or (0.0, 0.0, 0.0, 0.0)
    c = contrast or (0, 0, 0.0, 0.0, 1.0, 0)
    return (int(polar), freq_origin, freq_step, up[0], up[1], down[0], down[1], int(preserve), int(quant is not None),
            *q, int(contrast is not None), c[0], c[1], c[2], c[3], c[4], int(c[5]), _p(a), _p(b), a.shape[0],
            a.shape[1])


_FT_ARGTYPES = ([ctypes.c_int, ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 6 + [ctypes.c_float] * 4 +
                [ctypes.c_int] * 3 + [ctypes.c_float] * 3 + [ctypes.c_int, _f32p, _f32p, ctypes.c_int, ctypes.c_int])


def refsrc_ft_pipeline(a, b, polar=False, freq_origin=0.0, freq_step=1.0, up=(1, 1), down=(1, 1), preserve=True,
                       quant=None, contrast=None):
    """The reference's own ft_tools.hpp templates (oracle/_ref/libft_ref.so), same
    arguments and result as ft_pipeline."""
    global _ref_ft
    if _ref_ft is None:
        L = ctypes.CDLL(str(REF_FT_PATH))
        L.refsrc_ft_pipeline.argtypes = _FT_ARGTYPES
        L.refsrc_ft_pipeline.restype = ctypes.c_int
        _ref_ft = L
    a = np.ascontiguousarray(a, dtype=np.float32).copy()
    b = np.ascontiguousarray(b, dtype=np.float32).copy()
    ok = _ref_ft.refsrc_ft_pipeline(*_ft_args(a, b, polar, freq_origin, freq_step, up, down, preserve, quant,
                                              contrast))
    return int(ok), a, b


def ft_pipeline(a, b, polar=False, freq_origin=0.0, freq_step=1.0, up=(1, 1), down=(1, 1), preserve=True,
                quant=None, contrast=None):
    """Reference ft_apply_*_pipeline on copies of the two (time, freq) planes; quant =
    (base, step, mix, decay), contrast = (kernel_time, kernel_freq, brightness, contrast,
    gamma, clamp_zero). Returns (ok, a, b)."""
    a = np.ascontiguousarray(a, dtype=np.float32).copy()
    b = np.ascontiguousarray(b, dtype=np.float32).copy()
    ok = lib().ref_ft_pipeline(*_ft_args(a, b, polar, freq_origin, freq_step, up, down, preserve, quant, contrast))
    return int(ok), a, b


def metadata(N: int, kind: int, kernel: int = 0, inverse=False, transform=0, half_spectrum=False) -> np.ndarray:
    cap = max(4 * N * max(1, N.bit_length()), 16)
    out = np.zeros(cap, dtype=np.complex64)
    n = lib().ref_metadata(N, int(inverse), kernel, transform, int(half_spectrum), kind,
                           out.view(np.float32).ctypes.data_as(_f32p), cap)
    return out[:n]

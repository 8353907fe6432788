"""FP64 oracle for the fftfree batched FFT/STFT path. TEST INFRASTRUCTURE ONLY.

The reference computes in float32 (std::complex<float> plans). This module restates what
its float path is meant to compute, in float64, from the *same float32 inputs* the
reference feeds its butterflies:

* framed samples are the float32 products ``pcm[s+i] * win[i]`` (fft_cffi.cpp:772-775);
* windows are the float32 casts of the double-precision formulas (fft_cffi.cpp:163-208),
  built here with ``math.cos`` (libm double, as ``std::cos``) so they are bit-exact;
* the DFT itself is numpy's float64 FFT (checked against a direct O(N^2) sum in tests);
* inverse plans scale by 1/N (eigen_fft.hpp:3076-3085, 3540-3556);
* the post-transform precedence polar > magnitude > C2R-imag-clear and the R2C
  half-spectrum DC/Nyquist rule follow eigen_fft.hpp:2535-2596, and the output planes
  follow fft_cffi.cpp:811-825.
"""
from __future__ import annotations

import math

import numpy as np

WINDOW_RECT, WINDOW_HANN, WINDOW_HAMMING, WINDOW_BLACKMAN, WINDOW_TUKEY, WINDOW_KAISER = range(6)
NORM_NONE, NORM_L2, NORM_AREA = range(3)
C2C, R2C, C2R, R2R = range(4)


def dft_direct(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Direct O(N^2) DFT over the last axis in complex128 (no library FFT)."""
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[-1]
    k = np.arange(n)
    sgn = 1.0 if inverse else -1.0
    # exact integer phase index (k*m mod n) keeps the twiddle argument small
    idx = np.outer(k, k) % n
    mat = np.exp(sgn * 2j * np.pi * idx / n)
    y = x @ mat.T
    return y / n if inverse else y


def fft64(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """FP64 DFT over the last axis: X[k] = sum_n x[n] e^{-+2 pi i kn/N}, inverse scaled 1/N."""
    x = np.asarray(x, dtype=np.complex128)
    return np.fft.ifft(x, axis=-1) if inverse else np.fft.fft(x, axis=-1)


def build_window(kind: int, W: int, p1: float = 0.0, p2: float = 0.0, norm: int = NORM_NONE) -> np.ndarray:
    """float32 window taps, bit-exact restatement of build_window + normalize_window
    (fft_cffi.cpp:163-208). KAISER is the reference's Hann placeholder (200-203)."""
    f32 = np.float32
    w = np.ones(max(W, 0), dtype=np.float32)
    pi = 3.141592653589793238462643383279502884
    if kind in (WINDOW_HANN, WINDOW_KAISER):
        for n in range(W):
            w[n] = f32(0.5 * (1.0 - math.cos(2.0 * pi * n / (W - 1))))
    elif kind == WINDOW_HAMMING:
        for n in range(W):
            w[n] = f32(0.54 - 0.46 * math.cos(2.0 * pi * n / (W - 1)))
    elif kind == WINDOW_BLACKMAN:
        for n in range(W):
            w[n] = f32(0.42 - 0.5 * math.cos(2.0 * pi * n / (W - 1)) + 0.08 * math.cos(4.0 * pi * n / (W - 1)))
    elif kind == WINDOW_TUKEY:
        alpha = 0.5 if float(np.float32(p1)) == 0.0 else float(np.float32(p1))
        L = max(1, W - 1)
        for n in range(W):
            x = n / L
            if x < alpha / 2:
                w[n] = f32(0.5 * (1 + math.cos(pi * (2 * x / alpha - 1))))
            elif x <= 1 - alpha / 2:
                w[n] = f32(1.0)
            else:
                w[n] = f32(0.5 * (1 + math.cos(pi * (2 * x / alpha - 2 / alpha + 1))))
    if W > 0 and norm == NORM_L2:
        s2 = 0.0
        for v in w:
            s2 += float(v) * float(v)
        s2 = math.sqrt(max(1e-30, s2))
        w = np.array([f32(float(v) / s2) for v in w], dtype=np.float32)
    elif W > 0 and norm == NORM_AREA:
        s = 0.0
        for v in w:
            s += float(v)
        if abs(s) > 1e-30:
            w = np.array([f32(float(v) / s) for v in w], dtype=np.float32)
    return w


def frame_count(pcm_len: int, N: int, window: int, hop: int, ctx_pad: int, pad_mode: int, max_frames: int = 0) -> int:
    """Frames fft_execute_batched produces (fft_cffi.cpp:715-750); 0 = refused."""
    W = window if window else N
    H = hop if hop else W
    if W <= 0 or H <= 0 or W > N or pcm_len == 0:
        return 0
    pad = ctx_pad if pad_mode == 0 else pad_mode
    if pcm_len < W:
        if pad != 1:
            return 0
        frames = 1
    else:
        frames = 1 + (pcm_len - W) // H
        remaining = pcm_len - (frames - 1) * H
        if remaining < W:
            if pad == 1:
                pass
            elif pad == 2:
                frames -= 1
            else:
                return 0
    if frames == 0:
        return 0
    if max_frames and frames > max_frames:
        frames = max_frames
    return frames


def frame_signal(pcm: np.ndarray, N: int, W: int, H: int, frames: int, win: np.ndarray | None) -> np.ndarray:
    """frames x N float32 matrix of windowed, zero-padded frames (fft_cffi.cpp:766-791).
    The product pcm*win is rounded to float32, as the reference's float multiply is."""
    pcm = np.asarray(pcm, dtype=np.float32)
    L = pcm.shape[0]
    out = np.zeros((frames, N), dtype=np.float32)
    for f in range(frames):
        s = f * H
        c = min(W, L - s) if s < L else 0
        if c > 0:
            seg = pcm[s:s + c]
            out[f, :c] = seg * win[:c] if win is not None else seg
    return out


def post_planes(X: np.ndarray, N: int, transform: int, reduce_magnitude: bool, store_polar: bool,
                half_spectrum: bool):
    """(out_real, out_imag, out_mag) in float64 from spectra X (frames x N complex128),
    per eigen_fft.hpp:2535-2596 and fft_cffi.cpp:811-825."""
    X = np.array(X, dtype=np.complex128, copy=True)
    r2c_half = transform == R2C and half_spectrum
    if r2c_half:
        X[:, 0] = X[:, 0].real
        X[:, N // 2] = X[:, N // 2].real
        X[:, N // 2 + 1:] = 0
    bins = N // 2 + 1 if half_spectrum else N
    X = X[:, :bins]
    if store_polar:
        mag = np.abs(X)
        return mag, np.angle(X), mag
    if reduce_magnitude:
        mag = np.abs(X)
        return mag, np.zeros_like(mag), mag
    if transform == C2R:
        re = X.real
        return re, np.zeros_like(re), np.abs(re)
    return X.real, X.imag, np.abs(X)


def stft_planes(pcm, N, W, H, frames, win, inverse, transform, reduce_magnitude, store_polar, half_spectrum):
    """FP64 oracle of fft_execute_batched: framed float32 input, float64 DFT, planes."""
    x = frame_signal(pcm, N, W, H, frames, win)
    X = fft64(x.astype(np.float64), inverse)
    return post_planes(X, N, transform, reduce_magnitude, store_polar, half_spectrum)


def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 over all entries (b is the reference)."""
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def tolerance(N: int) -> float:
    """Relative-L2 bound against this oracle: 2e-7 * log2(N)."""
    return 2e-7 * max(1.0, math.log2(N))


def stft_magnitude_grad(pcm, y_mag, N, W, H, frames, win, grad_phase=None):
    """FP64 gradient of L = sum_{f,k} (|X_fk| - Y_fk)^2 with respect to the float32 PCM,
    for the half spectrum k = 0..N/2 of the windowed frames (hand derivation:
    dL/dx[fH+n] = w[n] * Re(sum_k U_fk e^{+2 pi i kn/N}), U = 2(|X|-Y) X/|X|).
    ``grad_phase`` (frames x bins) adds dL/d(arg X) terms: U += i * Gp * X/|X|^2."""
    pcm = np.asarray(pcm, dtype=np.float64)
    L = pcm.shape[0]
    win64 = np.asarray(win, dtype=np.float64) if win is not None else np.ones(W)
    x = np.zeros((frames, N))
    for f in range(frames):
        s = f * H
        c = min(W, L - s) if s < L else 0
        if c > 0:
            x[f, :c] = pcm[s:s + c] * win64[:c]
    X = np.fft.fft(x, axis=-1)[:, :N // 2 + 1]
    mag = np.abs(X)
    safe = np.where(mag > 0, mag, 1.0)
    U = 2.0 * (mag - y_mag) * X / safe
    if grad_phase is not None:
        U = U + 1j * grad_phase * X / (safe * safe)
    U = np.where(mag > 0, U, 0)
    n = np.arange(N)
    k = np.arange(N // 2 + 1)
    E = np.exp(2j * np.pi * np.outer(k, n) / N)  # bins x N
    g_frames = (U @ E).real  # frames x N
    g = np.zeros(L)
    for f in range(frames):
        s = f * H
        c = min(W, L - s) if s < L else 0
        if c > 0:
            g[s:s + c] += g_frames[f, :c] * win64[:c]
    return g


def _frames64(pcm, N, W, H, frames, win):
    """frames x N float64 matrix of windowed, zero-padded frames and the valid counts."""
    pcm = np.asarray(pcm, dtype=np.float64)
    L = pcm.shape[0]
    w = np.asarray(win, dtype=np.float64) if win is not None else np.ones(N)
    x = np.zeros((frames, N))
    cnt = np.zeros(frames, dtype=np.int64)
    for f in range(frames):
        s = f * H
        c = min(W, L - s) if s < L else 0
        cnt[f] = c
        if c > 0:
            x[f, :c] = pcm[s:s + c] * w[:c]
    return x, cnt, w


def value_planes64(pcm, N, W, H, frames, win, mode):
    """The two value planes of a forward R2C half-spectrum STFT in float64, per output mode
    ('complex': Re X, Im X; 'polar': |X|, arg X; 'magnitude': |X|, 0), with the post
    hook's real DC/Nyquist bins (eigen_fft.hpp:2535-2596)."""
    x, _, _ = _frames64(pcm, N, W, H, frames, win)
    X = np.fft.rfft(x, axis=-1)
    X[:, 0] = X[:, 0].real
    X[:, -1] = X[:, -1].real
    if mode == "complex":
        return X.real, X.imag
    if mode == "polar":
        return np.abs(X), np.angle(X)
    return np.abs(X), np.zeros(X.shape)


def stft_adjoint64(pcm, N, W, H, frames, win, mode, ga, gb=None, ola_window=True):
    """FP64 gradient of sum(ga * A + gb * B) over the value planes (A, B) of
    value_planes64 with respect to the samples (hand derivation):
      dL/dx[fH+n] = w[n] * sum_{k=0}^{N/2} Re(V_fk e^{+2 pi i kn/N}),
      V = ga + i gb (complex), ga X/|X| + i gb X/|X|^2 (polar), ga X/|X| (magnitude),
    V = 0 where |X| = 0 and no imaginary-part terms at bins 0 and N/2 (their imaginary
    parts are pinned to 0). The sum over k is (N/2) * irfft(V') with V'_0 = 2 V_0 and
    V'_{N/2} = 2 V_{N/2}. ola_window=False leaves the window out of the overlap-add: a
    planted mistake for the tests that show the finite-difference check catches it."""
    x, cnt, w = _frames64(pcm, N, W, H, frames, win)
    X = np.fft.rfft(x, axis=-1)
    X[:, 0] = X[:, 0].real
    X[:, -1] = X[:, -1].real
    ga = np.asarray(ga, dtype=np.float64)
    gb = np.zeros_like(ga) if gb is None else np.asarray(gb, dtype=np.float64)
    gb = gb.copy()
    gb[:, 0] = 0.0
    gb[:, -1] = 0.0
    if mode == "complex":
        V = ga + 1j * gb
    else:
        mag = np.abs(X)
        safe = np.where(mag > 0, mag, 1.0)
        V = ga * X / safe
        if mode == "polar":
            V = V + 1j * gb * X / (safe * safe)
        V = np.where(mag > 0, V, 0)
    Vp = V.copy()
    Vp[:, 0] *= 2
    Vp[:, -1] *= 2
    r = (N / 2) * np.fft.irfft(Vp, n=N, axis=-1)
    L = np.asarray(pcm).shape[0]
    g = np.zeros(L)
    for f in range(frames):
        c = cnt[f]
        if c > 0:
            g[f * H:f * H + c] += r[f, :c] * (w[:c] if ola_window else 1.0)
    return g


def complex_columns64(in_real, in_imag, N, transform, half_spectrum):
    """Column rebuild of fft_execute_complex_batched (fft_cffi.cpp:1062-1098) in float64
    from the float32 planes (frames x bins): for a C2R half spectrum imag(X_0) =
    imag(X_{N/2}) = 0 and X_{N-k} = conj X_k (0 < k < N/2); otherwise zeros above the
    given bins."""
    re = np.asarray(in_real, dtype=np.float64)
    im = np.asarray(in_imag, dtype=np.float64)
    frames, bins = re.shape
    cols = np.zeros((frames, N), dtype=np.complex128)
    cols[:, :bins] = re + 1j * im
    if half_spectrum and transform == C2R:
        ny = N // 2
        cols[:, 0] = cols[:, 0].real
        if ny < bins:
            cols[:, ny] = cols[:, ny].real
        maxk = ny if bins > ny + 1 else bins - 1
        if maxk > 1:
            cols[:, N - maxk + 1:] = np.conj(cols[:, 1:maxk][:, ::-1])
    return cols


def complex_batched64(in_real, in_imag, N, inverse=True, kernel=0, transform=C2R, reduce_magnitude=False,
                      store_polar=False, half_spectrum=True, windows_enabled=False, analysis=None, synthesis=None,
                      apply_ola=False, hop=0, cola_mode=0):
    """FP64 oracle of fft_execute_complex_batched (fft_cffi.cpp:1042-1157): column rebuild,
    the plan's pre hook (Cooley-Tukey only, kernel != 2; eigen_fft.hpp:2484-2533), the
    transform (inverse 1/N), the post hook's effect on the real parts
    (eigen_fft.hpp:2535-2596), then the output stage (1119-1157): synthesis window,
    overlap-add over frames*N samples, COLA division where the weight sum > 1e-12
    (weights fft_cffi.cpp:210-226). Returns frames*N floats (float64), shaped (frames, N)
    without overlap-add."""
    cols = complex_columns64(in_real, in_imag, N, transform, half_spectrum)
    frames = cols.shape[0]
    if kernel != 2:
        if transform == R2C:
            cols = cols.real.astype(np.complex128)
        elif transform == C2R and half_spectrum:
            ny = N // 2
            if store_polar:
                mag = cols[:, :ny + 1].real.copy()
                ph = cols[:, :ny + 1].imag.copy()
                cols[:, :ny + 1] = mag * np.cos(ph) + 1j * mag * np.sin(ph)
            cols[:, 0] = cols[:, 0].real
            cols[:, ny] = cols[:, ny].real
            if ny > 1:
                cols[:, N - ny + 1:] = np.conj(cols[:, 1:ny][:, ::-1])
    X = fft64(cols, inverse)
    if transform == R2C and half_spectrum:
        X[:, N // 2 + 1:] = 0
    upper = N // 2 + 1 if (transform == R2C and half_spectrum) else N
    re = X.real.copy()
    if store_polar or reduce_magnitude:
        re[:, :upper] = np.abs(X[:, :upper])
    win = windows_enabled and synthesis is not None
    syn = np.asarray(synthesis, dtype=np.float64) if synthesis is not None else None
    if win:
        re = re * syn[None, :]
    if not apply_ola:
        return re
    total = frames * N
    out = np.zeros(total)
    norm = np.zeros(total) if cola_mode == 1 else None
    if norm is not None:
        if windows_enabled:
            a = np.asarray(analysis, dtype=np.float64) if analysis is not None else np.ones(N)
            b = syn if syn is not None else np.ones(N)
            wgt = a * b
        else:
            wgt = np.ones(N)
    for f in range(frames):
        s = f * hop
        if s >= total:
            break
        c = min(N, total - s)
        out[s:s + c] += re[f, :c]
        if norm is not None:
            norm[s:s + c] += wgt[:c]
    if norm is not None:
        m = norm > 1e-12
        out[m] /= norm[m]
    return out


def cola_norm64(frames, N, hop, weights):
    """Summed COLA weights per output sample (frames*N samples) in float64."""
    total = frames * N
    norm = np.zeros(total)
    w = np.asarray(weights, dtype=np.float64)
    for f in range(frames):
        s = f * hop
        if s >= total:
            break
        c = min(N, total - s)
        norm[s:s + c] += w[:c]
    return norm


def griffin_lim64(mag, N, hop, angles, iterations, fwd_window=0, fwd_hop=0, pad=1, synth=None, analysis=None):
    """FP64 Griffin-Lim (phase_infer.hpp:69-204) for an inverse C2R half-spectrum context
    (synthesis taps or None) and a forward R2C half-spectrum context (window, hop, pad
    rule, analysis taps or None), from given start angles (frames x bins). Per iteration:
    x_f = IDFT_N of the Hermitian rebuild (inverse scale 1/N, imaginary parts of bins 0
    and N/2 dropped) times the synthesis taps; overlap-add averaged over the frames per
    sample; forward STFT; X = mag * (cos, sin)(arg X) for the frames the forward makes,
    imaginary parts of bins 0 and N/2 set to 0. Returns the final spectrum (complex128)."""
    mag = np.asarray(mag, dtype=np.float64)
    frames, bins = mag.shape
    W = fwd_window or N
    Hf = fwd_hop or hop
    X = mag * np.exp(1j * np.asarray(angles, dtype=np.float64).reshape(frames, bins))
    L = (frames - 1) * hop + N
    cnt = np.zeros(L)
    for f in range(frames):
        cnt[f * hop:f * hop + N] += 1
    syn = None if synth is None else np.asarray(synth, dtype=np.float64)
    ana = np.ones(N) if analysis is None else np.asarray(analysis, dtype=np.float64)
    fwd_frames = min(frame_count(L, N, W, Hf, pad, pad, frames), frames)
    for _ in range(iterations):
        fr = np.fft.irfft(X, n=N, axis=-1)
        if syn is not None:
            fr = fr * syn
        ola = np.zeros(L)
        for f in range(frames):
            ola[f * hop:f * hop + N] += fr[f]
        ola = np.where(cnt > 0, ola / np.where(cnt > 0, cnt, 1), ola)
        seg = np.zeros((fwd_frames, N))
        for f in range(fwd_frames):
            s = f * Hf
            c = min(W, L - s)
            seg[f, :c] = ola[s:s + c] * ana[:c]
        Y = np.fft.rfft(seg, axis=-1)
        Xn = mag[:fwd_frames] * np.exp(1j * np.angle(Y))
        Xn[:, 0] = Xn[:, 0].real
        if N % 2 == 0:
            Xn[:, -1] = Xn[:, -1].real
        X = X.copy()
        X[:fwd_frames] = Xn
    return X

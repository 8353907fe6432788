"""Closed forms of the reference's integer metadata. TEST INFRASTRUCTURE ONLY.

Derived from the loops that record them in the reference, not copied from any output:

* bit reversal                      eigen_fft.hpp:2244-2253
* LayoutMap (identity, both kernels) eigen_fft.hpp:2672-2682, 3239-3244
* StageParams (2^{s+1}, N/2^{s+1})   eigen_fft.hpp:2922-2923 (CT), 3476-3479 (Stockham)
* TwiddleIndexMap[s, p]              CT k*(N/len) at a, b (2981-3011);
                                     Stockham j*seg at o0, o1 (3478-3505)
* ButterflyPairs[s, p]               CT (a, b); Stockham (i0, i1)
* Twiddle[(2^s - 1) + j]             W[j * N / 2^{s+1}] (1566-1577)
* Omega[k] = float(2 pi) * k / N     CT only (1752-1762)

Kernel codes follow the C level (fft_cffi.cpp:799-800): 2 is Stockham, anything else is
the Cooley-Tukey default.
"""
from __future__ import annotations

import numpy as np

STOCKHAM = 2
KIND_TWIDDLE, KIND_LAYOUT, KIND_SNAPSHOT, KIND_STAGE_PARAMS, KIND_PAIRS, KIND_INVARIANT, KIND_TW_INDEX, KIND_OMEGA = range(8)


def log2_exact(N: int) -> int:
    s = N.bit_length() - 1
    if N <= 0 or (1 << s) != N:
        raise ValueError("N must be a power of two")
    return s


def bitrev(N: int) -> np.ndarray:
    s = log2_exact(N)
    i = np.arange(N, dtype=np.int64)
    r = np.zeros(N, dtype=np.int64)
    for b in range(s):
        r |= ((i >> b) & 1) << (s - 1 - b)
    return r


def layout_map(N: int) -> np.ndarray:
    return np.arange(N, dtype=np.int64)


def stage_params(N: int) -> np.ndarray:
    """(S, 2): (distance, twiddle step) per stage; identical for both kernels."""
    S = log2_exact(N)
    s = np.arange(S, dtype=np.int64)
    return np.stack([np.left_shift(2, s), N >> (s + 1)], axis=1)


def twiddle_index_map(N: int, kernel: int) -> np.ndarray:
    """(S, N) twiddle index used to produce each position at each stage."""
    S = log2_exact(N)
    p = np.arange(N, dtype=np.int64)
    out = np.zeros((S, N), dtype=np.int64)
    for s in range(S):
        if kernel == STOCKHAM:
            seg = N >> (s + 1)
            pp = p % (N // 2) if N > 1 else p
            out[s] = (pp // seg) * seg
        else:
            half = 1 << s
            out[s] = (p % half) * (N >> (s + 1))
    return out


def butterfly_pairs(N: int, kernel: int) -> np.ndarray:
    """(S, N, 2) input pair (lhs, rhs) feeding each position at each stage."""
    S = log2_exact(N)
    p = np.arange(N, dtype=np.int64)
    out = np.zeros((S, N, 2), dtype=np.int64)
    for s in range(S):
        if kernel == STOCKHAM:
            seg = N >> (s + 1)
            pp = p % (N // 2)
            j, g = pp // seg, pp % seg
            i0 = 2 * j * seg + g
            out[s, :, 0], out[s, :, 1] = i0, i0 + seg
        else:
            half = 1 << s
            a = np.where((p % (2 * half)) < half, p, p - half)
            out[s, :, 0], out[s, :, 1] = a, a + half
    return out


def twiddle_metadata_index(N: int) -> np.ndarray:
    """Plan-twiddle index stored at Twiddle metadata slot (2^s - 1) + j."""
    S = log2_exact(N)
    idx = []
    for s in range(S):
        step = N >> (s + 1)
        idx.extend(j * step for j in range(1 << s))
    return np.asarray(idx, dtype=np.int64)


def omega(N: int, transform: int = 0, half_spectrum: bool = False) -> np.ndarray:
    """float32 omega_k = float(2 pi) * k / N for k < upper (Cooley-Tukey only)."""
    upper = N // 2 + 1 if (transform == 1 and half_spectrum) else N
    two_pi = np.float32(2.0) * np.float32(np.arccos(np.float32(-1.0)))
    k = np.arange(upper, dtype=np.float32)
    return ((two_pi * k).astype(np.float32) / np.float32(N)).astype(np.float32)


def run_butterfly_program(x: np.ndarray, kernel: int, pairs: np.ndarray, tw_index: np.ndarray,
                          inverse: bool = False) -> np.ndarray:
    """Execute the radix-2 network the metadata describes, in complex128 with exact
    twiddles, and return the result (natural order).

    Stockham: out[p] = in[l] +- W^t in[r] (+ for p < N/2); Cooley-Tukey: bit-reversed
    input, then in place y[p] = y[l] +- W^t y[r] (+ when p is the left index). If the
    maps are right this is the DFT; a wrong index or twiddle makes it something else.
    """
    x = np.asarray(x, dtype=np.complex128).copy()
    N = x.shape[-1]
    S = pairs.shape[0]
    sgn = 1.0 if inverse else -1.0
    Wt = np.exp(sgn * 2j * np.pi * np.arange(max(1, N // 2)) / N)
    p = np.arange(N)
    if kernel != STOCKHAM:
        x = x[..., bitrev(N)]
    for s in range(S):
        l, r, t = pairs[s, :, 0], pairs[s, :, 1], tw_index[s]
        prod = Wt[t] * x[..., r]
        if kernel == STOCKHAM:
            plus = p < N // 2
        else:
            plus = p == l
        x = np.where(plus, x[..., l] + prod, x[..., l] - prod)
    return x / N if inverse else x

"""Seeded synthetic inputs of the five benchmark workloads (no datasets are used).

* complex64 batches with real and imaginary parts i.i.d. N(0, 1)
  (1024-point transforms x 64, and 2^20 / 2^22-point transforms x 16);
* mono float32 PCM: 0.1 * uniform[-1, 1] noise plus 8 sinusoids with frequencies
  uniform in (0, 0.5) cycles/sample, amplitudes uniform in [0.1, 1] and random phases;
* the ragged request set: 512 requests, lengths log-uniform in [2^12, 2^20], window
  drawn from {256, 512, 1024, 2048, 4096}, hop = window / 4.
"""
from __future__ import annotations

import numpy as np


def complex_normal(batch: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    re = rng.standard_normal((batch, n), dtype=np.float32)
    im = rng.standard_normal((batch, n), dtype=np.float32)
    return (re + 1j * im).astype(np.complex64)


def noise_plus_sines(length: int, seed: int, n_sines: int = 8) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = 0.1 * rng.uniform(-1.0, 1.0, size=length)
    f = rng.uniform(0.0, 0.5, size=n_sines)
    a = rng.uniform(0.1, 1.0, size=n_sines)
    ph = rng.uniform(0.0, 2 * np.pi, size=n_sines)
    n = np.arange(length, dtype=np.float64)
    for i in range(n_sines):
        # phase advance computed modulo 1 cycle in float64 keeps long signals exact enough
        x += a[i] * np.sin(2 * np.pi * np.mod(f[i] * n, 1.0) + ph[i])
    return x.astype(np.float32)


WINDOWS = (256, 512, 1024, 2048, 4096)


def ragged_requests(count: int = 512, seed: int = 5, lo: int = 12, hi: int = 20):
    """[(length, window, hop)] for the ragged batch."""
    rng = np.random.default_rng(seed)
    lengths = np.exp(rng.uniform(lo * np.log(2), hi * np.log(2), size=count)).astype(np.int64)
    windows = rng.choice(np.array(WINDOWS), size=count)
    return [(int(L), int(W), int(W) // 4) for L, W in zip(lengths, windows)]

"""Writes the golden fixtures of this directory from the pinned oracle (run from the
repository root: python tests/golden/make_golden.py).

The reference ships no golden vectors and its C++ tests draw inputs from libstdc++'s
std::normal_distribution, so these fixtures come from the FP64 oracle (oracle/fp64.py)
and the closed-form index maps (oracle/index_maps.py), both pinned in
tests/test_oracle.py. Each fixture holds the inputs and the expected outputs; float
outputs are the FP64 values rounded to float32 (at most 6e-8 relative, far below the
2e-7*log2(N) bound they are used with).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from fftfree_amd.workloads import complex_normal, noise_plus_sines  # noqa: E402
from oracle import fp64, index_maps as M  # noqa: E402

HERE = Path(__file__).resolve().parent


def c2c_case():
    # plain C2C forward, N=1024, 4 of the config-1 columns (seed 1)
    x = complex_normal(4, 1024, seed=1)
    X = fp64.fft64(x.astype(np.complex128)).astype(np.complex64)
    Xi = fp64.fft64(x.astype(np.complex128), True).astype(np.complex64)
    return {"x": x, "X": X, "X_inv": Xi}


def stft_case():
    # batched R2C STFT, half spectrum, Hann: N=1024/H=256 on 8192 samples and N=64/H=16
    out = {}
    for tag, n, h, length in (("n1024", 1024, 256, 8192), ("n64", 64, 16, 1000)):
        pcm = noise_plus_sines(length, seed=n)
        frames = fp64.frame_count(length, n, n, h, 0, 0)
        win = fp64.build_window(fp64.WINDOW_HANN, n)
        re, im, mag = fp64.stft_planes(pcm, n, n, h, frames, win, False, 1, False, False, True)
        out[f"{tag}_pcm"] = pcm
        out[f"{tag}_re"], out[f"{tag}_im"] = re.astype(np.float32), im.astype(np.float32)
    return out


def index_case():
    out = {}
    for n in (8, 64):
        out[f"bitrev_{n}"] = M.bitrev(n)
        out[f"layout_{n}"] = M.layout_map(n)
        out[f"stage_params_{n}"] = M.stage_params(n)
        for kernel in (0, 2):
            out[f"tw_index_{n}_k{kernel}"] = M.twiddle_index_map(n, kernel)
            out[f"pairs_{n}_k{kernel}"] = M.butterfly_pairs(n, kernel)
    # frame counts: (pcm_len, N, window, hop, ctx_pad, pad_mode, max_frames) -> frames
    rows = []
    for L in (0, 100, 1023, 1024, 1025, 1279, 1280, 5000, 1 << 24):
        for (n, w, h) in ((1024, 1024, 256), (1024, 512, 100), (64, 64, 16), (256, 512, 128)):
            for pad in (0, 1, 2):
                for mx in (0, 3):
                    rows.append((L, n, w, h, 0, pad, mx, fp64.frame_count(L, n, w, h, 0, pad, mx)))
    out["frame_counts"] = np.array(rows, dtype=np.int64)
    for kind in range(5):
        out[f"window_{kind}_33"] = fp64.build_window(kind, 33)
    return out


def main():
    np.savez_compressed(HERE / "c2c_n1024.npz", **c2c_case())
    np.savez_compressed(HERE / "stft_hann.npz", **stft_case())
    np.savez_compressed(HERE / "index_maps.npz", **index_case())


if __name__ == "__main__":
    main()

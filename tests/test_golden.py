"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
pinned oracle): the oracle still reproduces them, the restated reference path agrees
with them, and the C ABI's host-side frame rule matches them. CPU only."""
from pathlib import Path

import numpy as np
import pytest

from fftfree_amd.workloads import complex_normal, noise_plus_sines
from oracle import fp64, index_maps as M, ref

G = Path(__file__).resolve().parent / "golden"


def load(name):
    return np.load(G / name, allow_pickle=False)


def test_golden_c2c_reproduced_by_oracle_and_cpu_ref():
    g = load("c2c_n1024.npz")
    assert np.array_equal(g["x"], complex_normal(4, 1024, seed=1))
    assert fp64.rel_l2(fp64.fft64(g["x"].astype(np.complex128)), g["X"]) <= 1e-7
    assert fp64.rel_l2(ref.fft_batched(g["x"], kernel=2), g["X"]) <= fp64.tolerance(1024)
    assert fp64.rel_l2(ref.fft_batched(g["x"], inverse=True, kernel=0), g["X_inv"]) <= fp64.tolerance(1024)


@pytest.mark.parametrize("tag,n,h", [("n1024", 1024, 256), ("n64", 64, 16)])
def test_golden_stft_reproduced_by_oracle_and_cpu_ref(tag, n, h):
    g = load("stft_hann.npz")
    pcm = g[f"{tag}_pcm"]
    assert np.array_equal(pcm, noise_plus_sines(pcm.size, seed=n))
    frames = g[f"{tag}_re"].shape[0]
    win = fp64.build_window(fp64.WINDOW_HANN, n)
    re, im, _ = fp64.stft_planes(pcm, n, n, h, frames, win, False, 1, False, False, True)
    assert fp64.rel_l2(re + 1j * im, g[f"{tag}_re"] + 1j * g[f"{tag}_im"]) <= 1e-7
    got, rre, rim, _ = ref.execute_batched(pcm, n, n, h, pad_mode=0, kernel=2, transform=1, half_spectrum=True,
                                           window_taps=win)
    assert got == frames
    assert fp64.rel_l2(rre + 1j * rim, g[f"{tag}_re"] + 1j * g[f"{tag}_im"]) <= fp64.tolerance(n)


def test_golden_index_maps_and_windows():
    g = load("index_maps.npz")
    for n in (8, 64):
        assert np.array_equal(g[f"bitrev_{n}"], M.bitrev(n)) and np.array_equal(g[f"bitrev_{n}"], ref.bitrev(n))
        assert np.array_equal(g[f"layout_{n}"], M.layout_map(n))
        assert np.array_equal(g[f"stage_params_{n}"], M.stage_params(n))
        for k in (0, 2):
            assert np.array_equal(g[f"tw_index_{n}_k{k}"], M.twiddle_index_map(n, k))
            assert np.array_equal(g[f"pairs_{n}_k{k}"], M.butterfly_pairs(n, k))
    for kind in range(5):
        w = g[f"window_{kind}_33"]
        assert np.array_equal(w.view(np.uint32), fp64.build_window(kind, 33).view(np.uint32))
        assert np.array_equal(w.view(np.uint32), ref.build_window(kind, 33).view(np.uint32))


def test_golden_frame_counts_oracle_cpu_ref_and_abi():
    from fftfree_amd import _lib
    lib = _lib.load()
    g = load("index_maps.npz")["frame_counts"]
    for L, n, w, h, ctx_pad, pad, mx, want in g.tolist():
        assert fp64.frame_count(L, n, w, h, ctx_pad, pad, mx) == want
        assert ref.frame_count(L, n, w, h, ctx_pad, pad, mx) == want
        handle = lib.fft_init_ex(n, 1, 1, 0, 0, 0, None, 0, ctx_pad, w, h)
        assert handle
        try:
            assert lib.fft_frame_count(handle, L, pad, mx) == want, (L, n, w, h, pad, mx)
        finally:
            lib.fft_free(handle)

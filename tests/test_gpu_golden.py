"""The HIP path (through the C ABI) against the committed golden fixtures.

Tolerance for floats: relative L2 <= 2e-7 * log2(N) (fp64.tolerance); indices and
frame counts bit-exact."""
from pathlib import Path

import numpy as np
import pytest

from fftfree_amd import CffiFft
from oracle import fp64

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


@pytest.mark.parametrize("kernel", [0, 2])
def test_gpu_c2c_vs_golden(ff, kernel):
    g = np.load(G / "c2c_n1024.npz")
    fwd = ff.make_plan(1024, kernel=kernel, transform=0, half_spectrum=False)
    inv = ff.make_plan(1024, inverse=1, kernel=kernel, transform=0, half_spectrum=False)
    assert fp64.rel_l2(ff.c2c(fwd, g["x"]), g["X"]) <= fp64.tolerance(1024)
    assert fp64.rel_l2(ff.c2c(inv, g["x"]), g["X_inv"]) <= fp64.tolerance(1024)


@pytest.mark.parametrize("tag,n,h", [("n1024", 1024, 256), ("n64", 64, 16)])
@pytest.mark.parametrize("kernel", [0, 2])
def test_gpu_stft_vs_golden(ff, tag, n, h, kernel):
    g = np.load(G / "stft_hann.npz")
    pcm = g[f"{tag}_pcm"]
    p = ff.make_plan(n, window=n, hop=h, kernel=kernel, transform=1, half_spectrum=True, apply_windows=True,
                     analysis_window=1, pad_mode=0)
    got, re, im, mg = ff.stft_frames(p, pcm, pad_mode=0)
    assert got == g[f"{tag}_re"].shape[0]
    want = g[f"{tag}_re"] + 1j * g[f"{tag}_im"]
    assert fp64.rel_l2(re + 1j * im, want) <= fp64.tolerance(n)
    assert fp64.rel_l2(mg, np.abs(want)) <= fp64.tolerance(n)


@pytest.mark.parametrize("n", [8, 64])
@pytest.mark.parametrize("kernel", [0, 2])
def test_gpu_metadata_vs_golden(ff, n, kernel):
    g = np.load(G / "index_maps.npz")
    S = int(np.log2(n))
    p = ff.make_plan(n, kernel=kernel, transform=0, half_spectrum=False)
    lm = ff.metadata(p, 1)
    assert np.array_equal(lm.real.astype(np.int64), g[f"layout_{n}"]) and not lm.imag.any()
    ti = ff.metadata(p, 6)
    assert np.array_equal(ti.real.astype(np.int64).reshape(S, n), g[f"tw_index_{n}_k{kernel}"])
    bp = ff.metadata(p, 4)
    assert np.array_equal(np.stack([bp.real, bp.imag], -1).astype(np.int64).reshape(S, n, 2),
                          g[f"pairs_{n}_k{kernel}"])
    sp = ff.metadata(p, 3)
    assert np.array_equal(np.stack([sp.real, sp.imag], -1).astype(np.int64), g[f"stage_params_{n}"])

"""fft_stft_adjoint_device_noalloc, the allocation-free variant of fft_stft_adjoint_device:
through the C ABI it gives the same gradient, bit for bit (the same kernels on the same
inputs; only the halo slots live in the library's static scratch), on two calls in a
row, the second smaller; it refuses a context without device state and a shape whose
halo slots exceed the static scratch, and leaves the output untouched then."""
import ctypes

import numpy as np
import pytest

from fftfree_amd import CffiFft
from fftfree_amd.workloads import noise_plus_sines

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


def plan(ff, n=2048, W=2048, H=512):
    return ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, half_spectrum=True, store_polar=True,
                        apply_windows=True, analysis_window=1, pad_mode=1)


def call(ff, fn, p, x, gmag, out):
    import torch
    s = torch.cuda.current_stream().cuda_stream
    return fn(p.handle, x.data_ptr(), x.shape[0], x.shape[1], gmag.data_ptr(), None, out.data_ptr(), 1, s)


def test_noalloc_adjoint_matches_allocating_entry(ff):
    import torch
    p = plan(ff)
    lib = ff.lib
    for sig, L in ((32, 1 << 18), (4, 1 << 16)):  # the differentiable-STFT workload, then smaller
        print(f"noalloc adjoint signals={sig} len={L}", flush=True)
        x = torch.from_numpy(np.stack([noise_plus_sines(L, seed=100 + i) for i in range(sig)])).cuda()
        frames = ff.frame_count(p, L, 1)
        mag = torch.empty((sig, frames, 1025), device="cuda")
        ph = torch.empty_like(mag)
        ff.stft_signals_device(p, x, mag, ph, None, pad_mode=1)  # the forward: the context's device state
        gmag = torch.randn_like(mag)
        want = torch.full_like(x, float("nan"))
        got = torch.full_like(x, float("nan"))
        assert call(ff, lib.fft_stft_adjoint_device, p, x, gmag, want) == frames
        assert call(ff, lib.fft_stft_adjoint_device_noalloc, p, x, gmag, got) == frames
        torch.cuda.synchronize()
        assert torch.isfinite(want).all()
        assert torch.equal(got.view(torch.int32), want.view(torch.int32))


def test_noalloc_adjoint_refuses_without_device_state(ff):
    import torch
    p = plan(ff)  # fresh context: no device call made on it yet
    x = torch.from_numpy(noise_plus_sines(1 << 14, seed=1)[None, :].copy()).cuda()
    frames = ff.frame_count(p, 1 << 14, 1)
    gmag = torch.ones((1, frames, 1025), device="cuda")
    out = torch.full_like(x, 7.0)
    assert call(ff, ff.lib.fft_stft_adjoint_device_noalloc, p, x, gmag, out) == 0
    torch.cuda.synchronize()
    assert (out == 7.0).all()


def test_noalloc_adjoint_refuses_shapes_beyond_its_scratch(ff):
    """One signal of 2^27 samples: runs are capped at 64 steps of 2 frames, so it takes
    2^18 / 128 = 2048 runs and 2048 x 2 x (W - H) = 6.3M halo floats, beyond the 4M of
    the static scratch, whatever the occupancy. Refused before any launch; the
    allocating entry takes it."""
    import torch
    p = plan(ff)
    L = 1 << 27
    small = torch.from_numpy(noise_plus_sines(1 << 14, seed=1)[None, :].copy()).cuda()
    fs = ff.frame_count(p, 1 << 14, 1)
    ff.stft_signals_device(p, small, torch.empty((1, fs, 1025), device="cuda"), None, None, pad_mode=1)
    x = torch.zeros((1, L), device="cuda")
    x[0, ::7] = 1.0
    frames = ff.frame_count(p, L, 1)
    gmag = torch.ones((1, frames, 1025), device="cuda")
    out = torch.full_like(x, 7.0)
    assert call(ff, ff.lib.fft_stft_adjoint_device_noalloc, p, x, gmag, out) == 0
    torch.cuda.synchronize()
    assert (out == 7.0).all()
    assert call(ff, ff.lib.fft_stft_adjoint_device, p, x, gmag, out) == frames
    torch.cuda.synchronize()
    assert torch.isfinite(out).all()

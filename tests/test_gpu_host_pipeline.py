"""fft_execute_batched on host pointers (reference: fft_cffi.cpp:713-804) is staged: the
frames run in chunks of about 24 MiB of planes, the samples go up and the planes come
down through pinned bounce slots on their own streams, and host threads copy them in and
out. None of that may change a value: the planes of the host entry must equal, bit for
bit, those of one device-entry launch over all frames with the same kernel variant
pinned (chunks are a multiple of every variant's tile, so chunk edges are tile edges of
the single launch). Covered: the N = 1024 STFT workload at full size (16 chunks and a
partial one), a frame count of one chunk plus one, a zero-padded last frame (pad_mode 1), a window
shorter than N, a max_frames cut, a full-spectrum plan, a single tiny chunk, polar
planes, and one context called again with a larger and then a smaller input.
"""
import numpy as np
import pytest

from fftfree_amd import CffiFft
from fftfree_amd import _lib
from fftfree_amd.workloads import noise_plus_sines

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


def fp(a):
    return a.ctypes.data_as(_lib.f32p)


def chunk_frames(bins):
    return max(192, (24 << 20) // (12 * bins) // 192 * 192)


def host_vs_device(ff, p, pcm, pad_mode, max_frames, bins):
    frames = ff.frame_count(p, len(pcm), pad_mode, max_frames)
    assert frames > 0
    variant = ff.default_variant(p, frames)
    assert ff.set_variant(p, variant)
    host = np.full((3, frames * bins), np.nan, np.float32)
    got = _lib.load().fft_execute_batched(p.handle, fp(pcm), len(pcm), fp(host[0]), fp(host[1]), fp(host[2]),
                                          pad_mode, 0, max_frames)
    assert got == frames
    dev = [torch.empty((frames, bins), device="cuda") for _ in range(3)]
    ff.stft_device(p, torch.from_numpy(pcm).cuda(), dev[0], dev[1], dev[2], pad_mode=pad_mode, max_frames=max_frames)
    torch.cuda.synchronize()
    for i in range(3):
        d = dev[i].cpu().numpy().reshape(-1)
        assert np.array_equal(host[i], d), (i, int(np.argmax(host[i] != d)) // bins)
    return frames


CASES = [
    # n, W, H, samples, pad_mode, max_frames, half_spectrum, reduce_magnitude, polar
    (1024, 1024, 256, 1 << 24, 0, 0, True, True, False),
    (1024, 1024, 256, 4032 * 256 + 1024, 0, 0, True, True, False),
    (1024, 1024, 256, 4000 * 256 + 1024 + 700, 1, 0, True, False, False),
    (1024, 768, 200, 3_000_001, 1, 0, True, True, False),
    (1024, 1024, 256, 1 << 22, 0, 5000, True, True, False),
    (2048, 2048, 512, 1 << 22, 0, 0, False, False, False),
    (512, 512, 128, 1000, 1, 0, True, True, False),
    (2048, 2048, 512, 1 << 21, 1, 0, True, False, True),
]


@pytest.mark.parametrize("n,W,H,L,pad,maxf,half,mag,polar", CASES)
def test_host_entry_equals_one_device_launch(ff, n, W, H, L, pad, maxf, half, mag, polar):
    p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, reduce_magnitude=mag, store_polar=polar,
                     half_spectrum=half, apply_windows=True, analysis_window=1, pad_mode=pad)
    bins = n // 2 + 1 if half else n
    pcm = noise_plus_sines(L, seed=7)
    frames = host_vs_device(ff, p, pcm, pad, maxf, bins)
    print(f"n={n} W={W} H={H} samples={L} frames={frames} chunks={-(-frames // chunk_frames(bins))}")


def test_one_context_called_with_growing_and_shrinking_inputs(ff):
    p = ff.make_plan(1024, window=1024, hop=256, kernel=2, transform=1, reduce_magnitude=True, half_spectrum=True,
                     apply_windows=True, analysis_window=1, pad_mode=0)
    for L in (5000, 1 << 23, 300_000, 1 << 20):
        host_vs_device(ff, p, noise_plus_sines(L, seed=L % 97), 0, 0, 513)

"""fft_execute_complex_batched on the device (the fused inverse kernel and the general
path: column rebuild, the plan's transform with its hooks, gather overlap-add) against
the FP64 oracle and the float restatement of the reference.

Bar: rel-L2 <= 2 * fp64.tolerance(N) (2e-7 * log2 N per transform, doubled for the
overlap-add of several frames) against the FP64 oracle on the same float32 inputs. With
COLA normalisation a sample whose weight sum is tiny (the first taps of a Hann window
give sums near 1e-10) divides float rounding noise by that sum, whatever the algorithm:
those comparisons cover the samples whose sum is above 1e-3 of its maximum, and the
others are only required to be finite. Samples past the overlap-add, up to frames*N,
must be exactly zero, as the reference zero-fills them (fft_cffi.cpp:1124-1125)."""
import ctypes

import numpy as np
import pytest

from fftfree_amd import CffiFft
from fftfree_amd.workloads import noise_plus_sines
from oracle import fp64, ref

pytestmark = pytest.mark.gpu
C2C, R2C, C2R = 0, 1, 2


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


def taps(n, W, kind):
    return np.concatenate([fp64.build_window(kind, W), np.ones(n - W, np.float32)])


def planes(frames, bins, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((frames, bins)).astype(np.float32),
            rng.standard_normal((frames, bins)).astype(np.float32))


def run_device(ff, p, re, im):
    import torch
    d_re = torch.from_numpy(re).cuda()
    d_im = torch.from_numpy(im).cuda()
    out = torch.full((re.shape[0] * p.n,), float("nan"), device="cuda")
    ff.complex_batched_device(p, d_re, d_im, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def check(out, exp, n, frames, H, ola, cola, weights):
    exp = np.asarray(exp, np.float64).ravel()
    assert out.shape == exp.shape and np.isfinite(out).all()
    m = np.ones(out.size, bool)
    if ola:
        L = min((frames - 1) * H + n, frames * n)
        assert not out[L:].any(), "samples past the overlap-add must be zero"
        if cola:
            norm = fp64.cola_norm64(frames, n, H, weights)
            m = norm > 1e-3 * norm.max()
    err = fp64.rel_l2(out[m], exp[m])
    assert err <= 2 * fp64.tolerance(n), err


FUSED = [(512, 128, 37), (1024, 256, 700), (2048, 512, 300), (1024, 1024, 50), (1024, 512, 9), (1024, 255, 301),
         (512, 384, 40), (2048, 2048, 3), (1024, 256, 1), (1024, 256, 2), (1024, 100, 5000)]


@pytest.mark.parametrize("case", FUSED)
@pytest.mark.parametrize("mode", [(1, 1, 1), (1, 0, 1), (0, 0, 1), (1, 1, 0), (1, 0, 0), (0, 0, 0)])
@pytest.mark.parametrize("kernel", [0, 2])
def test_fused_inverse_vs_oracle(ff, case, mode, kernel):
    """Inverse C2R half-spectrum contexts (the fused kernel) across N, hop, frame counts
    (one run, many runs with warm-up frames, odd hop), overlap-add / COLA / frame-major
    output, windows on and off, both kernel codes."""
    n, H, frames = case
    ola, cola, windows = mode
    print(f"istft fused n={n} H={H} frames={frames} ola={ola} cola={cola} win={windows} kernel={kernel}", flush=True)
    p = ff.make_plan(n, window=n, hop=H, inverse=1, kernel=kernel, transform=C2R, half_spectrum=True,
                     apply_windows=bool(windows), apply_ola=bool(ola), analysis_window=fp64.WINDOW_HANN,
                     synthesis_window=fp64.WINDOW_HANN, cola=cola)
    re, im = planes(frames, n // 2 + 1, seed=n + H + frames)
    out = run_device(ff, p, re, im)
    w = taps(n, n, fp64.WINDOW_HANN) if windows else None
    exp = fp64.complex_batched64(re, im, n, kernel=kernel, windows_enabled=bool(windows), analysis=w, synthesis=w,
                                 apply_ola=bool(ola), hop=H, cola_mode=cola)
    weights = (w.astype(np.float64) ** 2) if windows else np.ones(n)
    check(out, exp, n, frames, H, ola, cola, weights)


GENERAL = [
    # (n, H, frames, inverse, kernel, transform, half, store_polar, reduce_magnitude)
    (256, 64, 40, 1, 2, C2R, True, False, False),
    (4096, 1024, 12, 1, 2, C2R, True, False, False),
    (8192, 2048, 6, 1, 2, C2R, True, False, False),
    (1024, 256, 30, 1, 0, C2R, True, True, False),
    (1024, 256, 30, 1, 2, C2R, True, True, False),
    (1024, 256, 30, 1, 2, C2R, True, False, True),
    (1024, 256, 30, 1, 2, C2C, True, False, False),
    (512, 128, 30, 1, 0, C2C, False, False, False),
    (1024, 256, 30, 0, 2, C2R, True, False, False),
    (1024, 1500, 20, 1, 2, C2R, True, False, False),
    (4096, 1024, 9, 1, 0, C2R, True, True, False),
]


@pytest.mark.parametrize("case", GENERAL)
@pytest.mark.parametrize("mode", [(1, 1, 1), (1, 0, 0), (0, 0, 1)])
def test_general_path_vs_oracle(ff, case, mode):
    """Contexts the fused kernel does not cover: other N (including the two-pass large
    transform), polar input through the Cooley-Tukey pre hook or straight through
    Stockham, polar/magnitude post hooks, complex plans on a zero-filled half spectrum or
    a full one, a forward plan, and a hop above N (gaps in the overlap-add)."""
    n, H, frames, inverse, kernel, transform, half, polar, mag = case
    ola, cola, windows = mode
    print(f"istft general case={case} mode={mode}", flush=True)
    p = ff.make_plan(n, window=n, hop=H, inverse=inverse, kernel=kernel, transform=transform, half_spectrum=half,
                     store_polar=polar, reduce_magnitude=mag, apply_windows=bool(windows), apply_ola=bool(ola),
                     analysis_window=fp64.WINDOW_HAMMING, synthesis_window=fp64.WINDOW_HAMMING, cola=cola)
    bins = n // 2 + 1 if half else n
    re, im = planes(frames, bins, seed=n + frames + 7 * kernel)
    if polar:
        re = np.abs(re)
    out = run_device(ff, p, re, im)
    w = taps(n, n, fp64.WINDOW_HAMMING) if windows else None
    exp = fp64.complex_batched64(re, im, n, inverse=bool(inverse), kernel=kernel, transform=transform,
                                 reduce_magnitude=mag, store_polar=polar, half_spectrum=half,
                                 windows_enabled=bool(windows), analysis=w, synthesis=w, apply_ola=bool(ola), hop=H,
                                 cola_mode=cola)
    weights = (w.astype(np.float64) ** 2) if windows else np.ones(n)
    check(out, exp, n, frames, H, ola, cola, weights)


def test_host_entry_matches_device_entry_and_float_restatement(ff):
    """The host-pointer entry runs the same kernels (bitwise equal output), and both agree
    with the float restatement of the reference path (cpu_ref)."""
    n, H, frames = 1024, 256, 123
    p = ff.make_plan(n, window=n, hop=H, inverse=1, kernel=2, transform=C2R, half_spectrum=True,
                     apply_windows=True, apply_ola=True, analysis_window=fp64.WINDOW_HANN,
                     synthesis_window=fp64.WINDOW_HANN, cola=1)
    re, im = planes(frames, n // 2 + 1, seed=3)
    dev = run_device(ff, p, re, im)
    host = np.full(frames * n, np.nan, np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert ff.lib.fft_execute_complex_batched(p.handle, fp(re), fp(im), frames, fp(host), 1, 0, 0) == frames
    assert np.array_equal(host, dev)
    w = taps(n, n, fp64.WINDOW_HANN)
    got, cref = ref.execute_complex_batched(re, im, n, kernel=2, synth_taps=w, analysis_taps=w, apply_ola=True,
                                            hop=H, cola_mode=1)
    assert got == frames
    check(dev, cref, n, frames, H, 1, 1, w.astype(np.float64) ** 2)


def test_max_frames_and_refusals(ff):
    import torch
    n, H = 1024, 256
    p = ff.make_plan(n, window=n, hop=H, inverse=1, kernel=2, transform=C2R, half_spectrum=True, apply_ola=True)
    re, im = planes(10, n // 2 + 1, seed=4)
    d_re, d_im = torch.from_numpy(re).cuda(), torch.from_numpy(im).cuda()
    out = torch.full((10 * n,), 7.0, device="cuda")
    lib = ff.lib
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.fft_execute_complex_batched_device(p.handle, vp(d_re), vp(d_im), 10, vp(out), 4, None) == 4
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    exp = fp64.complex_batched64(re[:4], im[:4], n, apply_ola=True, hop=H)
    check(o[:4 * n], exp, n, 4, H, 1, 0, np.ones(n))
    assert (o[4 * n:] == 7.0).all(), "nothing past frames*N is written"
    assert lib.fft_execute_complex_batched_device(p.handle, vp(d_re), vp(d_im), 0, vp(out), 0, None) == 0
    assert lib.fft_execute_complex_batched_device(p.handle, None, vp(d_im), 10, vp(out), 0, None) == 0
    assert lib.fft_execute_complex_batched_device(None, vp(d_re), vp(d_im), 10, vp(out), 0, None) == 0


def test_full_size_round_trip():
    """STFT then inverse STFT at the headline size (2^24 samples, N=1024, Hann, hop 256,
    65533 frames): with Hann analysis and synthesis and COLA normalisation the output is
    the input signal wherever the weight sum is not tiny; two runs are bitwise equal."""
    import torch
    ff = CffiFft()
    try:
        n, H, L = 1024, 256, 1 << 24
        fwd = ff.make_plan(n, window=n, hop=H, kernel=2, transform=R2C, half_spectrum=True, apply_windows=True,
                           analysis_window=fp64.WINDOW_HANN, pad_mode=0)
        inv = ff.make_plan(n, window=n, hop=H, inverse=1, kernel=2, transform=C2R, half_spectrum=True,
                           apply_windows=True, apply_ola=True, analysis_window=fp64.WINDOW_HANN,
                           synthesis_window=fp64.WINDOW_HANN, cola=1)
        pcm = noise_plus_sines(L, seed=2)
        d = torch.from_numpy(pcm).cuda()
        frames = 65533
        re = torch.empty((frames, n // 2 + 1), device="cuda")
        im = torch.empty_like(re)
        assert ff.stft_device(fwd, d, re, im, None, pad_mode=0) == frames
        out = ff.complex_batched_device(inv, re, im)
        out2 = ff.complex_batched_device(inv, re, im)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        assert np.array_equal(o, out2.cpu().numpy())
        span = (frames - 1) * H + n
        assert not o[span:].any()
        w = taps(n, n, fp64.WINDOW_HANN).astype(np.float64)
        norm = fp64.cola_norm64(frames, n, H, w * w)[:span]
        m = norm > 1e-3 * norm.max()
        err = fp64.rel_l2(o[:span][m], pcm[:span][m])
        print(f"round trip rel-L2 {err:.3e} over {m.sum()} samples", flush=True)
        assert err <= 4 * fp64.tolerance(n)
    finally:
        ff.close()

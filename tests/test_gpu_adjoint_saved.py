"""The adjoint from the forward's saved planes (fft_stft_adjoint_saved_device): V is formed
from the upstream gradients and the forward's own |X|, arg X (polar) or from the gradients
alone (complex), then one inverse DFT, the analysis window and the overlap-add per frame.

Oracle: the FP64 adjoint (fp64.stft_adjoint64), tolerance rel-L2 <= 4*fp64.tolerance(N)
as for the recomputing adjoint; that adjoint itself agrees to the same bound. Upstream
gradients of the polar map are weighted as in test_gpu_adjoint.py (bins below 1% of the
RMS magnitude out, the phase gradient scaled by |X|/RMS), since X/|X| is ill-conditioned
where |X| is near 0 whatever the algorithm."""
import numpy as np
import pytest

from fftfree_amd import CffiFft
from fftfree_amd.workloads import noise_plus_sines
from oracle import fp64

pytestmark = pytest.mark.gpu
MODES = {"complex": dict(), "polar": dict(store_polar=True), "magnitude": dict(reduce_magnitude=True)}

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


def make(ff, n, W, H, mode, pad=1, window_kind=1):
    return ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, half_spectrum=True, apply_windows=window_kind != 0,
                        analysis_window=window_kind, pad_mode=pad, **MODES[mode])


def signals(count, length, seed):
    return np.stack([noise_plus_sines(length, seed=seed + i) for i in range(count)])


def win_n(n, W, kind=1):
    if not kind:
        return None
    return np.concatenate([fp64.build_window(kind, W), np.ones(n - W, np.float32)])


def grads(x, n, W, H, frames, mode, seed, kind=1):
    bins = n // 2 + 1
    rng = np.random.default_rng(seed)
    ga = rng.standard_normal((x.shape[0], frames, bins))
    gb = rng.standard_normal((x.shape[0], frames, bins))
    if mode != "complex":
        for s in range(x.shape[0]):
            A, _ = fp64.value_planes64(x[s], n, W, H, frames, win_n(n, W, kind), "magnitude")
            rms = np.sqrt(np.mean(A ** 2))
            keep = A > 1e-2 * rms
            ga[s] *= keep
            gb[s] *= keep * A / rms
    return ga.astype(np.float32), gb.astype(np.float32)


def forward(ff, p, d, frames, n):
    bins = n // 2 + 1
    re = torch.full((d.shape[0], frames, bins), float("nan"), device="cuda")
    im = torch.full_like(re, float("nan"))
    assert ff.stft_signals_device(p, d, re, im, None) == frames
    return re, im


# (n, W, H, samples, context pad mode)
CASES = [(512, 512, 128, 5000, 1), (1024, 1024, 256, 9000, 1), (2048, 2048, 512, 20000, 1), (1024, 768, 200, 7001, 1),
         (2048, 2048, 512, 1500, 1), (1024, 1024, 512, 9000, 1), (1024, 1024, 1024, 9000, 1), (512, 300, 77, 6000, 1),
         (2048, 2048, 128, 50000, 1), (512, 512, 64, 30000, 1), (1024, 1024, 256, 9100, 2), (2048, 1500, 333, 12345, 2),
         (512, 256, 700, 6000, 1)]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("mode", ["complex", "polar"])
def test_saved_adjoint_vs_fp64_and_recompute(ff, case, mode):
    n, W, H, length, pad = case
    print(f"saved adjoint n={n} W={W} H={H} len={length} pad={pad} mode={mode}", flush=True)
    p = make(ff, n, W, H, mode, pad=pad)
    x = signals(3, length, seed=3 * n + W)
    frames = ff.frame_count(p, length, 0)
    ga, gb = grads(x, n, W, H, frames, mode, seed=n + H)
    d = torch.from_numpy(x).cuda()
    re, im = forward(ff, p, d, frames, n)
    tga, tgb = torch.from_numpy(ga).cuda(), torch.from_numpy(gb).cuda()
    g = torch.full_like(d, float("nan"))
    ff.stft_adjoint_saved_device(p, length, re, im, tga, tgb, g)
    rec = ff.stft_adjoint_device(p, d, tga, tgb)
    torch.cuda.synchronize()
    got, got_rec = g.cpu().numpy(), rec.cpu().numpy()
    for s in range(3):
        want = fp64.stft_adjoint64(x[s], n, W, H, frames, win_n(n, W), mode, ga[s], gb[s])
        assert fp64.rel_l2(got[s], want) <= 4 * fp64.tolerance(n), s
        assert fp64.rel_l2(got[s], got_rec[s]) <= 4 * fp64.tolerance(n), s
    # samples no frame reaches (pad mode 2 drops the partial last frame) are zero
    covered = (frames - 1) * H + W
    if covered < length:
        assert not got[:, covered:].any()
    # a second run gives the same bits (fixed summation order)
    g2 = torch.empty_like(d)
    ff.stft_adjoint_saved_device(p, length, re, im, tga, tgb, g2)
    torch.cuda.synchronize()
    assert torch.equal(g, g2)


def test_saved_adjoint_single_gradients_no_window_and_zero_frames(ff):
    """Polar map: only the phase gradient, only the magnitude gradient, none (all zero,
    from a fill kernel: the plane kernel is not launched without a gradient plane);
    no analysis window; signals with silent stretches, where whole frames have |X| = 0
    (V = 0 there, as the recomputing adjoint and the oracle define it)."""
    n, W, H, length = 1024, 1024, 256, 8000
    p = make(ff, n, W, H, "polar", window_kind=0)
    x = signals(2, length, seed=91)
    x[0, 2000:4500] = 0.0
    x[1, :3000] = 0.0
    frames = ff.frame_count(p, length, 0)
    ga, gb = grads(x, n, W, H, frames, "polar", seed=5, kind=0)
    d = torch.from_numpy(x).cuda()
    re, im = forward(ff, p, d, frames, n)
    for use_a, use_b in ((False, True), (True, False)):
        ta = torch.from_numpy(ga).cuda() if use_a else None
        tb = torch.from_numpy(gb).cuda() if use_b else None
        g = ff.stft_adjoint_saved_device(p, length, re, im, ta, tb)
        rec = ff.stft_adjoint_device(p, d, ta, tb)
        torch.cuda.synchronize()
        for s in range(2):
            want = fp64.stft_adjoint64(x[s], n, W, H, frames, None, "polar", ga[s] if use_a else np.zeros_like(ga[s]),
                                       gb[s] if use_b else None)
            assert fp64.rel_l2(g[s].cpu().numpy(), want) <= 4 * fp64.tolerance(n), (use_a, use_b, s)
            assert fp64.rel_l2(g[s].cpu().numpy(), rec[s].cpu().numpy()) <= 4 * fp64.tolerance(n)
    # frames entirely inside the silent stretch have |X| = 0: their V is 0, so a gradient
    # placed only there gives dL/dx = 0
    zf = [f for f in range(frames) if not np.abs(x[1, f * H:f * H + W]).any()]
    assert zf
    only = np.zeros_like(ga)
    only[1, zf] = 1.0
    g = ff.stft_adjoint_saved_device(p, length, re, im, torch.from_numpy(only).cuda(), None)
    torch.cuda.synchronize()
    assert not g.cpu().numpy().any()
    g = torch.full_like(d, float("nan"))
    ff.stft_adjoint_saved_device(p, length, re, im, None, None, g)
    torch.cuda.synchronize()
    assert not g.cpu().numpy().any()


def test_saved_adjoint_refusals(ff):
    d = torch.zeros((1, 10000), device="cuda")
    z = torch.zeros((1, 7, 513), device="cuda")
    p = make(ff, 1024, 1024, 256, "magnitude")  # no direction of X in magnitude planes
    with pytest.raises(RuntimeError):
        ff.stft_adjoint_saved_device(p, 10000, z, z, z, None)
    p = make(ff, 1024, 1024, 256, "polar")  # polar planes required
    with pytest.raises(RuntimeError):
        ff.stft_adjoint_saved_device(p, 10000, None, None, z, None, out=torch.empty_like(d))
    p = make(ff, 1024, 1024, 2048, "polar")  # one step spans more than the adjoint's 8N ring
    with pytest.raises(RuntimeError):
        ff.stft_adjoint_saved_device(p, 10000, z, z, z, None, out=torch.empty_like(d))


def test_config4_saved_adjoint_full_size(ff):
    """The config-4 shape (32 signals x 2^18, Hann 2048, hop 512, polar), magnitude-loss
    gradient: the saved-plane adjoint against the FP64 adjoint on three signals and the
    recomputing adjoint on all."""
    n, W, H, B, length = 2048, 2048, 512, 32, 1 << 18
    p = make(ff, n, W, H, "polar")
    x = signals(B, length, seed=1000)
    y = signals(B, length, seed=2000)
    d = torch.from_numpy(x).cuda()
    frames = ff.frame_count(p, length, 0)
    assert frames == 509
    re, im = forward(ff, p, d, frames, n)
    yre, _ = forward(ff, p, torch.from_numpy(y).cuda(), frames, n)
    gm = 2.0 * (re - yre)
    g = ff.stft_adjoint_saved_device(p, length, re, im, gm, None)
    rec = ff.stft_adjoint_device(p, d, gm, None)
    torch.cuda.synchronize()
    got, got_rec, gmn = g.cpu().numpy(), rec.cpu().numpy(), gm.cpu().numpy()
    assert fp64.rel_l2(got, got_rec) <= 4 * fp64.tolerance(n)
    for s in (0, 13, 31):
        want = fp64.stft_adjoint64(x[s], n, W, H, frames, win_n(n, W), "magnitude", gmn[s])
        err = fp64.rel_l2(got[s], want)
        print(f"config4 saved adjoint signal {s}: rel_l2 = {err:.3e}", flush=True)
        assert err <= 4 * fp64.tolerance(n)

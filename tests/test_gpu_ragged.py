"""Ragged batched STFT (fft_ragged_*): one persistent launch in a graph over requests
with different lengths, windows and hops, against per-request plain launches and the
FP64 oracle. Tolerance: relative L2 <= 2e-7*log2(N) (fp64.tolerance); frame counts exact."""
import numpy as np
import pytest

from fftfree_amd import CffiFft
from fftfree_amd.cffi_fft import RaggedBatch
from fftfree_amd.workloads import noise_plus_sines, ragged_requests
from oracle import fp64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


def make_batch(ff, count, seed, lo=12, hi=16, **kw):
    import torch
    reqs = ragged_requests(count, seed=seed, lo=lo, hi=hi)
    pcms = [torch.from_numpy(noise_plus_sines(L, seed=seed * 1000 + i)).cuda() for i, (L, W, H) in enumerate(reqs)]
    return reqs, pcms, RaggedBatch(ff, [(p, W, H) for p, (L, W, H) in zip(pcms, reqs)], **kw)


def oracle_planes(pcm, W, H, frames, mode):
    win = fp64.build_window(1, W)
    flags = {"polar": (False, True), "magnitude": (True, False), "complex": (False, False)}[mode]
    return fp64.stft_planes(pcm, W, W, H, frames, win, False, 1, flags[0], flags[1], True)


def check_planes(outs, want, mode, n):
    re, im, mg = (o.cpu().numpy() for o in outs)
    ore, oim, omg = want
    tol = fp64.tolerance(n)
    if mode == "polar":
        assert fp64.rel_l2(re * np.exp(1j * im), ore * np.exp(1j * oim)) <= tol
    else:
        assert fp64.rel_l2(re + 1j * im, ore + 1j * oim) <= tol
    assert fp64.rel_l2(mg, omg) <= tol


@pytest.mark.parametrize("mode", ["polar", "magnitude", "complex"])
def test_ragged_graph_matches_oracle_and_plain(ff, mode):
    import torch
    kw = dict(store_polar=mode == "polar", reduce_magnitude=mode == "magnitude")
    reqs, pcms, rb = make_batch(ff, 40, seed=3, **kw)
    print(f"ragged mode={mode}: {rb.describe()}", flush=True)
    for (L, W, H), f in zip(reqs, rb.frames):
        assert f == fp64.frame_count(L, W, W, H, 0, 1)
    rb.execute()
    torch.cuda.synchronize()
    got = [[o.clone() for o in outs] for outs in rb.outputs]
    for i in range(0, 40, 7):
        L, W, H = reqs[i]
        check_planes(got[i], oracle_planes(pcms[i].cpu().numpy(), W, H, rb.frames[i], mode), mode, W)
    rb.execute_plain()
    torch.cuda.synchronize()
    for i in range(40):
        a = [o.cpu().numpy() for o in got[i]]
        b = [o.cpu().numpy() for o in rb.outputs[i]]
        tol = fp64.tolerance(reqs[i][1])
        if mode == "polar":  # the phase of near-zero bins is noise: compare values
            assert fp64.rel_l2(a[0] * np.exp(1j * a[1]), b[0] * np.exp(1j * b[1])) <= tol
        else:
            assert fp64.rel_l2(a[0] + 1j * a[1], b[0] + 1j * b[1]) <= tol
        assert fp64.rel_l2(a[2], b[2]) <= tol
    rb.close()


def test_ragged_autotune_and_graph_replay(ff):
    import torch
    reqs, pcms, rb = make_batch(ff, 24, seed=4)
    rb.autotune(2)
    print(f"tuned: {rb.describe()}", flush=True)
    rb.execute()
    torch.cuda.synchronize()
    first = [[o.clone() for o in outs] for outs in rb.outputs]
    rb.execute()
    torch.cuda.synchronize()
    for a, b in zip(first, rb.outputs):
        for x, y in zip(a, b):
            assert torch.equal(x, y)  # replays are deterministic
    # the graph reads the live buffers: halve every input and replay
    for p in pcms:
        p.mul_(0.5)
    rb.execute()
    torch.cuda.synchronize()
    for a, b in zip(first, rb.outputs):
        assert fp64.rel_l2(b[0].cpu().numpy(), 0.5 * a[0].cpu().numpy()) <= 1e-6  # |X| halves
    for i in range(0, 24, 5):
        L, W, H = reqs[i]
        check_planes(rb.outputs[i], oracle_planes(pcms[i].cpu().numpy(), W, H, rb.frames[i], "polar"), "polar", W)
    rb.close()


def test_ragged_config5_batch(ff):
    """512 requests, lengths log-uniform in [2^12, 2^20], window from {256..4096}, hop
    W/4, R2C half spectrum, polar, pad_mode 1: exact frame counts for all, and every frame
    of every request against the FP64 oracle (framed as the reference frames: pcm * win
    rounded to float32, zero padding past the end)."""
    import torch
    reqs, pcms, rb = make_batch(ff, 512, seed=5, lo=12, hi=20)
    rb.execute()
    torch.cuda.synchronize()
    for (L, W, H), f in zip(reqs, rb.frames):
        assert f == fp64.frame_count(L, W, W, H, 0, 1)
    worst = 0.0
    for i, (L, W, H) in enumerate(reqs):
        f = rb.frames[i]
        x = fp64.frame_signal(pcms[i].cpu().numpy(), W, W, H, f, fp64.build_window(1, W))
        X = np.fft.rfft(x.astype(np.float64), axis=-1)
        X[:, 0] = X[:, 0].real
        X[:, -1] = X[:, -1].real
        mag = rb.outputs[i][0].cpu().numpy()
        ph = rb.outputs[i][1].cpu().numpy()
        err = fp64.rel_l2(mag * np.exp(1j * ph), X)
        worst = max(worst, err / fp64.tolerance(W))
        assert err <= fp64.tolerance(W), (i, L, W, err)
        assert np.array_equal(rb.outputs[i][2].cpu().numpy().view(np.uint32), mag.view(np.uint32))  # mag = |X| plane
    print(f"ragged 512 requests: worst rel_l2 / tolerance = {worst:.3f}", flush=True)
    rb.close()


def test_ragged_every_engine_matches_oracle(ff):
    """Every engine of every window size (tile shapes of each radix pattern, and the
    line-aligned engine at N = 256 .. 4096, N >= 2048 a block per workgroup), forced
    through fft_ragged_set_engine at 1 and 3 blocks per work item, on a batch whose
    requests end in padded frames: every frame of a sample of requests against the FP64
    oracle."""
    import torch
    reqs, pcms, rb = make_batch(ff, 30, seed=8, lo=12, hi=15)
    sizes = sorted({W for L, W, H in reqs})
    count = ff.lib.fft_ragged_engine_count()
    ran = {}
    for n in sizes:
        for engine in range(count):
            for per in (1, 3):
                if not rb.set_engine(n, engine, per):
                    continue
                ran.setdefault(n, set()).add(engine)
                rb.execute()
                torch.cuda.synchronize()
                for i, (L, W, H) in enumerate(reqs):
                    if W != n:
                        continue
                    check_planes(rb.outputs[i], oracle_planes(pcms[i].cpu().numpy(), W, H, rb.frames[i], "polar"),
                                 "polar", W)
    print(f"engines run per N: {ran}", flush=True)
    for n in sizes:
        assert len(ran.get(n, ())) >= 3, (n, ran.get(n))
    rb.close()


def test_ragged_misaligned_request_takes_the_tile_engine(ff):
    """A request whose samples start one float past an 8-byte boundary does not qualify
    for the line-aligned engine: its window size runs a tile shape, and the results
    still match the oracle."""
    import torch
    W, H = 1024, 256
    big = torch.from_numpy(noise_plus_sines(60001, seed=9)).cuda()
    pcms = [big[1:], torch.from_numpy(noise_plus_sines(40000, seed=10)).cuda()]
    rb = RaggedBatch(ff, [(p, W, H) for p in pcms])
    assert "lines" not in rb.describe()
    accepted = []
    for e in range(ff.lib.fft_ragged_engine_count()):
        if rb.set_engine(W, e, 2):  # the line-aligned engine of N = 1024 is refused
            accepted.append(e)
            assert "lines" not in rb.describe()
    assert len(accepted) == 2  # the two tile shapes of N = 1024
    rb.execute()
    torch.cuda.synchronize()
    for i, p in enumerate(pcms):
        check_planes(rb.outputs[i], oracle_planes(p.cpu().numpy(), W, H, rb.frames[i], "polar"), "polar", W)
    rb.close()

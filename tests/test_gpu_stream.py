"""Streaming STFT (fft_stream_push_pcm / fft_stream_push_spectrum and the pending /
backlog queries, fft_cffi.cpp:1342-1742) against the restated state machine
(oracle/stream.py over the float restatement of the batched paths).

Counts (frames produced, samples emitted, backlog, pending) must be equal; values within
2 * fp64.tolerance(N) in relative L2 per call. With COLA normalisation, samples whose
weight sum is below 1e-3 of its maximum are left out of the value comparison (dividing
rounding noise by ~1e-10 at a Hann edge), as in tests/test_gpu_istft.py."""
import ctypes

import numpy as np
import pytest

from fftfree_amd import CffiFft
from oracle import fp64
from oracle.stream import FLUSH_FINAL, FLUSH_NONE, Stream

pytestmark = pytest.mark.gpu
R2C, C2R = 1, 2
fp = lambda a: None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


def close(got, exp, N):
    got, exp = np.asarray(got, np.float64).ravel(), np.asarray(exp, np.float64).ravel()
    assert got.shape == exp.shape
    if exp.size:
        assert fp64.rel_l2(got, exp) <= 2 * fp64.tolerance(N)


@pytest.mark.parametrize("geom", [(1024, 1024, 256), (512, 384, 128), (2048, 2048, 512), (256, 256, 100)])
@pytest.mark.parametrize("windows", [False, True])
def test_stream_push_pcm_vs_oracle(ff, geom, windows):
    N, W, H = geom
    print(f"stream pcm N={N} W={W} H={H} windows={windows}", flush=True)
    p = ff.make_plan(N, window=W, hop=H, kernel=2, transform=R2C, half_spectrum=True, apply_windows=windows,
                     analysis_window=fp64.WINDOW_HANN)
    taps = np.concatenate([fp64.build_window(fp64.WINDOW_HANN, W), np.ones(N - W, np.float32)]) if windows else None
    st = Stream(N, window=W, hop=H, transform=R2C, taps=taps)
    lib = ff.lib
    lib.fft_stream_reset(p.handle)
    rng = np.random.default_rng(N + H)
    bins = N // 2 + 1
    for call in range(25):
        k = int(rng.integers(0, 3 * W))
        x = rng.standard_normal(k).astype(np.float32)
        maxf = int(rng.integers(0, 4)) if call % 3 == 0 else 0
        cap = max(1, (k + st.backlog()) // H + 2)
        planes = [np.full(cap * bins, np.nan, np.float32) for _ in range(3)]
        got = lib.fft_stream_push_pcm(p.handle, fp(x), k, fp(planes[0]), fp(planes[1]),
                                      None if call % 4 == 1 else fp(planes[2]), maxf, FLUSH_NONE)
        n, exp = st.push_pcm(x, max_frames=maxf)
        assert got == n
        assert lib.fft_stream_backlog_samples(p.handle) == st.backlog()
        assert lib.fft_stream_pending_frames(p.handle) == st.pending_frames()
        if n:
            close(planes[0][:n * bins], exp[0], N)
            close(planes[1][:n * bins], exp[1], N)
            if call % 4 != 1:
                close(planes[2][:n * bins], exp[2], N)
    assert lib.fft_stream_pending_pcm(p.handle) == 0  # forward context


@pytest.mark.parametrize("cola", [0, 1])
@pytest.mark.parametrize("geom", [(1024, 256), (512, 128), (256, 256)])
def test_stream_push_spectrum_vs_oracle(ff, geom, cola):
    N, H = geom
    print(f"stream spectrum N={N} H={H} cola={cola}", flush=True)
    p = ff.make_plan(N, window=N, hop=H, inverse=1, kernel=2, transform=C2R, half_spectrum=True,
                     apply_windows=True, analysis_window=fp64.WINDOW_HANN, synthesis_window=fp64.WINDOW_HANN,
                     cola=cola)
    w = fp64.build_window(fp64.WINDOW_HANN, N)
    st = Stream(N, window=N, hop=H, inverse=True, transform=C2R, analysis=w, synthesis=w, windows_enabled=True,
                cola_mode=cola)
    lib = ff.lib
    lib.fft_stream_reset(p.handle)
    rng = np.random.default_rng(N + cola)
    bins = N // 2 + 1
    full = float(np.max(w.astype(np.float64) ** 2))  # scale of a COLA weight sum
    for call in range(20):
        k = int(rng.integers(0, 6))
        re = rng.standard_normal((k, bins)).astype(np.float32)
        im = rng.standard_normal((k, bins)).astype(np.float32)
        flush = FLUSH_FINAL if call in (9, 19) else FLUSH_NONE
        ms = int(rng.integers(1, 3 * H)) if call % 4 == 2 else 0
        # room for everything that may come out: what is queued plus this call's frames
        # (a full flush adds N more)
        out = np.full(st.pending_pcm() + (k + 2) * N, np.nan, np.float32)
        got = lib.fft_stream_push_spectrum(p.handle, fp(re) if k else None, fp(im) if k else None, k, fp(out), ms,
                                           flush)
        exp = st.push_spectrum(re if k else None, im if k else None, max_samples=ms, flush=flush)
        assert got == exp.size
        assert lib.fft_stream_pending_pcm(p.handle) == st.pending_pcm()
        if got:
            g, e = out[:got].astype(np.float64), exp.astype(np.float64)
            assert np.isfinite(g).all()
            m = st.last_den > 1e-3 * full if cola else np.ones(got, bool)
            if m.any():
                assert fp64.rel_l2(g[m], e[m]) <= 2 * fp64.tolerance(N)
    # a forward context refuses the inverse stream
    fwd = ff.make_plan(N, window=N, hop=H, transform=R2C, half_spectrum=True)
    out = np.zeros(N, np.float32)
    re = np.zeros((1, bins), np.float32)
    assert lib.fft_stream_push_spectrum(fwd.handle, fp(re), fp(re), 1, fp(out), 0, FLUSH_NONE) == 0


def test_stream_inverse_matches_batched_overlap_add(ff):
    """The library's own two paths: frames streamed then flushed equal the fused batched
    overlap-add (fft_execute_complex_batched with apply_ola and COLA) up to rounding, and
    the samples past it are the flush's zeros."""
    N, H, F = 1024, 256, 77
    kw = dict(window=N, hop=H, inverse=1, kernel=2, transform=C2R, half_spectrum=True, apply_windows=True,
              analysis_window=fp64.WINDOW_HANN, synthesis_window=fp64.WINDOW_HANN, cola=1)
    ps = ff.make_plan(N, **kw)
    pb = ff.make_plan(N, apply_ola=True, **kw)
    rng = np.random.default_rng(8)
    re = rng.standard_normal((F, N // 2 + 1)).astype(np.float32)
    im = rng.standard_normal((F, N // 2 + 1)).astype(np.float32)
    lib = ff.lib
    lib.fft_stream_reset(ps.handle)
    outs, f = [], 0
    while f < F:
        k = int(rng.integers(1, 9))
        buf = np.zeros(k * N + N, np.float32)
        n = lib.fft_stream_push_spectrum(ps.handle, fp(np.ascontiguousarray(re[f:f + k])),
                                         fp(np.ascontiguousarray(im[f:f + k])), min(k, F - f), fp(buf), 0, FLUSH_NONE)
        outs.append(buf[:n].copy())
        f += k
    buf = np.zeros(2 * N, np.float32)
    n = lib.fft_stream_push_spectrum(ps.handle, None, None, 0, fp(buf), 0, FLUSH_FINAL)
    outs.append(buf[:n].copy())
    got = np.concatenate(outs)
    L = (F - 1) * H + N
    batch = np.zeros(F * N, np.float32)
    assert lib.fft_execute_complex_batched(pb.handle, fp(re), fp(im), F, fp(batch), 1, 0, 0) == F
    w = fp64.build_window(fp64.WINDOW_HANN, N).astype(np.float64)
    norm = fp64.cola_norm64(F, N, H, w * w)[:L]
    m = norm > 1e-3 * norm.max()
    assert got.size == L + H and not got[L:].any()
    assert fp64.rel_l2(got[:L][m], batch[:L][m]) <= 2 * fp64.tolerance(N)


@pytest.mark.parametrize("hop", [256, 384, 1024])
def test_stream_spectrum_reference_additions(ff, hop):
    """The spectrum push's overlap-add is the reference's accumulator loop
    (fft_cffi.cpp:1602-1742) bit for bit: per frame, add the frame's N samples to the
    accumulator, queue the first min(hop, N) (all N after the last frame of a final
    flush), shift. Checked on the library's own frames (fft_execute_complex_batched of
    the same context), with the loop run in float32 here; pushes of 0 .. 16 frames, the
    accumulator carried across them."""
    N = 1024
    p = ff.make_plan(N, window=N, hop=hop, inverse=1, kernel=2, transform=C2R, half_spectrum=True, apply_windows=True,
                     analysis_window=fp64.WINDOW_HANN, synthesis_window=fp64.WINDOW_HANN, cola=0)
    lib = ff.lib
    lib.fft_stream_reset(p.handle)
    bins = N // 2 + 1
    e = min(hop, N)
    rng = np.random.default_rng(hop)
    acc = np.zeros(N, np.float32)
    for call, k in enumerate([1, 3, 0, 16, 5, 2, 7]):
        final = call == 6
        re = rng.standard_normal((k, bins)).astype(np.float32)
        im = rng.standard_normal((k, bins)).astype(np.float32)
        want = []
        if k:
            fr = np.empty(k * N, np.float32)
            assert lib.fft_execute_complex_batched(p.handle, fp(re), fp(im), k, fp(fr), 2, 0, k) == k
            for f in range(k):
                acc += fr[f * N:(f + 1) * N]
                emit = N if (final and f == k - 1) else e
                want.append(acc[:emit].copy())
                acc = np.concatenate([acc[emit:], np.zeros(emit, np.float32)])
        want = np.concatenate(want) if want else np.zeros(0, np.float32)
        out = np.full(k * N + N, np.nan, np.float32)
        got = lib.fft_stream_push_spectrum(p.handle, fp(re) if k else None, fp(im) if k else None, k, fp(out), 0,
                                           FLUSH_FINAL if final else FLUSH_NONE)
        assert got == want.size, call
        assert np.array_equal(out[:got], want), call

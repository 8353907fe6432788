"""Pins of the oracle itself (CPU only).

Each check is chosen so a plausible mistake in the oracle fails it:
* the FP64 FFT is checked against a direct O(N^2) DFT (no shared code path);
* the float restatement of the reference (cpu_ref) is checked against the FP64 oracle
  for both kernels and both directions, and on the reference's own test shapes;
* the integer metadata closed forms are checked against the maps cpu_ref records
  inside its butterfly loops, against the N=8 values written out by hand, and by
  executing the butterfly network they describe, which must reproduce the DFT;
* windows are compared bit-for-bit between the two restatements and against scipy;
* the frame count is compared with a brute-force enumeration.
"""
import math

import numpy as np
import pytest
import scipy.signal

from oracle import fp64, index_maps as M, ref

POW2 = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]


def cnormal(b, n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((b, n)) + 1j * rng.standard_normal((b, n))).astype(np.complex64)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 64, 256, 1024])
@pytest.mark.parametrize("inverse", [False, True])
def test_fp64_fft_matches_direct_dft(n, inverse):
    x = cnormal(3, n, 7).astype(np.complex128)
    a = fp64.fft64(x, inverse)
    b = fp64.dft_direct(x, inverse)
    assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(b)))


def test_fp64_closed_forms():
    # impulse -> constant, constant -> impulse, single tone -> one bin
    n = 64
    e = np.zeros(n); e[0] = 1
    assert np.allclose(fp64.fft64(e), np.ones(n))
    assert np.allclose(fp64.fft64(np.ones(n)), n * e)
    k0 = 5
    tone = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    X = fp64.fft64(tone)
    assert abs(X[k0] - n) < 1e-9 and np.max(np.abs(np.delete(X, k0))) < 1e-9
    # Parseval
    x = cnormal(1, n, 3)[0].astype(np.complex128)
    assert math.isclose(np.sum(np.abs(fp64.fft64(x)) ** 2), n * np.sum(np.abs(x) ** 2), rel_tol=1e-12)


@pytest.mark.parametrize("n", POW2)
@pytest.mark.parametrize("kernel", [0, 1, 2])
@pytest.mark.parametrize("inverse", [False, True])
def test_cpu_ref_fft_within_tolerance(n, kernel, inverse):
    x = cnormal(5, n, 11 + n)
    y = ref.fft_batched(x, inverse=inverse, kernel=kernel)
    assert fp64.rel_l2(y, fp64.fft64(x, inverse)) <= fp64.tolerance(n)


def test_cpu_ref_kernel_codes_ct_vs_stockham_equal_up_to_rounding():
    # C-level codes 0 and 1 both run Cooley-Tukey (fft_cffi.cpp:799-800): identical bits.
    x = cnormal(4, 256, 5)
    assert np.array_equal(ref.fft_batched(x, kernel=0), ref.fft_batched(x, kernel=1))
    assert fp64.rel_l2(ref.fft_batched(x, kernel=2), ref.fft_batched(x, kernel=0)) < 1e-6


def test_reference_roundtrip_small_batches():
    # fftfree_tests.cpp:177-220: N=16, B=8 forward then inverse, tolerance 1e-5.
    x = cnormal(8, 16, 42)
    y = ref.fft_batched(ref.fft_batched(x), inverse=True)
    assert np.max(np.abs(y - x)) < 1e-5


def test_reference_stockham_parallel_large_batch_real_inputs():
    # fftfree_tests.cpp:719-844: N=256, B=128, real-only inputs, CT and Stockham agree
    # with an exact DFT within 1e-5 * 256.
    rng = np.random.default_rng(1337)
    x = rng.standard_normal((128, 256)).astype(np.float32).astype(np.complex64)
    want = fp64.fft64(x)
    for k in (1, 2):
        assert np.max(np.abs(ref.fft_batched(x, kernel=k) - want)) < 1e-5 * 256


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32])
@pytest.mark.parametrize("frames", [1, 3, 16])
def test_reference_extended_consensus(n, frames):
    # fftfree_tests.cpp:1089-1271: CT vs Stockham vs exact DFT, seed N*1009+frames*97.
    x = cnormal(frames, n, n * 1009 + frames * 97)
    want = fp64.fft64(x)
    for k in (0, 2):
        assert fp64.rel_l2(ref.fft_batched(x, kernel=k), want) <= fp64.tolerance(n)


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 256, 1024])
@pytest.mark.parametrize("kernel", [0, 2])
def test_index_maps_match_recorded_metadata(n, kernel):
    S = M.log2_exact(n)
    ti = ref.metadata(n, M.KIND_TW_INDEX, kernel).real.astype(np.int64).reshape(S, n)
    assert np.array_equal(ti, M.twiddle_index_map(n, kernel))
    bp = ref.metadata(n, M.KIND_PAIRS, kernel)
    bp = np.stack([bp.real, bp.imag], -1).astype(np.int64).reshape(S, n, 2)
    assert np.array_equal(bp, M.butterfly_pairs(n, kernel))
    sp = ref.metadata(n, M.KIND_STAGE_PARAMS, kernel)
    assert np.array_equal(np.stack([sp.real, sp.imag], -1).astype(np.int64), M.stage_params(n))
    lm = ref.metadata(n, M.KIND_LAYOUT, kernel)
    assert np.array_equal(lm.real.astype(np.int64), M.layout_map(n)) and not lm.imag.any()
    tw = ref.metadata(n, M.KIND_TWIDDLE, kernel)
    plan_w = ref.twiddles(n)
    assert np.array_equal(tw, plan_w[M.twiddle_metadata_index(n)])
    assert np.array_equal(ref.bitrev(n), M.bitrev(n))


def test_index_maps_known_values_n8():
    # written out by hand from the loops at eigen_fft.hpp:2981-3011 and 3478-3505
    assert M.twiddle_index_map(8, 2).tolist() == [[0] * 8, [0, 0, 2, 2, 0, 0, 2, 2], [0, 1, 2, 3, 0, 1, 2, 3]]
    assert M.twiddle_index_map(8, 0).tolist() == [[0] * 8, [0, 2, 0, 2, 0, 2, 0, 2], [0, 1, 2, 3, 0, 1, 2, 3]]
    assert M.bitrev(8).tolist() == [0, 4, 2, 6, 1, 5, 3, 7]
    assert M.butterfly_pairs(8, 0)[0].tolist() == [[0, 1], [0, 1], [2, 3], [2, 3], [4, 5], [4, 5], [6, 7], [6, 7]]
    assert M.butterfly_pairs(8, 2)[0].tolist() == [[0, 4], [1, 5], [2, 6], [3, 7], [0, 4], [1, 5], [2, 6], [3, 7]]
    assert M.stage_params(8).tolist() == [[2, 4], [4, 2], [8, 1]]


@pytest.mark.parametrize("n", [2, 4, 8, 32, 256, 1024])
@pytest.mark.parametrize("kernel", [0, 2])
@pytest.mark.parametrize("inverse", [False, True])
def test_index_maps_describe_the_dft(n, kernel, inverse):
    x = cnormal(2, n, 99).astype(np.complex128)
    y = M.run_butterfly_program(x, kernel, M.butterfly_pairs(n, kernel), M.twiddle_index_map(n, kernel), inverse)
    assert np.max(np.abs(y - fp64.fft64(x, inverse))) < 1e-9 * n


def test_index_map_check_has_teeth():
    # a single wrong twiddle index or pair must break the network (plant in the oracle)
    n, k = 64, 2
    x = cnormal(1, n, 5).astype(np.complex128)
    tw = M.twiddle_index_map(n, k).copy()
    tw[3, 17] = (tw[3, 17] + 1) % (n // 2)
    y = M.run_butterfly_program(x, k, M.butterfly_pairs(n, k), tw)
    assert np.max(np.abs(y - fp64.fft64(x))) > 1e-3
    pairs = M.butterfly_pairs(n, 0).copy()
    pairs[2, 5, 1] += 1
    y = M.run_butterfly_program(x, 0, pairs, M.twiddle_index_map(n, 0))
    assert np.max(np.abs(y - fp64.fft64(x))) > 1e-3


def test_omega_float_formula():
    n = 1024
    om = M.omega(n)
    assert om.dtype == np.float32 and om[0] == 0.0
    assert np.max(np.abs(om.astype(np.float64) - 2 * np.pi * np.arange(n) / n)) < 1e-5
    assert len(M.omega(n, transform=1, half_spectrum=True)) == n // 2 + 1


@pytest.mark.parametrize("kind", range(6))
@pytest.mark.parametrize("norm", range(3))
@pytest.mark.parametrize("W", [2, 7, 256, 1024])
def test_windows_two_restatements_bit_exact(kind, norm, W):
    a = fp64.build_window(kind, W, 0.25 if kind == 4 else 0.0, 0.0, norm)
    b = ref.build_window(kind, W, 0.25 if kind == 4 else 0.0, 0.0, norm)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("W", [16, 513, 1024, 4096])
def test_windows_against_scipy(W):
    # symmetric Hann/Hamming/Blackman; the reference's Kaiser is a Hann placeholder
    for kind, sp in [(1, scipy.signal.windows.hann(W, sym=True)), (2, scipy.signal.windows.hamming(W, sym=True)),
                     (3, scipy.signal.windows.blackman(W, sym=True)), (5, scipy.signal.windows.hann(W, sym=True))]:
        w = fp64.build_window(kind, W)
        assert np.max(np.abs(w.astype(np.float64) - sp)) < 1e-6
    tk = fp64.build_window(4, W, 0.5)
    assert np.max(np.abs(tk.astype(np.float64) - scipy.signal.windows.tukey(W, 0.5, sym=True))) < 1e-6
    l2 = fp64.build_window(1, W, norm=1)
    assert abs(float(np.sum(l2.astype(np.float64) ** 2)) - 1.0) < 1e-5
    area = fp64.build_window(1, W, norm=2)
    assert abs(float(np.sum(area.astype(np.float64))) - 1.0) < 1e-5


def brute_frames(L, N, W, H, pad, max_frames):
    W = W or N
    H = H or W
    if W > N or L == 0:
        return 0
    if L < W:
        f = 1 if pad == 1 else 0
    else:
        f = sum(1 for s in range(0, L, H) if s + W <= L)
    if max_frames and f > max_frames:
        f = max_frames
    return f


def test_frame_count_rule():
    cases = 0
    for L in [0, 1, 5, 255, 256, 257, 1000, 1024, 1025, 4096, 5000]:
        for N, W, H in [(256, 0, 0), (256, 256, 64), (512, 256, 100), (256, 512, 64), (1024, 1024, 256), (8, 3, 5)]:
            for pad in (0, 1, 2):
                for mx in (0, 1, 3):
                    want = brute_frames(L, N, W, H, pad, mx)
                    assert fp64.frame_count(L, N, W, H, 0, pad, mx) == want
                    assert ref.frame_count(L, N, W, H, pad, 0, mx) == want  # pad via the context
                    cases += 1
    assert cases > 500
    # headline workload shapes
    assert fp64.frame_count(1 << 24, 1024, 1024, 256, 0, 0) == 65533
    assert fp64.frame_count(1 << 18, 2048, 2048, 512, 0, 0) == 509


@pytest.mark.parametrize("transform", [0, 1, 2, 3])
@pytest.mark.parametrize("flags", [(0, 0, 0), (1, 0, 1), (0, 1, 1), (1, 1, 0), (0, 0, 1)])
@pytest.mark.parametrize("kernel", [0, 2])
@pytest.mark.parametrize("inverse", [False, True])
def test_cpu_ref_stft_planes_vs_fp64(transform, flags, kernel, inverse):
    reduce_mag, polar, half = flags
    if transform == 2 and half and kernel != 2:
        pytest.skip("Cooley-Tukey C2R half rebuilds the input spectrum; covered separately")
    N, W, H = 64, 48, 16
    rng = np.random.default_rng(3)
    pcm = rng.standard_normal(500).astype(np.float32)
    win = fp64.build_window(1, W)
    taps = np.ones(N, np.float32); taps[:W] = win
    got, re, im, mg = ref.execute_batched(pcm, N, W, H, ctx_pad=1, pad_mode=0, inverse=inverse, kernel=kernel,
                                          transform=transform, reduce_magnitude=reduce_mag, store_polar=polar,
                                          half_spectrum=half, window_taps=taps)
    frames = fp64.frame_count(pcm.size, N, W, H, 1, 0)
    assert got == frames
    ore, oim, omg = fp64.stft_planes(pcm, N, W, H, frames, win, inverse, transform, reduce_mag, polar, half)
    tol = 4 * fp64.tolerance(N)
    if polar:
        z, oz = re * np.exp(1j * im), ore * np.exp(1j * oim)
        assert fp64.rel_l2(z, oz) <= tol
    else:
        assert fp64.rel_l2(re + 1j * im, ore + 1j * oim) <= tol
    assert fp64.rel_l2(mg, omg) <= tol


# -- the FP64 adjoint (gradient) oracle, pinned by finite differences ---------------------

def _adj_case(seed, L=300, N=64, W=64, H=16, win_kind=1):
    rng = np.random.default_rng(seed)
    pcm = rng.standard_normal(L).astype(np.float32)
    frames = fp64.frame_count(L, N, W, H, 0, 1)
    win = fp64.build_window(win_kind, W) if win_kind else None
    win_n = None if win is None else np.concatenate([win, np.ones(N - W, np.float32)])
    return rng, pcm, frames, win_n


def _fd_check(pcm, N, W, H, frames, win, mode, ga, gb, grad, coords, h=1e-6):
    def loss(x):
        A, B = fp64.value_planes64(x, N, W, H, frames, win, mode)
        return np.sum(ga * A) + np.sum(gb * B)
    x0 = pcm.astype(np.float64)
    worst = 0.0
    for j in coords:
        xp = x0.copy(); xp[j] += h
        xm = x0.copy(); xm[j] -= h
        fd = (loss(xp) - loss(xm)) / (2 * h)
        worst = max(worst, abs(fd - grad[j]) / (1e-6 + abs(fd)))
    return worst


@pytest.mark.parametrize("mode", ["complex", "polar", "magnitude"])
@pytest.mark.parametrize("geom", [(64, 64, 16), (64, 48, 12), (32, 32, 8)])
def test_adjoint_oracle_matches_finite_differences(mode, geom):
    N, W, H = geom
    rng, pcm, frames, win = _adj_case(7 + N + W, N=N, W=W, H=H)
    bins = N // 2 + 1
    ga = rng.standard_normal((frames, bins))
    gb = rng.standard_normal((frames, bins))
    g = fp64.stft_adjoint64(pcm, N, W, H, frames, win, mode, ga, gb)
    coords = rng.choice(pcm.size, 40, replace=False)
    assert _fd_check(pcm, N, W, H, frames, win, mode, ga, gb, g, coords) < 1e-5


def test_adjoint_oracle_fd_check_has_teeth():
    # plausible mistakes planted in the oracle: the window left out of the overlap-add,
    # and the complex-output V used for a polar loss; the FD check must reject both
    N, W, H = 64, 64, 16
    rng, pcm, frames, win = _adj_case(3)
    bins = N // 2 + 1
    ga = rng.standard_normal((frames, bins))
    gb = rng.standard_normal((frames, bins))
    coords = rng.choice(pcm.size, 20, replace=False)
    good = fp64.stft_adjoint64(pcm, N, W, H, frames, win, "polar", ga, gb)
    no_win = fp64.stft_adjoint64(pcm, N, W, H, frames, win, "polar", ga, gb, ola_window=False)
    wrong_v = fp64.stft_adjoint64(pcm, N, W, H, frames, win, "complex", ga, gb)
    assert _fd_check(pcm, N, W, H, frames, win, "polar", ga, gb, good, coords) < 1e-5
    assert _fd_check(pcm, N, W, H, frames, win, "polar", ga, gb, no_win, coords) > 1e-2
    assert _fd_check(pcm, N, W, H, frames, win, "polar", ga, gb, wrong_v, coords) > 1e-2


def test_adjoint_oracle_agrees_with_direct_magnitude_gradient():
    # the direct O(N^2) restatement (stft_magnitude_grad) and the irfft form agree
    N, W, H = 64, 64, 16
    rng, pcm, frames, win = _adj_case(11)
    y = np.abs(rng.standard_normal((frames, N // 2 + 1)))
    A, _ = fp64.value_planes64(pcm, N, W, H, frames, win, "magnitude")
    direct = fp64.stft_magnitude_grad(pcm, y, N, W, H, frames, win)
    fast = fp64.stft_adjoint64(pcm, N, W, H, frames, win, "magnitude", 2 * (A - y))
    assert np.max(np.abs(direct - fast)) <= 1e-9 * np.max(np.abs(direct))


# ---- fft_execute_complex_batched: rebuild, hooks, overlap-add, COLA -------------------

def _planes64(x, N, H, win):
    """R2C half-spectrum planes of W = N frames, all in float64 (no float32 rounding)."""
    frames = fp64.frame_count(x.size, N, N, H, 1, 1)
    fr = np.stack([np.asarray(x[f * H:f * H + N], np.float64) for f in range(frames)])
    X = np.fft.rfft(fr * np.asarray(win, np.float64), axis=-1)
    return frames, X.real.copy(), X.imag.copy()


@pytest.mark.parametrize("N", [16, 64, 256])
@pytest.mark.parametrize("hop_div", [1, 2, 4])
@pytest.mark.parametrize("kinds", [(fp64.WINDOW_HANN, fp64.WINDOW_HANN), (fp64.WINDOW_HAMMING, fp64.WINDOW_RECT),
                                   (fp64.WINDOW_RECT, fp64.WINDOW_BLACKMAN)])
def test_complex_oracle_reconstructs_the_signal(N, hop_div, kinds):
    """Analysis STFT then the inverse with overlap-add: each frame returns wa*x, so the
    output is x * sum(wa*ws) without COLA and x itself with it (where the sum is not
    tiny), and frame-major output is wa*ws*frame. Pins the Hermitian rebuild, the 1/N
    scale, the hop and the frame order of the sums."""
    H = N // hop_div
    wa = fp64.build_window(kinds[0], N)
    ws = fp64.build_window(kinds[1], N)
    x = np.random.default_rng(N + H).standard_normal(N * 6).astype(np.float64)
    frames, re, im = _planes64(x, N, H, wa)
    L = (frames - 1) * H + N
    norm = fp64.cola_norm64(frames, N, H, wa.astype(np.float64) * ws)
    kw = dict(windows_enabled=True, analysis=wa, synthesis=ws, hop=H)
    raw = fp64.complex_batched64(re, im, N, apply_ola=True, cola_mode=0, **kw)
    assert np.max(np.abs(raw[:L] - x[:L] * norm[:L])) <= 1e-10
    assert not raw[L:].any()
    cola = fp64.complex_batched64(re, im, N, apply_ola=True, cola_mode=1, **kw)
    m = norm[:L] > 1e-3 * norm.max()
    assert m.sum() > L // 2
    assert np.max(np.abs(cola[:L][m] - x[:L][m])) <= 1e-9
    fm = fp64.complex_batched64(re, im, N, apply_ola=False, **kw)
    framed = np.stack([x[f * H:f * H + N] for f in range(frames)])
    assert np.max(np.abs(fm - framed * wa * ws)) <= 1e-10


def test_complex_oracle_reconstruction_check_has_teeth():
    """Planted mistakes (a hop off by one, a mirror without the conjugate) must fail the
    reconstruction check above."""
    N, H = 64, 16
    w = fp64.build_window(fp64.WINDOW_HANN, N)
    x = np.random.default_rng(5).standard_normal(N * 6)
    frames, re, im = _planes64(x, N, H, w)
    norm = fp64.cola_norm64(frames, N, H, w.astype(np.float64) ** 2)
    L = (frames - 1) * H + N
    m = norm[:L] > 1e-3 * norm.max()
    bad_hop = fp64.complex_batched64(re, im, N, windows_enabled=True, analysis=w, synthesis=w, hop=H + 1,
                                     apply_ola=True, cola_mode=1)
    assert np.max(np.abs(bad_hop[:L][m] - x[:L][m])) > 1e-2
    cols = fp64.complex_columns64(re, im, N, fp64.C2R, True)
    cols[:, N // 2 + 1:] = cols[:, 1:N // 2][:, ::-1]  # mirror without conj
    y = fp64.fft64(cols, True).real
    assert np.max(np.abs(y[0, :N] - x[:N] * w)) > 1e-2


@pytest.mark.parametrize("N", [8, 32])
def test_complex_oracle_closed_forms(N):
    """One bin k0 set: a C2R half-spectrum plan gives (2/N) cos(2 pi k0 n/N + phi), a
    complex plan on the zero-filled half spectrum (1/N) e^{i...}; with store_polar the
    Cooley-Tukey pre hook reads (mag, phase) (eigen_fft.hpp:2511-2519), Stockham does not,
    and the polar post hook leaves |x| in the real part; DC and Nyquist imaginary parts
    are dropped by the rebuild."""
    k0, phi = 3, 0.7
    n = np.arange(N)
    th = 2 * np.pi * k0 * n / N
    re = np.zeros((1, N // 2 + 1))
    im = np.zeros((1, N // 2 + 1))
    re[0, k0], im[0, k0] = np.cos(phi), np.sin(phi)
    c2r = fp64.complex_batched64(re, im, N, transform=fp64.C2R)
    assert np.allclose(c2r[0], (2 / N) * np.cos(th + phi), atol=1e-14)
    c2c = fp64.complex_batched64(re, im, N, transform=fp64.C2C)
    assert np.allclose(c2c[0], (1 / N) * np.cos(th + phi), atol=1e-14)
    pre, pim = np.zeros_like(re), np.zeros_like(im)
    pre[0, k0], pim[0, k0] = 1.0, phi
    ct = fp64.complex_batched64(pre, pim, N, kernel=0, transform=fp64.C2R, store_polar=True)
    assert np.allclose(ct[0], np.abs((2 / N) * np.cos(th + phi)), atol=1e-14)
    st = fp64.complex_batched64(pre, pim, N, kernel=2, transform=fp64.C2R, store_polar=True)
    assert np.allclose(st[0], np.abs((2 / N) * (np.cos(th) - phi * np.sin(th))), atol=1e-14)
    dc = np.zeros_like(re)
    dci = np.zeros_like(im)
    dc[0, 0], dci[0, 0], dc[0, -1], dci[0, -1] = 1.0, 5.0, 1.0, -3.0
    out = fp64.complex_batched64(dc, dci, N, transform=fp64.C2R)
    assert np.allclose(out[0], (1 + (-1.0) ** n) / N, atol=1e-14)


@pytest.mark.parametrize("kernel", [0, 2])
@pytest.mark.parametrize("mode", [(0, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)])
def test_cpu_ref_complex_batched_vs_fp64(kernel, mode):
    """The float restatement (rebuild, plan, output stage) against the FP64 oracle:
    (apply_ola, cola, windows) combinations. Samples whose COLA weight sum is tiny are
    ill-conditioned (division by ~1e-10 at a Hann edge) whatever the precision, so the
    normalised comparison covers samples with a sum above 1e-3 of its maximum."""
    ola, cola, windows = mode
    N, H = 256, 64
    w = fp64.build_window(fp64.WINDOW_HANN, N)
    x = np.random.default_rng(9).standard_normal(N * 12).astype(np.float32)
    frames, re, im = _planes64(x, N, H, w)
    re, im = re.astype(np.float32), im.astype(np.float32)
    taps = w if windows else None
    got, out = ref.execute_complex_batched(re, im, N, kernel=kernel, synth_taps=taps, analysis_taps=taps,
                                           apply_ola=bool(ola), hop=H, cola_mode=cola)
    exp = fp64.complex_batched64(re, im, N, kernel=kernel, windows_enabled=bool(windows), analysis=taps,
                                 synthesis=taps, apply_ola=bool(ola), hop=H, cola_mode=cola)
    assert got == frames
    out, exp = np.asarray(out).ravel(), np.asarray(exp).ravel()
    m = np.ones(out.size, bool)
    if ola and cola:
        norm = fp64.cola_norm64(frames, N, H, (w.astype(np.float64) ** 2) if windows else np.ones(N))
        m = norm > 1e-3 * norm.max()
    assert fp64.rel_l2(out[m], exp[m]) <= fp64.tolerance(N)


# ---- Griffin-Lim ----------------------------------------------------------------------

def _gl_problem(N=256, H=64, periods=40, seed=3):
    from fftfree_amd.workloads import noise_plus_sines
    x = noise_plus_sines(N * periods, seed=seed)
    frames = fp64.frame_count(x.size, N, N, H, 1, 1)
    _, _, mag = fp64.stft_planes(x, N, N, H, frames, None, False, fp64.R2C, False, False, True)
    return mag.astype(np.float32), frames


def _gl_inconsistency(X, mag, N, H):
    """|| |STFT(averaged OLA of IDFT(X))| - mag || / ||mag|| (spectral convergence)."""
    frames = mag.shape[0]
    fr = np.fft.irfft(np.asarray(X, np.complex128), n=N, axis=-1)
    L = (frames - 1) * H + N
    ola, w = np.zeros(L), np.zeros(L)
    for f in range(frames):
        ola[f * H:f * H + N] += fr[f]
        w[f * H:f * H + N] += 1
    ola /= w
    Y = np.abs(np.fft.rfft(np.stack([ola[f * H:f * H + N] for f in range(frames)]), axis=-1))
    return np.linalg.norm(Y - mag) / np.linalg.norm(mag)


@pytest.mark.parametrize("seed", [0, 1, 12345])
def test_gl_start_angles_are_the_fused_uniform(seed):
    """The start angles of the reference (std::uniform_real_distribution<float> over
    std::mt19937, built with the reference's -ffast-math -march flags) equal
    fma(u, 2 pi, -pi) with u = float(word) / 2^32, on the raw mt19937 words (numpy's
    MT19937 with the same init_genrand seeding). The GPU start kernel computes that
    formula from the same words; an unfused multiply-add would differ in half the
    draws."""
    n = 50000
    a = ref.gl_angles(seed, n)
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed if seed else 0x47d7bfad)
    w = bg.random_raw(n).astype(np.uint64)
    u = (w.astype(np.float32) * np.float32(2.0 ** -32)).astype(np.float32)
    u = np.where(u >= 1, np.nextafter(np.float32(1), np.float32(0)), u)
    kpi = np.float32(np.pi)
    fused = (u.astype(np.float64) * np.float64(kpi - (-kpi)) + np.float64(-kpi)).astype(np.float32)
    assert np.array_equal(fused, a)
    unfused = (u * (kpi - (-kpi))).astype(np.float32) + (-kpi)
    assert np.sum(unfused != a) > n // 10


@pytest.mark.parametrize("iterations", [1, 8, 32])
def test_gl_two_restatements_agree(iterations):
    """The float restatement (cpu_ref, the reference loop over the restated inverse and
    forward paths) and the FP64 one (numpy) agree within 2*tol(N)*sqrt(iterations) from
    the same start angles."""
    N, H = 256, 64
    mag, frames = _gl_problem(N, H)
    ok, r, i = ref.griffin_lim(mag, N, H, iterations=iterations, seed=7)
    assert ok == 1
    X64 = fp64.griffin_lim64(mag, N, H, ref.gl_angles(7, mag.size).reshape(mag.shape), iterations)
    assert fp64.rel_l2(r + 1j * i, X64) <= 2 * fp64.tolerance(N) * max(1.0, np.sqrt(iterations))


def test_gl_agreement_check_has_teeth():
    """A planted mistake in the FP64 restatement (the overlap-add hop off by one) breaks
    the agreement."""
    N, H = 256, 64
    mag, _ = _gl_problem(N, H)
    ok, r, i = ref.griffin_lim(mag, N, H, iterations=4, seed=7)
    ang = ref.gl_angles(7, mag.size).reshape(mag.shape)
    bad = fp64.griffin_lim64(mag, N, H + 1, ang, 4, fwd_hop=H)
    assert fp64.rel_l2(r + 1j * i, bad) > 1e-2


def test_gl_oracle_properties():
    """|X| stays the given magnitude, bins 0 and N/2 are real, and the spectral
    convergence of the iterates does not increase (Griffin-Lim's monotonicity)."""
    N, H = 256, 64
    mag, _ = _gl_problem(N, H)
    prev = np.inf
    for it in [1, 4, 16, 64]:
        ok, r, i = ref.griffin_lim(mag, N, H, iterations=it, seed=11)
        assert ok == 1
        assert np.max(np.abs(np.hypot(r, i) - mag)) <= 1e-6 * np.max(mag)
        assert not i[:, 0].any() and not i[:, -1].any()
        sc = _gl_inconsistency(r + 1j * i, mag, N, H)
        assert sc < prev
        prev = sc
    assert prev < 0.2


# ---- streaming ------------------------------------------------------------------------

def test_stream_oracle_forward_equals_batch():
    """Pushing a signal in uneven chunks yields, in order, exactly the frames the batched
    call makes of the whole signal (partial frames dropped): the buffering neither
    skips nor repeats samples."""
    from oracle.stream import Stream
    N, H = 256, 64
    x = np.random.default_rng(1).standard_normal(N * 30).astype(np.float32)
    st = Stream(N, window=N, hop=H, transform=fp64.R2C)
    got_re = []
    pos, rng = 0, np.random.default_rng(2)
    while pos < x.size:
        k = int(rng.integers(1, 700))
        n, planes = st.push_pcm(x[pos:pos + k])
        pos += k
        if n:
            got_re.append(planes[0])
    got = np.concatenate(got_re)
    full, re, _, _ = ref.execute_batched(x, N, N, H, ctx_pad=0, pad_mode=2, transform=fp64.R2C, half_spectrum=True)
    assert got.shape == re.shape and np.array_equal(got, re)
    assert st.backlog() == x.size - full * H


@pytest.mark.parametrize("cola", [0, 1])
def test_stream_oracle_inverse_equals_batch_overlap_add(cola):
    """Frames pushed in uneven groups, then a final flush: the samples that come out are
    the batched overlap-add (with the same COLA division) over (F-1)*H + N samples,
    followed by the H zeros the flush of the whole N-sample accumulator adds
    (fft_cffi.cpp:1717-1736)."""
    from oracle.stream import Stream, FLUSH_FINAL
    N, H, F = 256, 64, 37
    w = fp64.build_window(fp64.WINDOW_HANN, N)
    rng = np.random.default_rng(3)
    re = rng.standard_normal((F, N // 2 + 1)).astype(np.float32)
    im = rng.standard_normal((F, N // 2 + 1)).astype(np.float32)
    st = Stream(N, window=N, hop=H, inverse=True, transform=fp64.C2R, analysis=w, synthesis=w,
                windows_enabled=True, cola_mode=cola)
    outs, f = [], 0
    while f < F:
        k = int(rng.integers(1, 6))
        outs.append(st.push_spectrum(re[f:f + k], im[f:f + k]))
        f += k
    outs.append(st.push_spectrum(None, None, flush=FLUSH_FINAL))
    got = np.concatenate(outs)
    L = (F - 1) * H + N
    _, exp = ref.execute_complex_batched(re, im, N, synth_taps=w, analysis_taps=w, apply_ola=True, hop=H,
                                         cola_mode=cola)
    assert got.size == L + H
    # the batched path fuses out += x*w into one rounding under the reference's flags, the
    # streaming path rounds the windowed frame first: a few ulps apart
    assert np.max(np.abs(got[:L] - exp[:L])) <= 4 * np.finfo(np.float32).eps * np.max(np.abs(exp))
    assert not got[L:].any()


# ---- phase inference ------------------------------------------------------------------

def test_phase_linear_oracle_closed_form():
    """Mode 0 equals m * e^{i f 2 pi k hop / N}, the accumulated phase agreeing with the
    product to double rounding (checked in float64 on small frame counts); bin 0 is real."""
    N, H, F = 64, 16, 50
    mag = np.random.default_rng(4).random((F, N // 2 + 1)).astype(np.float32) + 0.1
    re, im = ref.phase_infer(mag, N, H, half=True, mode=0)
    f = np.arange(F)[:, None]
    k = np.arange(N // 2 + 1)[None, :]
    exp = mag.astype(np.float64) * np.exp(1j * f * 2 * np.pi * k * H / N)
    assert np.max(np.abs((re + 1j * im) - exp)) <= 1e-6
    assert not im[:, 0].any()


@pytest.mark.parametrize("N", [64, 256])
@pytest.mark.parametrize("a", [0.5, -0.7])
def test_phase_minimum_oracle_one_pole(N, a):
    """Mode 1 on |1 / (1 - a e^{-i w})| returns the minimum-phase filter itself,
    1 / (1 - a e^{-i w}) (its cepstrum is causal: log H = sum a^n e^{-i n w} / n), up to
    the cepstral aliasing a^N and float rounding; |out| equals the magnitude given."""
    k = np.arange(N // 2 + 1)
    h = 1.0 / (1.0 - a * np.exp(-2j * np.pi * k / N))
    mag = np.abs(h)[None, :].astype(np.float32).repeat(3, axis=0)
    re, im = ref.phase_infer(mag, N, 0, half=True, mode=1)
    got = re + 1j * im
    assert np.max(np.abs(got - h[None, :])) <= 1e-5 * np.max(np.abs(h))
    assert np.max(np.abs(np.abs(got) - mag)) <= 1e-5 * np.max(mag)


def test_phase_minimum_oracle_check_has_teeth():
    """A planted mistake (the anti-causal half of the cepstrum kept, i.e. zero phase)
    misses the one-pole filter by far."""
    N, a = 256, 0.5
    k = np.arange(N // 2 + 1)
    h = 1.0 / (1.0 - a * np.exp(-2j * np.pi * k / N))
    assert np.max(np.abs(np.abs(h) - h)) > 0.3 * np.max(np.abs(h))

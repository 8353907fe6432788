"""Large transforms (N above one workgroup's LDS): two-pass C2C through the C ABI.

Tolerance: relative L2 <= 2e-7 * log2(N) against the FP64 oracle (fp64.tolerance).
"""
import numpy as np
import pytest

from fftfree_amd import CffiFft
from fftfree_amd.workloads import complex_normal, noise_plus_sines
from oracle import fp64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


@pytest.mark.parametrize("lg", [13, 14, 15, 16, 17, 18, 19, 20, 21])
@pytest.mark.parametrize("inverse", [0, 1])
@pytest.mark.parametrize("kernel", [0, 2])
def test_c2c_large_sizes(ff, lg, inverse, kernel):
    n = 1 << lg
    batch = max(2, (1 << 21) // n)
    print(f"large c2c n={n} batch={batch} inverse={inverse} kernel={kernel}", flush=True)
    p = ff.make_plan(n, inverse=inverse, kernel=kernel, transform=0, half_spectrum=False)
    x = complex_normal(batch, n, seed=lg + 31 * inverse)
    y = ff.c2c(p, x)
    assert fp64.rel_l2(y, fp64.fft64(x, bool(inverse))) <= fp64.tolerance(n)


def test_c2c_large_impulses_and_tones(ff):
    """Closed forms: DFT of delta[n - m] is W^{mk}; DFT of a pure tone e^{2 pi i f n / N}
    is N at bin f and 0 elsewhere."""
    n = 1 << 16
    p = ff.make_plan(n, transform=0, half_spectrum=False)
    x = np.zeros((4, n), np.complex64)
    ms = [0, 1, 12345, n - 1]
    for b, m in enumerate(ms):
        x[b, m] = 1.0
    y = ff.c2c(p, x)
    k = np.arange(n)
    for b, m in enumerate(ms):
        want = np.exp(-2j * np.pi * ((m * k) % n) / n)
        assert np.max(np.abs(y[b] - want)) <= 1e-5
    f = 777
    tone = np.exp(2j * np.pi * ((f * k) % n) / n).astype(np.complex64)[None, :]
    yt = ff.c2c(p, tone)
    want = np.zeros(n)
    want[f] = n
    assert np.max(np.abs(yt[0] - want)) <= 1e-5 * n


def test_c2c_large_out_of_place_device(ff):
    import torch
    n = 1 << 15
    p = ff.make_plan(n, transform=0, half_spectrum=False)
    x = complex_normal(8, n, seed=3)
    d_in = torch.from_numpy(x).cuda()
    keep = d_in.clone()
    d_out = torch.empty_like(d_in)
    ff.c2c_device(p, d_in, d_out)
    torch.cuda.synchronize()
    assert torch.equal(d_in, keep)  # input untouched
    assert fp64.rel_l2(d_out.cpu().numpy(), fp64.fft64(x)) <= fp64.tolerance(n)
    ff.c2c_device(p, d_in)  # in place gives the same bits
    torch.cuda.synchronize()
    assert torch.equal(d_in, d_out)


FLAGS = [  # transform, reduce_magnitude, store_polar, half_spectrum
    (1, 0, 0, 1), (1, 1, 0, 1), (1, 0, 1, 1), (1, 0, 0, 0), (2, 0, 0, 0), (0, 1, 0, 0)]


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("kernel", [0, 2])
def test_c2c_large_hooks(ff, flags, kernel):
    """Pre/post hooks on the two-pass path against the FP64 restatement of the hooks."""
    transform, mag, polar, half = flags
    n = 1 << 14
    p = ff.make_plan(n, kernel=kernel, transform=transform, reduce_magnitude=bool(mag), store_polar=bool(polar),
                     half_spectrum=bool(half))
    x = complex_normal(6, n, seed=17)
    y = ff.c2c(p, x)
    xin = x.real.astype(np.complex128) if (transform == 1 and kernel == 0) else x.astype(np.complex128)
    X = fp64.fft64(xin)
    ore, oim, omg = fp64.post_planes(X, n, transform, bool(mag), bool(polar), bool(half))
    bins = ore.shape[1]
    if bins < n:  # R2C half spectrum: bins above N/2 are cleared
        assert not np.any(y[:, bins:])
    y = y[:, :bins]
    tol = fp64.tolerance(n)
    if polar:
        # compare as complex values where the magnitude is not tiny (phase is ill-posed at 0)
        assert fp64.rel_l2(y.real * np.exp(1j * y.imag), ore * np.exp(1j * oim)) <= tol
    else:
        assert fp64.rel_l2(y.real, ore) <= tol
        assert fp64.rel_l2(y.imag, oim) <= tol if np.any(oim) else not np.any(y.imag)


@pytest.mark.parametrize("polar", [0, 1])
def test_c2c_large_c2r_half_rebuild(ff, polar):
    """Cooley-Tukey C2R half-spectrum plans rebuild the Hermitian spectrum from bins
    0..N/2 before transforming (eigen_fft.hpp:2505-2531): the inverse of a real signal's
    half spectrum gives the signal back."""
    n = 1 << 15
    rng = np.random.default_rng(polar)
    sig = rng.standard_normal((3, n))
    X = np.fft.fft(sig, axis=-1)
    half = np.zeros((3, n), np.complex64)
    h = X[:, : n // 2 + 1]
    half[:, : n // 2 + 1] = (np.abs(h) + 1j * np.angle(h)) if polar else h
    p = ff.make_plan(n, inverse=1, kernel=0, transform=2, store_polar=bool(polar), half_spectrum=True)
    y = ff.c2c(p, half)
    if polar:  # the post hook stores (|x|, arg x) of the result (phase ~0 or ~pi)
        assert np.all(y.real >= 0)
        y = y.real * np.cos(y.imag)
    else:
        assert not np.any(y.imag)
        y = y.real
    assert fp64.rel_l2(y, sig) <= 2 * fp64.tolerance(n)


@pytest.mark.parametrize("lg", [20, 21])
@pytest.mark.parametrize("inverse", [0, 1])
def test_c2c_large_every_variant(ff, lg, inverse):
    """Every (column pass, row pass) kernel pair of the two-pass transform meets the FP64
    bound; the pipelined persistent passes do the arithmetic of the one-tile-per-workgroup
    LDS-table passes in the same order, so their pairs give the same bits as those."""
    n = 1 << lg
    batch = 5  # tiles of 8 lines: 5 * 128 or 5 * 256 of them, not a multiple of the CUs
    p = ff.make_plan(n, inverse=inverse, kernel=2, transform=0, half_spectrum=False)
    names = ff.variants(p)
    x = complex_normal(batch, n, seed=lg + 7 * inverse)
    want = fp64.fft64(x, bool(inverse))
    outs = {}
    for v, name in enumerate(names):
        print(f"large variant n={n} inverse={inverse} {v}: {name}", flush=True)
        assert ff.set_variant(p, v)
        outs[name] = ff.c2c(p, x)
        assert fp64.rel_l2(outs[name], want) <= fp64.tolerance(n), name
    cols = sorted({nm.split(" | ")[0] for nm in names})
    rows = sorted({nm.split(" | ")[1] for nm in names})
    pcol = [c for c in cols if "pipelined" in c]
    prow = [r for r in rows if "pipelined" in r]
    assert pcol or prow
    plain_col = [c for c in cols if c.endswith("lds-tw e=16 lines=8") and "pipelined" not in c]
    plain_row = [r for r in rows if r.endswith("lds-tw e=16 lines=8") and "pipelined" not in r]
    for c in pcol:
        for r in plain_row:
            assert np.array_equal(outs[f"{c} | {r}"], outs[f"{plain_col[0]} | {r}"])
    for r in prow:
        for c in plain_col:
            assert np.array_equal(outs[f"{c} | {r}"], outs[f"{c} | {plain_row[0]}"])


@pytest.mark.parametrize("lg", [20, 22])
def test_large_c2c_batch16_roundtrip(ff, lg):
    """2^20 and 2^22 points, batch 16, complex normal input: every forward transform meets
    the FP64 bound and forward then inverse returns the input."""
    import torch
    n = 1 << lg
    print(f"large roundtrip n={n} batch=16", flush=True)
    fwd = ff.make_plan(n, transform=0, half_spectrum=False, kernel=2)
    inv = ff.make_plan(n, inverse=1, transform=0, half_spectrum=False, kernel=2)
    x = complex_normal(16, n, seed=lg)
    d = torch.from_numpy(x).cuda()
    y = torch.empty_like(d)
    ff.c2c_device(fwd, d, y)
    z = torch.empty_like(d)
    ff.c2c_device(inv, y, z)
    torch.cuda.synchronize()
    yh = y.cpu().numpy()
    for b in range(16):  # every transform of the batch against the FP64 DFT
        assert fp64.rel_l2(yh[b], fp64.fft64(x[b])) <= fp64.tolerance(n), b
    zh = z.cpu().numpy()
    assert fp64.rel_l2(zh, x) <= 2 * fp64.tolerance(n)
    # Parseval on every transform: ||X||^2 = N ||x||^2
    ex = np.sum(np.abs(x.astype(np.complex128)) ** 2, axis=1)
    ey = np.sum(np.abs(yh.astype(np.complex128)) ** 2, axis=1)
    assert np.max(np.abs(ey / (n * ex) - 1)) <= 1e-5


@pytest.mark.parametrize("n", [16384, 65536])
def test_stft_large_n(ff, n):
    """STFT with N above the fused kernel's LDS limit (framing + two-pass transform +
    plane kernel) against the FP64 oracle."""
    W, H = n, n // 4
    pcm = noise_plus_sines(8 * n, seed=n)
    p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, reduce_magnitude=True, half_spectrum=True,
                     apply_windows=True, analysis_window=1, pad_mode=0)
    got, re, im, mg = ff.stft_frames(p, pcm, pad_mode=0)
    assert got == fp64.frame_count(pcm.size, n, W, H, 0, 0)
    win = fp64.build_window(1, W)
    ore, oim, omg = fp64.stft_planes(pcm, n, W, H, got, win, False, 1, True, False, True)
    assert fp64.rel_l2(mg, omg) <= fp64.tolerance(n)
    assert np.array_equal(re, mg) and not im.any()

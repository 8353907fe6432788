"""fft_griffin_lim / fft_griffin_lim_device (phase_infer.hpp:69-204) against the float
restatement of the reference loop (oracle/cpu_ref.cpp ref_griffin_lim).

Bar: the complex spectra agree within 2 * fp64.tolerance(N) * sqrt(iterations) in
relative L2; the start alone is exact up to cos/sin rounding (its angles are the
reference's bit for bit, tests/test_oracle.py), and each iteration adds transform
rounding of order tolerance(N) (the two oracle restatements drift apart at that rate).
Properties at a larger size: |X| = mag, real bins 0 and N/2, spectral convergence that
does not increase."""
import ctypes

import numpy as np
import pytest

from fftfree_amd import CffiFft
from fftfree_amd.workloads import noise_plus_sines
from oracle import fp64, ref

pytestmark = pytest.mark.gpu
R2C, C2R = 1, 2


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


def spectrogram(N, H, length, seed, window=None):
    x = noise_plus_sines(length, seed=seed)
    frames = fp64.frame_count(x.size, N, N, H, 1, 1)
    _, _, mag = fp64.stft_planes(x, N, N, H, frames, window, False, R2C, False, False, True)
    return mag.astype(np.float32)


def plans(ff, N, H, inv_kernel=0, windows=False):
    kw = dict(apply_windows=windows, analysis_window=fp64.WINDOW_HANN, synthesis_window=fp64.WINDOW_HANN)
    fwd = ff.make_plan(N, window=N, hop=H, inverse=0, kernel=0, transform=R2C, half_spectrum=True, pad_mode=1, **kw)
    inv = ff.make_plan(N, window=N, hop=H, inverse=1, kernel=inv_kernel, transform=C2R, half_spectrum=True,
                       pad_mode=1, **kw)
    return fwd, inv


CASES = [(512, 128, 60, 1), (1024, 256, 50, 8), (1024, 256, 50, 1), (2048, 512, 30, 4), (256, 64, 80, 4),
         (512, 512, 40, 3), (1024, 200, 45, 2)]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("windows", [False, True])
def test_gl_vs_oracle(ff, case, windows):
    """Fused path (inverse C2R half N in {512, 1024, 2048}, hop <= N) and general path
    (N = 256), with and without Hann analysis/synthesis windows, against the float
    restatement of the reference and the FP64 Griffin-Lim from the same start angles.

    Bound: 2 * tolerance(N) * sqrt(iterations), or twice the float restatement's own
    distance from FP64 where that is larger: the phase normalisation X/|X| amplifies
    rounding at bins of small |X|, so some inputs are ill-conditioned for any float32
    implementation (N=2048, 4 iterations, Hann: the reference restatement is 1.6e-5 from
    FP64 against a bound of 8.8e-6); the device result must then be no further from FP64
    than twice the reference's own float error, and no further from the reference."""
    N, H, frames, iters = case
    print(f"griffin-lim N={N} H={H} frames={frames} iters={iters} windows={windows}", flush=True)
    w = np.concatenate([fp64.build_window(fp64.WINDOW_HANN, N)]) if windows else None
    mag = spectrogram(N, H, (frames - 1) * H + N, seed=N + H, window=w)
    fwd, inv = plans(ff, N, H, windows=windows)
    re, im = ff.griffin_lim(fwd, inv, mag, hop=H, iterations=iters, seed=99)
    ok, rr, ri = ref.griffin_lim(mag, N, H, iterations=iters, seed=99, inv_synth=w, fwd_taps=w)
    assert ok == 1
    X64 = fp64.griffin_lim64(mag, N, H, ref.gl_angles(99, mag.size).reshape(mag.shape), iters, synth=w, analysis=w)
    ref_err = fp64.rel_l2(rr + 1j * ri, X64)
    bound = max(2 * fp64.tolerance(N) * max(1.0, np.sqrt(iters)), 2 * ref_err)
    err = fp64.rel_l2(re + 1j * im, rr + 1j * ri)
    err64 = fp64.rel_l2(re + 1j * im, X64)
    print(f"  vs reference {err:.3e}, vs FP64 {err64:.3e}, reference vs FP64 {ref_err:.3e}, bound {bound:.3e}")
    assert err <= bound, err
    assert err64 <= bound, err64


def test_gl_device_entry_matches_host_entry(ff):
    import torch
    N, H = 1024, 256
    mag = spectrogram(N, H, 100 * H + N, seed=5)
    fwd, inv = plans(ff, N, H)
    re, im = ff.griffin_lim(fwd, inv, mag, hop=H, iterations=5, seed=3)
    d = torch.from_numpy(mag).cuda()
    dr, di = ff.griffin_lim_device(fwd, inv, d, hop=H, iterations=5, seed=3)
    torch.cuda.synchronize()
    assert np.array_equal(dr.cpu().numpy(), re) and np.array_equal(di.cpu().numpy(), im)


def test_gl_repeated_calls_reuse_the_start_words(ff):
    """The start words depend only on the seed and their count, so a context keeps them on
    the device between calls: calls that repeat a seed give the same bits as the first,
    whatever ran in between (another seed, another frame count, the host entry)."""
    N, H = 1024, 256
    mag = spectrogram(N, H, 60 * H + N, seed=6)
    short = np.ascontiguousarray(mag[:37])
    fwd, inv = plans(ff, N, H)
    first = ff.griffin_lim(fwd, inv, mag, hop=H, iterations=3, seed=11)
    again = ff.griffin_lim(fwd, inv, mag, hop=H, iterations=3, seed=11)
    other = ff.griffin_lim(fwd, inv, mag, hop=H, iterations=3, seed=12)
    cut = ff.griffin_lim(fwd, inv, short, hop=H, iterations=3, seed=11)
    third = ff.griffin_lim(fwd, inv, mag, hop=H, iterations=3, seed=11)
    fresh_f, fresh_i = plans(ff, N, H)
    fresh = ff.griffin_lim(fresh_f, fresh_i, short, hop=H, iterations=3, seed=11)
    for x in (again, third):
        assert np.array_equal(x[0], first[0]) and np.array_equal(x[1], first[1])
    assert not np.array_equal(other[0], first[0])
    assert np.array_equal(cut[0], fresh[0]) and np.array_equal(cut[1], fresh[1])


def test_gl_refusals(ff):
    """The reference's argument checks (fft_cffi.cpp:2187-2212): NULL or aliased buffers,
    a half_spectrum flag that differs from the contexts; plus contexts of different N."""
    N, H = 512, 128
    mag = spectrogram(N, H, 20 * H + N, seed=1)
    fwd, inv = plans(ff, N, H)
    lib = ff.lib
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    re, im = np.zeros_like(mag), np.zeros_like(mag)
    f = mag.shape[0]
    assert lib.fft_griffin_lim(fwd.handle, inv.handle, fp(mag), f, H, 1, 2, 1, 0, fp(mag), fp(im)) == 0
    assert lib.fft_griffin_lim(fwd.handle, inv.handle, fp(mag), f, H, 1, 2, 1, 0, fp(re), fp(re)) == 0
    assert lib.fft_griffin_lim(fwd.handle, inv.handle, fp(mag), f, H, 0, 2, 1, 0, fp(re), fp(im)) == 0
    assert lib.fft_griffin_lim(None, inv.handle, fp(mag), f, H, 1, 2, 1, 0, fp(re), fp(im)) == 0
    assert lib.fft_griffin_lim(fwd.handle, inv.handle, fp(mag), 0, H, 1, 2, 1, 0, fp(re), fp(im)) == 0
    other = ff.make_plan(2 * N, window=2 * N, hop=H, transform=R2C, half_spectrum=True)
    assert lib.fft_griffin_lim(other.handle, inv.handle, fp(mag), f, H, 1, 2, 1, 0, fp(re), fp(im)) == 0
    assert lib.fft_griffin_lim(fwd.handle, inv.handle, fp(mag), f, H, 1, 2, 1, 0, fp(re), fp(im)) == 1


def test_gl_properties_at_size(ff):
    """4093 frames of N=1024, hop 256: |X| = mag, real bins 0 and N/2, and the spectral
    convergence after 32 iterations below that after 4."""
    import torch
    N, H = 1024, 256
    mag = spectrogram(N, H, 1 << 20, seed=21)
    fwd, inv = plans(ff, N, H)
    d = torch.from_numpy(mag).cuda()
    scs = []
    for it in (4, 32):
        dr, di = ff.griffin_lim_device(fwd, inv, d, hop=H, iterations=it, seed=1)
        torch.cuda.synchronize()
        r, i = dr.cpu().numpy(), di.cpu().numpy()
        assert np.max(np.abs(np.hypot(r, i) - mag)) <= 2e-6 * np.max(mag)
        assert not i[:, 0].any() and not i[:, -1].any()
        X = r.astype(np.float64) + 1j * i
        fr = np.fft.irfft(X, n=N, axis=-1)
        frames = mag.shape[0]
        L = (frames - 1) * H + N
        ola, w = np.zeros(L), np.zeros(L)
        for f in range(frames):
            ola[f * H:f * H + N] += fr[f]
            w[f * H:f * H + N] += 1
        ola /= w
        Y = np.abs(np.fft.rfft(np.stack([ola[f * H:f * H + N] for f in range(frames)]), axis=-1))
        scs.append(np.linalg.norm(Y - mag) / np.linalg.norm(mag))
    print(f"spectral convergence 4 iterations {scs[0]:.4f}, 32 iterations {scs[1]:.4f}", flush=True)
    assert scs[1] < scs[0]

"""Pins of the grid-pipeline restatement (oracle ref_ft_pipeline), CPU only.

Each expectation is computed independently of the oracle's loops, in float64 numpy:
* a constant field passes through up- and downsampling unchanged;
* a field linear in (t, f), upsampled by (u, v) and box-averaged back, is the field at
  (t + (u-1)/2u, f + (v-1)/2v) except in the last row/column, where the bilinear weights
  meet the clamped neighbour and the value stays put; without energy preservation the
  box sums are u*v times that;
* a ragged downsampling block is checked against a vectorised bilinear + block mean;
* the quantizer with weight 1 (mix 1, base <= 0) rounds magnitudes to the step grid and
  keeps phases; with a base frequency, the weight mix * exp(-decay * |r - round r|)
  takes its closed values at harmonics and half-way between them;
* the contrast booster with a 1x1 kernel adds the brightness only; the global mode and a
  local box are checked against cumulative-sum box means; gamma is a power law;
* a result of another shape returns 0 and leaves the planes as they were.
"""
import numpy as np
import pytest

from oracle import ref


def _cplx(t, f, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((t, f)) + 1j * rng.standard_normal((t, f))).astype(np.complex64)


def _run_complex(z, **kw):
    ok, re, im = ref.ft_pipeline(z.real, z.imag, polar=False, **kw)
    return ok, re.astype(np.float64) + 1j * im.astype(np.float64)


def test_constant_field_passes_resampling():
    z = np.full((6, 5), 0.75 - 1.25j, dtype=np.complex64)
    ok, out = _run_complex(z, up=(2, 3), down=(2, 3))
    assert ok == 1
    assert np.max(np.abs(out - z)) <= 1e-6
    mag = np.full((6, 5), 2.0, np.float32)
    ph = np.full((6, 5), 0.5, np.float32)
    ok, m, p = ref.ft_pipeline(mag, ph, polar=True, up=(3, 2), down=(3, 2))
    assert ok == 1
    assert np.max(np.abs(m - 2.0)) <= 1e-5 and np.max(np.abs(p - 0.5)) <= 1e-5


@pytest.mark.parametrize("u,v", [(2, 2), (3, 1), (1, 4), (4, 3)])
@pytest.mark.parametrize("preserve", [True, False])
def test_linear_field_up_down_closed_form(u, v, preserve):
    T, F = 7, 9
    a, b, c = 0.5 + 0.25j, -0.125 + 0.5j, 1.0 - 2.0j
    tt, ff = np.meshgrid(np.arange(T, dtype=np.float64), np.arange(F, dtype=np.float64), indexing="ij")
    z = (a * tt + b * ff + c).astype(np.complex64)
    ok, out = _run_complex(z, up=(u, v), down=(u, v), preserve=preserve)
    assert ok == 1
    te = np.where(tt < T - 1, tt + (u - 1) / (2 * u), tt)
    fe = np.where(ff < F - 1, ff + (v - 1) / (2 * v), ff)
    want = a * te + b * fe + c
    if not preserve:
        want = want * (u * v)
    assert np.max(np.abs(out - want)) <= 2e-5 * np.max(np.abs(want))


def _bilinear64(z, u, v):
    T, F = z.shape
    st = np.arange(T * u) / u
    sf = np.arange(F * v) / v
    t0 = np.floor(st).astype(int)
    f0 = np.floor(sf).astype(int)
    t1 = np.minimum(t0 + 1, T - 1)
    f1 = np.minimum(f0 + 1, F - 1)
    wt = (st - t0)[:, None]
    wf = (sf - f0)[None, :]
    z = z.astype(np.complex128)
    lo = z[t0][:, f0] + (z[t1][:, f0] - z[t0][:, f0]) * wt
    hi = z[t0][:, f1] + (z[t1][:, f1] - z[t0][:, f1]) * wt
    return lo + (hi - lo) * wf


def _block_mean(z, d, e, preserve=True):
    T, F = z.shape
    To, Fo = -(-T // d), -(-F // e)
    out = np.zeros((To, Fo), np.complex128)
    for i in range(To):
        for j in range(Fo):
            blk = z[i * d:(i + 1) * d, j * e:(j + 1) * e]
            out[i, j] = blk.mean() if preserve else blk.sum()
    return out


@pytest.mark.parametrize("preserve", [True, False])
def test_ragged_downsample_block(preserve):
    # 4x3 grid, upsampled (4, 3) to 16x9, box-downsampled (5, 4) back to 4x3: the last
    # block in time holds one row, the last in frequency one column.
    z = _cplx(4, 3, 11)
    ok, out = _run_complex(z, up=(4, 3), down=(5, 4), preserve=preserve)
    assert ok == 1
    want = _block_mean(_bilinear64(z, 4, 3), 5, 4, preserve)
    assert want.shape == (4, 3)
    assert np.max(np.abs(out - want)) <= 2e-5 * np.max(np.abs(want))


def _off_tie_mags(shape, step, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 40, shape)
    frac = rng.choice([0.15, 0.3, 0.7, 0.85], shape)
    return ((k + frac) * step).astype(np.float32), np.where(frac < 0.5, k, k + 1)


def test_quantizer_weight_one_rounds_to_step_grid():
    step = np.float32(0.05)
    mag, k = _off_tie_mags((8, 17), step, 3)
    ph = np.random.default_rng(4).uniform(-3, 3, mag.shape).astype(np.float32)
    ok, m, p = ref.ft_pipeline(mag, ph, polar=True, quant=(0.0, float(step), 1.0, 8.0))
    assert ok == 1
    assert np.array_equal(p, ph)
    assert np.max(np.abs(m - k * np.float64(step))) <= 1e-6
    # complex grid: |z| rounds, the phase stays, exact zeros stay zero
    z = (mag * np.exp(1j * ph.astype(np.float64))).astype(np.complex64)
    z[0, :3] = 0
    ok, out = _run_complex(z, quant=(-1.0, float(step), 1.0, 8.0))
    assert ok == 1
    want = k * np.float64(step) * np.exp(1j * np.angle(z.astype(np.complex128)))
    want[0, :3] = 0
    assert np.max(np.abs(out - want)) <= 1e-5


@pytest.mark.parametrize("mix,decay", [(0.5, 8.0), (0.8, 2.0), (1.5, 0.5)])
def test_quantizer_harmonic_weight_closed_form(mix, decay):
    base, step, fstep = 55.0, 0.1, 13.75  # harmonics every 4 bins, half-way every 2
    F = 33
    mag, _ = _off_tie_mags((5, F), np.float32(step), 9)
    ph = np.zeros_like(mag)
    ok, m, _ = ref.ft_pipeline(mag, ph, polar=True, freq_step=fstep, quant=(base, step, mix, decay))
    assert ok == 1
    r = np.arange(F) * fstep / base
    w = np.clip(mix * np.exp(-decay * np.abs(r - np.floor(r + 0.5))), 0.0, 1.0)
    mag64 = mag.astype(np.float64)
    target = np.floor(mag64 / step + 0.5) * step
    want = mag64 + (target - mag64) * w[None, :]
    assert np.max(np.abs(m - want)) <= 1e-5
    # at the harmonics the weight is min(mix, 1); half-way it is mix * exp(-decay / 2)
    assert np.allclose(w[::4], min(mix, 1.0)) and np.allclose(w[2::4], min(mix * np.exp(-decay / 2), 1.0))


def test_quantizer_off_when_mix_not_positive():
    mag = np.random.default_rng(5).uniform(0, 2, (4, 6)).astype(np.float32)
    ph = np.zeros_like(mag)
    ok, m, _ = ref.ft_pipeline(mag, ph, polar=True, quant=(0.0, 0.05, 0.0, 8.0))
    assert ok == 1 and np.array_equal(m, mag)


def _box_mean(m, kt, kf):
    T, F = m.shape
    rt = T if kt <= 0 else kt // 2
    rf = F if kf <= 0 else kf // 2
    c = np.zeros((T + 1, F + 1))
    c[1:, 1:] = np.cumsum(np.cumsum(m.astype(np.float64), 0), 1)
    t0 = np.clip(np.arange(T) - rt, 0, T)
    t1 = np.clip(np.arange(T) + rt + 1, 0, T)
    f0 = np.clip(np.arange(F) - rf, 0, F)
    f1 = np.clip(np.arange(F) + rf + 1, 0, F)
    s = c[t1][:, f1] - c[t0][:, f1] - c[t1][:, f0] + c[t0][:, f0]
    n = (t1 - t0)[:, None] * (f1 - f0)[None, :]
    return s / n


def test_contrast_unit_kernel_adds_brightness():
    mag = np.random.default_rng(6).uniform(0, 1, (5, 7)).astype(np.float32)
    ph = np.random.default_rng(7).uniform(-3, 3, mag.shape).astype(np.float32)
    ok, m, p = ref.ft_pipeline(mag, ph, polar=True, contrast=(1, 1, -0.25, 0.5, 1.0, 1))
    assert ok == 1 and np.array_equal(p, ph)
    assert np.max(np.abs(m - np.maximum(mag.astype(np.float64) - 0.25, 0.0))) <= 1e-6
    ok, m, _ = ref.ft_pipeline(mag, ph, polar=True, contrast=(1, 1, -0.25, 0.5, 1.0, 0))
    assert np.max(np.abs(m - (mag.astype(np.float64) - 0.25))) <= 1e-6


@pytest.mark.parametrize("kt,kf", [(0, 0), (3, 3), (5, 2), (0, 3), (4, 0)])
@pytest.mark.parametrize("gain,bright,gamma,clamp", [(0.5, 0.0, 1.0, 1), (1.5, -0.3, 1.0, 0), (0.5, 0.1, 2.0, 1)])
def test_contrast_box_mean_closed_form(kt, kf, gain, bright, gamma, clamp):
    mag = np.random.default_rng(8).uniform(0, 1, (9, 12)).astype(np.float32)
    ph = np.zeros_like(mag)
    ok, m, _ = ref.ft_pipeline(mag, ph, polar=True, contrast=(kt, kf, bright, gain, gamma, clamp))
    assert ok == 1
    mean = _box_mean(mag, kt, kf)
    adj = mean + (mag - mean) * (1.0 + gain) + bright
    if clamp:
        adj = np.maximum(adj, 0.0)
    if gamma != 1.0:
        adj = np.maximum(adj, 1e-12) ** (1.0 / gamma)
    assert np.max(np.abs(m - adj)) <= 2e-6 * max(1.0, np.max(np.abs(adj)))


def test_contrast_on_complex_keeps_phase():
    z = _cplx(6, 8, 12)
    ok, out = _run_complex(z, contrast=(3, 3, 0.0, 0.5, 1.0, 1))
    assert ok == 1
    mag = np.abs(z.astype(np.complex128))
    mean = _box_mean(mag, 3, 3)
    want = np.maximum(mean + (mag - mean) * 1.5, 0.0) * np.exp(1j * np.angle(z.astype(np.complex128)))
    assert np.max(np.abs(out - want)) <= 1e-5


def test_shape_mismatch_leaves_planes():
    z = _cplx(4, 6, 13)
    a, b = z.real.copy(), z.imag.copy()
    ok, ra, rb = ref.ft_pipeline(a, b, up=(2, 1), down=(1, 1))
    assert ok == 0 and np.array_equal(ra, a) and np.array_equal(rb, b)
    ok, _, _ = ref.ft_pipeline(a, b, up=(1, 1), down=(1, 3))
    assert ok == 0
    # factors <= 0 read as 1: the identity
    ok, ra, rb = ref.ft_pipeline(a, b, up=(0, -2), down=(0, 0))
    assert ok == 1 and np.array_equal(ra, a) and np.array_equal(rb, b)


# The oracle against the reference's own ft_tools.hpp, compiled in place with the
# reference's flags (oracle/_ref/libft_ref.so; skipped where that is not built).
# Without the harmonic weight (base frequency <= 0 or no quantizer) the two agree bit for
# bit. The weight mix * exp(-decay * distance) goes through expf, which -ffast-math lets
# the compiler vectorise (glibc's vector expf) in one build and not in the other; there
# they agree to 1e-5 of the grid's largest value.
REF_CASES = [
    ({}, True),
    ({"up": (2, 2), "down": (2, 2)}, True),
    ({"up": (4, 3), "down": (5, 4)}, True),
    ({"up": (3, 1), "down": (3, 1), "preserve": False}, True),
    ({"quant": (0.0, 0.5, 1.0, 8.0), "contrast": (3, 3, 0.0, 2.0, 1.0, 0)}, True),
    ({"contrast": (0, 0, -0.1, 0.8, 0.5, 0)}, True),
    ({"contrast": (5, 2, 0.1, 1.0, 2.0, 1)}, True),
    ({"quant": (55.0, 0.05, 0.5, 8.0), "contrast": (3, 3, 0.0, 0.5, 1.0, 1), "freq_step": 43.06640625}, False),
    ({"up": (2, 2), "down": (2, 2), "quant": (55.0, 0.05, 0.5, 8.0), "contrast": (5, 2, 0.1, 1.0, 2.0, 1),
      "freq_step": 43.06640625}, False),
]


@pytest.mark.skipif(not ref.ref_ft_available(), reason="oracle/_ref/libft_ref.so not built (needs the reference tree)")
@pytest.mark.parametrize("polar", [False, True], ids=["complex", "polar"])
@pytest.mark.parametrize("case", range(len(REF_CASES)))
def test_oracle_matches_reference_header(polar, case):
    cfg, exact = REF_CASES[case]
    T, F = (4, 3) if cfg.get("up") == (4, 3) else (37, 65)
    rng = np.random.default_rng(100 + case)
    a = rng.standard_normal((T, F)).astype(np.float32)
    b = rng.standard_normal((T, F)).astype(np.float32)
    if polar:
        a = np.abs(a)
    ok, oa, ob = ref.ft_pipeline(a, b, polar=polar, **cfg)
    rok, ra, rb = ref.refsrc_ft_pipeline(a, b, polar=polar, **cfg)
    assert ok == rok == 1
    if exact:
        assert np.array_equal(oa, ra) and np.array_equal(ob, rb)
    else:
        scale = max(1.0, float(np.max(np.abs(ra))))
        assert np.max(np.abs(oa - ra)) <= 1e-5 * scale and np.max(np.abs(ob - rb)) <= 1e-5 * scale
    # a result of another shape: both refuse
    assert ref.refsrc_ft_pipeline(a, b, polar=polar, up=(2, 1))[0] == ref.ft_pipeline(a, b, polar=polar, up=(2, 1))[0] == 0

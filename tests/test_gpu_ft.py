"""ft_apply_complex_pipeline / ft_apply_polar_pipeline (ft_tools_cffi.cpp:76-170) on the
GPU against the float restatement (oracle/cpu_ref.cpp ref_ft_pipeline).

Tolerance: every step is a float computation per cell in the reference's order, and |z|
is the reference build's sqrt(fma(re, re, im * im)) bit for bit, but the device's
atan2f/sincosf/expf, its reciprocals in the box mean and the direction, and the
placement of fused multiply-adds differ from the host's by an ulp or two, so cells agree
to 2e-5 of the grid's largest value. The quantizer
rounds |z| / step to the nearest integer; where |z| lies within an ulp of a half step the
two sides may round to neighbouring steps. Such a cell differs by one step times the
quantizer weight, scaled by at most 1 + |contrast gain| when the contrast stage follows,
and its 3x3 neighbourhood shifts through the box mean: up to 9 cells per 10^4 may
differ, by at most 3 steps. Polar grids are compared as m * exp(i phase), since a
phase of +pi and -pi name the same cell.
"""
import numpy as np
import pytest

from fftfree_amd import ft_tools as ft
from oracle import ref

pytestmark = pytest.mark.gpu

TOL = 2e-5


def _grid(t, f, seed, kind="complex"):
    rng = np.random.default_rng(seed)
    if kind == "complex":
        return rng.standard_normal((t, f)).astype(np.float32), rng.standard_normal((t, f)).astype(np.float32)
    mag = rng.gamma(2.0, 0.5, (t, f)).astype(np.float32)
    ph = rng.uniform(-np.pi, np.pi, (t, f)).astype(np.float32)
    return mag, ph


def _as_complex(a, b, polar):
    a, b = a.astype(np.float64), b.astype(np.float64)
    return a * np.exp(1j * b) if polar else a + 1j * b


def _check(got, want, quant_step=0.0):
    scale = max(1.0, float(np.max(np.abs(want))))
    err = np.abs(got - want)
    bad = err > TOL * scale
    if quant_step > 0.0:
        assert bad.sum() <= 9 * max(1, bad.size // 10000), f"{bad.sum()} cells off (max {err.max():.3g})"
        assert np.all(err[bad] <= 3 * quant_step + TOL * scale), f"max {err.max():.3g}"
    else:
        assert not bad.any(), f"{bad.sum()} cells off, max {err.max():.3g} (scale {scale:.3g})"


def _run(polar, a, b, freq_step=1.0, up=(1, 1), down=(1, 1), preserve=True, quant=None, contrast=None):
    ga, gb = a.copy(), b.copy()
    spec = ft.grid_spec(freq_origin=0.0, freq_step=freq_step)
    plan = ft.resample_plan(up, down, preserve)
    q = ft.quantizer(*quant) if quant is not None else None
    c = ft.contrast(*contrast) if contrast is not None else None
    fn = ft.apply_polar_pipeline if polar else ft.apply_complex_pipeline
    ok = fn(ga, gb, spec, plan, q, c)
    rok, ra, rb = ref.ft_pipeline(a, b, polar=polar, freq_step=freq_step, up=up, down=down, preserve=preserve,
                                  quant=quant, contrast=contrast)
    return ok, ga, gb, rok, ra, rb


# (time bins, freq bins, up, down, preserve, quant, contrast): the reference tool's defaults
# (quantizer 55 Hz / 0.05 / 0.5 / 8, 3x3 contrast 0.5), STFT-shaped grids of N = 1024 and
# 2048, the largest contrast box of the one-pass kernel (17 x 33) and one past it, a
# quantizer that sets many cells to zero ahead of an unclamped contrast, ragged
# downsampling blocks (a grid of at most u rows upsampled by u and
# downsampled by u + 1 keeps its shape with a short last block), the global and semi-global contrast modes, gamma.
CASES = [
    (256, 513, (1, 1), (1, 1), True, (55.0, 0.05, 0.5, 8.0), (3, 3, 0.0, 0.5, 1.0, 1)),
    (128, 513, (2, 2), (2, 2), True, (55.0, 0.05, 0.5, 8.0), (3, 3, 0.0, 0.5, 1.0, 1)),
    (64, 1025, (1, 2), (1, 2), False, (110.0, 0.1, 0.9, 2.0), None),
    (4, 3, (4, 3), (5, 4), True, None, (5, 2, 0.1, 1.0, 2.0, 1)),
    (3, 2, (3, 2), (4, 3), False, (55.0, 0.05, 0.5, 8.0), None),
    (40, 30, (4, 3), (4, 3), True, None, (5, 2, 0.1, 1.0, 2.0, 1)),
    (33, 47, (3, 1), (3, 1), True, (0.0, 0.02, 1.0, 8.0), (0, 0, -0.1, 0.8, 0.5, 0)),
    (50, 70, (2, 2), (2, 2), True, (55.0, 0.05, 0.5, 8.0), (0, 3, 0.0, 0.5, 1.0, 1)),
    (17, 9, (1, 1), (1, 1), True, None, (4, 0, 0.2, -0.5, 1.0, 1)),
    (70, 300, (1, 1), (1, 1), True, (55.0, 0.05, 0.5, 8.0), (17, 33, 0.0, 0.5, 1.0, 1)),
    (70, 300, (1, 1), (1, 1), True, (55.0, 0.05, 0.5, 8.0), (19, 35, 0.05, 0.5, 1.5, 1)),
    (37, 260, (1, 1), (1, 1), True, (0.0, 0.5, 1.0, 8.0), (3, 3, 0.0, 2.0, 1.0, 0)),
    (1, 1, (3, 3), (3, 3), True, (55.0, 0.05, 0.5, 8.0), (3, 3, 0.0, 0.5, 1.0, 1)),
]


@pytest.mark.parametrize("polar", [False, True], ids=["complex", "polar"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}-up{c[2]}-down{c[3]}" for c in CASES])
def test_pipeline_vs_oracle(polar, case):
    T, F, up, down, preserve, quant, contrast = case
    print(f"ft pipeline polar={polar} {T}x{F} up={up} down={down} quant={quant} contrast={contrast}", flush=True)
    a, b = _grid(T, F, T * 1000 + F, "polar" if polar else "complex")
    ok, ga, gb, rok, ra, rb = _run(polar, a, b, freq_step=44100.0 / 1024.0, up=up, down=down, preserve=preserve,
                                   quant=quant, contrast=contrast)
    assert ok and rok == 1
    step = quant[1] * min(1.0, quant[2]) if quant is not None else 0.0
    _check(_as_complex(ga, gb, polar), _as_complex(ra, rb, polar), step)
    if polar and not (up != (1, 1) or down != (1, 1)):
        # no resampling: the phase plane is passed through untouched
        assert np.array_equal(gb, b)


def test_quantizer_only_off_tie_exact_steps():
    # magnitudes away from half steps with weight 1: every cell lands on k * step
    rng = np.random.default_rng(21)
    k = rng.integers(0, 60, (96, 257))
    frac = rng.choice([0.2, 0.35, 0.65, 0.8], k.shape)
    mag = ((k + frac) * 0.05).astype(np.float32)
    ph = rng.uniform(-3, 3, k.shape).astype(np.float32)
    gm, gp = mag.copy(), ph.copy()
    assert ft.apply_polar_pipeline(gm, gp, None, None, ft.quantizer(0.0, 0.05, 1.0, 8.0), None)
    want = np.where(frac < 0.5, k, k + 1) * np.float64(np.float32(0.05))
    assert np.max(np.abs(gm - want)) <= 1e-5 and np.array_equal(gp, ph)


def test_shape_mismatch_and_null_leave_planes():
    a, b = _grid(12, 20, 5)
    ga, gb = a.copy(), b.copy()
    assert not ft.apply_complex_pipeline(ga, gb, None, ft.resample_plan((2, 1), (1, 1)), None, None)
    assert np.array_equal(ga, a) and np.array_equal(gb, b)
    assert not ft.apply_polar_pipeline(ga, gb, None, ft.resample_plan((1, 1), (1, 3)), None, None)
    assert np.array_equal(ga, a) and np.array_equal(gb, b)
    lib = ft._lib.load()
    assert lib.ft_apply_complex_pipeline(None, None, None, None, None, None, 4, 4) == 0
    assert lib.ft_apply_polar_pipeline(None, None, None, None, ga.ctypes.data_as(ft._lib.f32p),
                                       gb.ctypes.data_as(ft._lib.f32p), 0, 20) == 0
    # every setting NULL: plan defaults (no resampling), no units: the identity
    assert ft.apply_complex_pipeline(ga, gb)
    assert np.array_equal(ga, a) and np.array_equal(gb, b)


def test_disabled_units_are_skipped():
    a, b = _grid(20, 33, 6)
    ga, gb = a.copy(), b.copy()
    assert ft.apply_complex_pipeline(ga, gb, None, None, ft.quantizer(enabled=False), ft.contrast(enabled=False))
    assert np.array_equal(ga, a) and np.array_equal(gb, b)


def _torch():
    import torch
    return torch


DEVICE_CASES = [
    ((1, 1), (1, 1), (55.0, 0.05, 0.5, 8.0), (3, 3, 0.0, 0.5, 1.0, 1)),  # one pass
    ((1, 1), (1, 1), (55.0, 0.05, 0.5, 8.0), None),  # one pass, elementwise
    ((1, 1), (1, 1), None, (0, 0, 0.0, 0.5, 1.0, 1)),  # global mean: step by step
    ((2, 2), (2, 2), (55.0, 0.05, 0.5, 8.0), (3, 3, 0.0, 0.5, 1.0, 1)),  # resampling: step by step
]


@pytest.mark.parametrize("polar", [False, True], ids=["complex", "polar"])
@pytest.mark.parametrize("case", DEVICE_CASES, ids=["fused", "fused-elementwise", "global", "resampled"])
def test_device_entry_matches_host_entry(polar, case):
    """The device form gives the host form's planes bit for bit: out of place, in place,
    and (polar) with the phase plane alone in place."""
    torch = _torch()
    up, down, quant, contrast = case
    print(f"ft device entry polar={polar} up={up} down={down} quant={quant} contrast={contrast}", flush=True)
    a, b = _grid(96, 513, 77, "polar" if polar else "complex")
    spec = ft.grid_spec(freq_step=44100.0 / 1024.0)
    plan = ft.resample_plan(up, down)
    q = ft.quantizer(*quant) if quant is not None else None
    c = ft.contrast(*contrast) if contrast is not None else None
    ha, hb = a.copy(), b.copy()
    host = ft.apply_polar_pipeline if polar else ft.apply_complex_pipeline
    dev = ft.apply_polar_pipeline_device if polar else ft.apply_complex_pipeline_device
    assert host(ha, hb, spec, plan, q, c)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    oa, ob = torch.full_like(da, float("nan")), torch.full_like(db, float("nan"))
    assert dev(da, db, oa, ob, spec, plan, q, c)
    torch.cuda.synchronize()
    assert np.array_equal(oa.cpu().numpy(), ha) and np.array_equal(ob.cpu().numpy(), hb)
    assert np.array_equal(da.cpu().numpy(), a) and np.array_equal(db.cpu().numpy(), b)  # inputs untouched
    ia, ib = da.clone(), db.clone()
    assert dev(ia, ib, None, None, spec, plan, q, c)
    torch.cuda.synchronize()
    assert np.array_equal(ia.cpu().numpy(), ha) and np.array_equal(ib.cpu().numpy(), hb)
    if polar:
        pa, pb = torch.full_like(da, float("nan")), db.clone()
        assert dev(da, pb, pa, pb, spec, plan, q, c)
        torch.cuda.synchronize()
        assert np.array_equal(pa.cpu().numpy(), ha) and np.array_equal(pb.cpu().numpy(), hb)


def test_device_entry_rejects_shape_change():
    torch = _torch()
    a = torch.zeros((8, 8), device="cuda")
    b = torch.zeros((8, 8), device="cuda")
    assert not ft.apply_complex_pipeline_device(a, b, None, None, None, ft.resample_plan((2, 2), (1, 1)))
    lib = ft._lib.load()
    assert lib.ft_apply_polar_pipeline_device(None, None, None, None, None, None, None, None, 8, 8, None) == 0

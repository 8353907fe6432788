"""Parity of the HIP path (through the C ABI) against the oracle.

Tolerance for floating-point outputs: relative L2 <= 2e-7 * log2(N) against the FP64
oracle (fp64.tolerance); integer/index outputs and frame counts: bit-exact.
"""
import numpy as np
import pytest

from fftfree_amd import CffiFft
from fftfree_amd.workloads import complex_normal, noise_plus_sines
from oracle import fp64, index_maps as M, ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    f = CffiFft()
    yield f
    f.close()


def spectra_close(re, im, ore, oim, polar, tol):
    if polar:
        return fp64.rel_l2(re * np.exp(1j * im), ore * np.exp(1j * oim)) <= tol
    return fp64.rel_l2(re + 1j * im, ore + 1j * oim) <= tol


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("inverse", [0, 1])
@pytest.mark.parametrize("kernel", [0, 2])
def test_c2c_all_lds_sizes(ff, n, inverse, kernel):
    print(f"c2c n={n} inverse={inverse} kernel={kernel}", flush=True)
    p = ff.make_plan(n, inverse=inverse, kernel=kernel, transform=0, half_spectrum=False)
    x = complex_normal(37, n, seed=n + 7 * inverse)
    y = ff.c2c(p, x)
    assert fp64.rel_l2(y, fp64.fft64(x, bool(inverse))) <= fp64.tolerance(n)


@pytest.mark.parametrize("inverse", [0, 1])
def test_c2c_1024_register_exchanges_same_bits(ff, inverse):
    """The N=1024 C2C variants with one wave per transform and the exchanges between the
    radix-16, 16, 4 passes in registers (1 and 4 transforms per workgroup) run the LDS
    16x16x4 kernel's butterflies and twiddles on permuted lanes: the same bits, within
    the FP64 bound, and with the magnitude post hook too."""
    n = 1024
    x = complex_normal(333, n, seed=91 + inverse)
    for magnitude in (False, True):
        p = ff.make_plan(n, inverse=inverse, kernel=2, transform=0, half_spectrum=False, reduce_magnitude=magnitude)
        names = ff.variants(p)
        lds = [i for i, m in enumerate(names) if m.startswith("16x16x4 fpw")]
        regs = [i for i, m in enumerate(names) if m.startswith("16x16x4 dpp")]
        assert len(lds) == 1 and len(regs) == 2, names
        ys = []
        for vi in lds + regs:
            assert ff.set_variant(p, vi)
            ys.append(ff.c2c(p, x))
        ff.set_variant(p, -1)
        for y in ys[1:]:
            assert np.array_equal(y.view(np.uint32), ys[0].view(np.uint32))
        want = fp64.fft64(x, bool(inverse))
        if magnitude:
            assert fp64.rel_l2(ys[-1].real, np.abs(want)) <= fp64.tolerance(n) and not ys[-1].imag.any()
        else:
            assert fp64.rel_l2(ys[-1], want) <= fp64.tolerance(n)


@pytest.mark.parametrize("inverse", [0, 1])
def test_c2c_512_register_exchanges_same_bits(ff, inverse):
    """The N=512 C2C variant whose exchanges between radix-8 passes are DPP / permlane
    swaps in registers runs the LDS kernel's butterflies and twiddles on permuted lanes:
    every output bit equal to the LDS kernel's, and within the FP64 bound."""
    n = 512
    p = ff.make_plan(n, inverse=inverse, kernel=2, transform=0, half_spectrum=False)
    names = ff.variants(p)
    lds = [i for i, m in enumerate(names) if m.startswith("8x8x8 fpw")]
    regs = [i for i, m in enumerate(names) if m.startswith("8x8x8 dpp")]
    assert len(lds) == 1 and regs == [0], names  # the register kernel is the default
    x = complex_normal(1000, n, seed=71 + inverse)
    ys = []
    for vi in (lds[0], regs[0]):
        assert ff.set_variant(p, vi)
        ys.append(ff.c2c(p, x))
    ff.set_variant(p, -1)
    assert np.array_equal(ys[0].view(np.uint32), ys[1].view(np.uint32))
    assert fp64.rel_l2(ys[1], fp64.fft64(x, bool(inverse))) <= fp64.tolerance(n)


def test_config1_c2c_1024_stockham_and_metadata(ff):
    """Plain C2C forward, Stockham (kernel code 2), N=1024, radix 4, batch 64, complex64
    N(0,1): FP64-oracle error within 2e-7*10; LayoutMap and TwiddleIndexMap bit-exact."""
    n = 1024
    p = ff.make_plan(n, kernel=2, radix=4, transform=0, half_spectrum=False)
    x = complex_normal(64, n, seed=1)
    y = ff.c2c(p, x)
    err = fp64.rel_l2(y, fp64.fft64(x))
    print(f"config1 rel_l2={err:.3e} tol={fp64.tolerance(n):.1e}")
    assert err <= fp64.tolerance(n)
    S = 10
    lm = ff.metadata(p, 1)
    assert np.array_equal(lm.real.astype(np.int64), M.layout_map(n)) and not lm.imag.any()
    ti = ff.metadata(p, 6)
    assert np.array_equal(ti.real.astype(np.int64).reshape(S, n), M.twiddle_index_map(n, 2)) and not ti.imag.any()
    bp = ff.metadata(p, 4)
    assert np.array_equal(np.stack([bp.real, bp.imag], -1).astype(np.int64).reshape(S, n, 2), M.butterfly_pairs(n, 2))
    sp = ff.metadata(p, 3)
    assert np.array_equal(np.stack([sp.real, sp.imag], -1).astype(np.int64), M.stage_params(n))
    assert ff.metadata(p, 7).size == 0  # Omega is recorded by Cooley-Tukey only


@pytest.mark.parametrize("n", [2, 8, 64, 1024, 4096])
@pytest.mark.parametrize("kernel", [0, 2])
@pytest.mark.parametrize("inverse", [0, 1])
def test_metadata_all_kinds_bit_exact(ff, n, kernel, inverse):
    p = ff.make_plan(n, kernel=kernel, inverse=inverse, transform=1, half_spectrum=True)
    S = M.log2_exact(n)
    ti = ff.metadata(p, 6)
    assert np.array_equal(ti.real.astype(np.int64).reshape(S, n), M.twiddle_index_map(n, kernel))
    bp = ff.metadata(p, 4)
    assert np.array_equal(np.stack([bp.real, bp.imag], -1).astype(np.int64).reshape(S, n, 2),
                          M.butterfly_pairs(n, kernel))
    om = ff.metadata(p, 7)
    if kernel == 2:
        assert om.size == 0
    else:
        assert np.array_equal(om.real.view(np.uint32), M.omega(n, 1, True).view(np.uint32))
    # Twiddle metadata holds the plan's FP64-built twiddles: exact to float rounding
    tw = ff.metadata(p, 0)
    idx = M.twiddle_metadata_index(n)
    sgn = 1.0 if inverse else -1.0
    assert np.max(np.abs(tw - np.exp(sgn * 2j * np.pi * idx / n))) <= 6e-8


def stft_case(ff, pcm, n, W, H, *, kernel=2, inverse=0, transform=1, reduce_mag=0, polar=0, half=1, windows=1,
              win_kind=1, pad_mode=1, ctx_pad=0, max_frames=0):
    p = ff.make_plan(n, window=W, hop=H, kernel=kernel, inverse=inverse, transform=transform,
                     reduce_magnitude=bool(reduce_mag), store_polar=bool(polar), half_spectrum=bool(half),
                     apply_windows=bool(windows), analysis_window=win_kind, pad_mode=ctx_pad)
    got, re, im, mg = ff.stft_frames(p, pcm, pad_mode=pad_mode, max_frames=max_frames)
    frames = fp64.frame_count(pcm.size, n, W, H, ctx_pad, pad_mode, max_frames)
    assert got == frames
    win = fp64.build_window(win_kind, W or n) if windows else None
    return p, got, (re, im, mg), win


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192])
def test_stft_r2c_half_magnitude_all_sizes(ff, n):
    print(f"stft n={n}", flush=True)
    pcm = noise_plus_sines(20 * n + 3, seed=n)
    W, H = n, max(1, n // 4)
    # the symmetric Hann formula divides by W-1, so windows start at W=4
    p, got, (re, im, mg), win = stft_case(ff, pcm, n, W, H, reduce_mag=1, windows=int(n >= 4))
    ore, oim, omg = fp64.stft_planes(pcm, n, W, H, got, win, False, 1, True, False, True)
    assert fp64.rel_l2(mg, omg) <= fp64.tolerance(n)
    assert fp64.rel_l2(re, ore) <= fp64.tolerance(n)
    assert not im.any()


FLAGS = [(t, r, pz, h) for t in (0, 1, 2, 3) for (r, pz, h) in [(0, 0, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 0)]]


@pytest.mark.parametrize("transform,reduce_mag,polar,half", FLAGS)
@pytest.mark.parametrize("kernel", [0, 2])
@pytest.mark.parametrize("inverse", [0, 1])
def test_stft_flag_matrix_vs_oracle_and_reference(ff, transform, reduce_mag, polar, half, kernel, inverse):
    n, W, H = 256, 200, 70
    pcm = noise_plus_sines(3000, seed=transform * 10 + kernel)
    p, got, (re, im, mg), win = stft_case(ff, pcm, n, W, H, kernel=kernel, inverse=inverse, transform=transform,
                                          reduce_mag=reduce_mag, polar=polar, half=half)
    taps = np.ones(n, np.float32)
    taps[:W] = win
    rgot, rre, rim, rmg = ref.execute_batched(pcm, n, W, H, ctx_pad=0, pad_mode=1, inverse=bool(inverse),
                                              kernel=kernel, transform=transform, reduce_magnitude=reduce_mag,
                                              store_polar=polar, half_spectrum=half, window_taps=taps)
    assert rgot == got
    tol = fp64.tolerance(n)
    if transform == 2 and half and kernel != 2:
        # Cooley-Tukey's C2R half hook rebuilds a Hermitian input from the framed samples:
        # the restated reference path is the oracle here
        assert spectra_close(re, im, rre, rim, polar, 4 * tol)
        assert fp64.rel_l2(mg, rmg) <= 4 * tol
        return
    ore, oim, omg = fp64.stft_planes(pcm, n, W, H, got, win, bool(inverse), transform, reduce_mag, polar, half)
    assert spectra_close(re, im, ore, oim, polar, tol)
    assert fp64.rel_l2(mg, omg) <= tol
    assert spectra_close(re, im, rre, rim, polar, 4 * tol)


def test_reference_batched_stft_vs_eigen_case(ff):
    """cffi_stft_compare_test.cpp:406-475: N=W=512, H=128, 8 frames, Hann through v2
    apply_windows=1, pad_mode=1, R2C full spectrum; max abs error <= 1e-4."""
    n, W, H = 512, 512, 128
    pcm = np.random.default_rng(42).standard_normal(W + 7 * H).astype(np.float32)
    p = ff.make_plan(n, window=W, hop=H, transform=1, half_spectrum=False, apply_windows=True, analysis_window=1,
                     synthesis_window=1, pad_mode=1)
    got, re, im, mg = ff.stft_frames(p, pcm, pad_mode=1)
    assert got == 8
    ore, oim, _ = fp64.stft_planes(pcm, n, W, H, 8, fp64.build_window(1, W), False, 1, False, False, False)
    assert np.max(np.abs((re + 1j * im) - (ore + 1j * oim))) <= 1e-4


def test_reference_plain_fft_vs_eigen_case(ff):
    """cffi_stft_compare_test.cpp:344-404: fft_execute, N=1024 normal PCM, R2C, 1e-4."""
    n = 1024
    pcm = np.random.default_rng(1337).standard_normal(n).astype(np.float32)
    p = ff.make_plan(n, transform=1, half_spectrum=False, pad_mode=1)
    re = np.zeros(n, np.float32)
    im = np.zeros(n, np.float32)
    mg = np.zeros(n, np.float32)
    fp = lambda a: a.ctypes.data_as(ff.lib.fft_execute.argtypes[1])
    assert ff.lib.fft_execute(p.handle, fp(pcm), fp(re), fp(im), fp(mg), n) == 1
    X = fp64.fft64(pcm.astype(np.float64))
    assert np.max(np.abs((re + 1j * im) - X)) <= 1e-4 * max(1.0, np.max(np.abs(X))) or \
        fp64.rel_l2(re + 1j * im, X) <= fp64.tolerance(n)
    assert fp64.rel_l2(mg, np.abs(X)) <= fp64.tolerance(n)


@pytest.mark.parametrize("case", ["short_pad", "short_refuse", "max_frames", "w_lt_n", "odd_hop", "no_window",
                                  "unaligned"])
def test_stft_edge_cases(ff, case):
    n = 1024
    if case == "short_pad":
        pcm = noise_plus_sines(300, seed=1)
        p, got, (re, im, mg), win = stft_case(ff, pcm, n, 1024, 256, pad_mode=1)
        assert got == 1
    elif case == "short_refuse":
        pcm = noise_plus_sines(300, seed=1)
        p = ff.make_plan(n, window=1024, hop=256)
        assert ff.stft_frames(p, pcm, pad_mode=2)[0] == 0
        return
    elif case == "max_frames":
        pcm = noise_plus_sines(50000, seed=2)
        p, got, (re, im, mg), win = stft_case(ff, pcm, n, 1024, 256, max_frames=17)
        assert got == 17
    elif case == "w_lt_n":
        pcm = noise_plus_sines(20000, seed=3)
        p, got, (re, im, mg), win = stft_case(ff, pcm, n, 600, 150)
    elif case == "odd_hop":
        pcm = noise_plus_sines(20000, seed=4)
        p, got, (re, im, mg), win = stft_case(ff, pcm, n, 1024, 257)
    elif case == "no_window":
        pcm = noise_plus_sines(20000, seed=5)
        p, got, (re, im, mg), win = stft_case(ff, pcm, n, 1024, 256, windows=0)
    else:  # unaligned: the PCM view starts one float past an 8-byte boundary
        big = noise_plus_sines(20001, seed=6)
        pcm = big[1:]
        p, got, (re, im, mg), win = stft_case(ff, pcm, n, 1024, 256)
    W = p.window
    ore, oim, omg = fp64.stft_planes(pcm, n, W, p.hop, got, win, False, 1, False, False, True)
    assert fp64.rel_l2(re + 1j * im, ore + 1j * oim) <= fp64.tolerance(n)
    assert fp64.rel_l2(mg, omg) <= fp64.tolerance(n)


def _every_frame_rel_l2(pcm, mg, n, W, H, frames, win):
    """Relative L2 of a magnitude plane over every frame against the FP64 oracle, framed
    with pcm * win rounded to float32, in chunks of 4096 frames."""
    num = den = 0.0
    for f0 in range(0, frames, 4096):
        f1 = min(frames, f0 + 4096)
        x = fp64.frame_signal(pcm[f0 * H:(f1 - 1) * H + W], n, W, H, f1 - f0, win)
        want = np.abs(np.fft.rfft(x.astype(np.float64), axis=-1))
        num += float(np.sum((mg[f0:f1].astype(np.float64) - want) ** 2))
        den += float(np.sum(want ** 2))
    return (num / den) ** 0.5


def test_config2_full_size_every_frame(ff):
    """Batched R2C STFT, Hann 1024, hop 256, half + magnitude, 2^24 samples: frame count
    exact, every one of the 65533 frames of the magnitude plane against the FP64 oracle,
    re = mag and im = 0 bit for bit, for both plane layouts a caller can hand over:
    three separate 16-byte aligned planes (the line-aligned kernel) and one (3, frames,
    bins) tensor whose second and third planes start 4 bytes past a 16-byte boundary
    (the general fused kernel, which rounds differently). The host entry, whose staging
    slots are aligned, equals the aligned device launch bit for bit."""
    import torch
    n, W, H = 1024, 1024, 256
    pcm = noise_plus_sines(1 << 24, seed=2)
    p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, reduce_magnitude=True, half_spectrum=True,
                     apply_windows=True, analysis_window=1, pad_mode=0)
    frames = ff.frame_count(p, pcm.size, 0)
    assert frames == 65533
    bins = n // 2 + 1
    assert (frames * bins * 4) % 16 == 4  # the packed layout really is misaligned
    d_pcm = torch.from_numpy(pcm).cuda()
    win = fp64.build_window(1, W)
    layouts = {
        "aligned": [torch.empty((frames, bins), device="cuda", dtype=torch.float32) for _ in range(3)],
        "packed": list(torch.empty((3, frames, bins), device="cuda", dtype=torch.float32).unbind(0)),
    }
    mags = {}
    for name, out in layouts.items():
        assert ff.stft_device(p, d_pcm, out[0], out[1], out[2], pad_mode=0) == frames
        torch.cuda.synchronize()
        re, im, mg = (o.cpu().numpy() for o in out)
        err = _every_frame_rel_l2(pcm, mg, n, W, H, frames, win)
        print(f"config 2 every frame ({name} planes): rel_l2 = {err:.3e} (tolerance {fp64.tolerance(n):.1e})",
              flush=True)
        assert err <= fp64.tolerance(n)
        assert np.array_equal(re.view(np.uint32), mg.view(np.uint32)) and not im.any()
        mags[name] = mg
    # the host entry and the aligned device launch produce the same bits
    got, hre, him, hmg = ff.stft_frames(p, pcm[: 1 << 20], pad_mode=0)
    assert got == 4093
    assert np.array_equal(hmg, mags["aligned"][:got])


def test_device_entry_matches_host_entry_bitwise(ff):
    import torch
    n = 2048
    pcm = noise_plus_sines(100000, seed=9)
    p = ff.make_plan(n, window=2048, hop=512, transform=1, store_polar=True, half_spectrum=True,
                     apply_windows=True, analysis_window=1)
    got, re, im, mg = ff.stft_frames(p, pcm, pad_mode=1)
    d = torch.from_numpy(pcm).cuda()
    out = torch.zeros((3, got, n // 2 + 1), device="cuda")
    assert ff.stft_device(p, d, out[0], out[1], None, pad_mode=1) == got
    torch.cuda.synchronize()
    assert np.array_equal(out[0].cpu().numpy(), re)
    assert np.array_equal(out[1].cpu().numpy(), im)
    assert not out[2].any()  # the skipped plane was not written


@pytest.mark.parametrize("kernel", [0, 2])
def test_complex_batched_inverse_roundtrip(ff, kernel):
    """fft_execute_complex_batched on an inverse C2R half-spectrum plan reproduces the
    framed signal (Hermitian rebuild, fft_cffi.cpp:1062-1098) and matches the restated
    reference path."""
    n, H = 512, 128
    pcm = noise_plus_sines(10 * n, seed=11)
    fwd = ff.make_plan(n, window=n, hop=H, transform=1, half_spectrum=True)
    inv = ff.make_plan(n, window=n, hop=H, inverse=1, kernel=kernel, transform=2, half_spectrum=True)
    got, re, im, _ = ff.stft_frames(fwd, pcm, pad_mode=1)
    out = np.zeros(got * n, np.float32)
    fp = lambda a: a.ctypes.data_as(ff.lib.fft_execute.argtypes[1])
    assert ff.lib.fft_execute_complex_batched(inv.handle, fp(re), fp(im), got, fp(out), 1, 0, 0) == got
    frames = np.stack([pcm[f * H:f * H + n] for f in range(got)])
    assert fp64.rel_l2(out.reshape(got, n), frames) <= 2 * fp64.tolerance(n)
    rgot, rout = ref.execute_complex_batched(re, im, n, inverse=True, kernel=kernel, transform=2, half_spectrum=True)
    assert rgot == got and fp64.rel_l2(out.reshape(got, n), rout) <= 2 * fp64.tolerance(n)


@pytest.mark.parametrize("mode", [(1, 0, 1), (0, 0, 0), (0, 1, 1), (0, 0, 1)])
@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_stft_every_variant_matches_oracle(ff, mode, n):
    """Every kernel variant the autotuner may pick (radix pattern, frames per
    workgroup, multi-frame) meets the oracle bound; variants with the same radix pattern
    agree to well inside it (different kernels may fuse multiply-adds differently, so
    bits can differ); misaligned output planes take the scalar path."""
    import torch
    reduce_mag, polar, half = mode
    W, H = n, n // 4
    pcm = noise_plus_sines(200 * H + W + 77, seed=21)
    p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, reduce_magnitude=bool(reduce_mag),
                     store_polar=bool(polar), half_spectrum=bool(half), apply_windows=True, analysis_window=1)
    frames = ff.frame_count(p, pcm.size, 1)
    bins = p.bins
    ore, oim, omg = fp64.stft_planes(pcm, n, W, H, frames, fp64.build_window(1, W), False, 1, reduce_mag, polar,
                                     half)
    d = torch.from_numpy(pcm).cuda()
    names = ff.variants(p)
    by_radix = {}
    for vi, name in enumerate(names):
        assert ff.set_variant(p, vi)
        for offset in (0, 1):
            buf = torch.zeros((3, frames * bins + offset), device="cuda")
            planes = [buf[i, offset:].view(frames, bins) for i in range(3)]
            assert ff.stft_device(p, d, *planes, pad_mode=1) == frames
            torch.cuda.synchronize()
            re, im, mg = (x.cpu().numpy() for x in planes)
            assert spectra_close(re, im, ore, oim, polar, fp64.tolerance(n)), name
            assert fp64.rel_l2(mg, omg) <= fp64.tolerance(n), name
            key = name.split(" ")[0]
            if key in by_radix:
                assert fp64.rel_l2(mg, by_radix[key]) <= fp64.tolerance(n) / 8, name
            else:
                by_radix[key] = mg
    ff.set_variant(p, -1)


@pytest.mark.parametrize("frames", [1, 7, 8, 31, 32, 33, 63, 64, 65, 127, 129, 200, 1000])
@pytest.mark.parametrize("half", [1, 0])
@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_stft_multiframe_remainders(ff, frames, half, n):
    """Frame counts around the default kernels' tile sizes (32-frame line-aligned tiles;
    whole tiles and the remainder share one launch) and a PCM shorter than one frame,
    through the device entry with separately allocated (aligned) planes, against the FP64
    oracle."""
    import torch
    W, H = n, n // 4
    L = W + (frames - 1) * H + 17 if frames > 1 else W - 100
    pcm = noise_plus_sines(L, seed=frames)
    p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, half_spectrum=bool(half), apply_windows=True,
                     analysis_window=1, pad_mode=1)
    got = ff.frame_count(p, L, 1)
    assert got == fp64.frame_count(L, n, W, H, 0, 1) == frames
    bins = n // 2 + 1 if half else n
    d = torch.from_numpy(pcm).cuda()
    out = [torch.full((got, bins), float("nan"), device="cuda") for _ in range(3)]
    assert ff.stft_device(p, d, out[0], out[1], out[2], pad_mode=1) == got
    torch.cuda.synchronize()
    re, im, mg = (o.cpu().numpy() for o in out)
    ore, oim, omg = fp64.stft_planes(pcm, n, W, H, got, fp64.build_window(1, W), False, 1, False, False, bool(half))
    assert fp64.rel_l2(re + 1j * im, ore + 1j * oim) <= fp64.tolerance(n)
    assert fp64.rel_l2(mg, omg) <= fp64.tolerance(n)


@pytest.mark.parametrize("n", [1024, 2048, 4096])
@pytest.mark.parametrize("case", ["short_window", "odd_hop", "pcm_offset", "plane_offset"])
def test_stft_default_falls_back_where_lines_do_not_fit(ff, n, case):
    """The line-aligned default needs W = N, an even hop, 8-byte aligned samples and
    16-byte aligned planes; otherwise the launch hands its frames to the tile kernel.
    Each case against the FP64 oracle, polar planes, 100+ frames."""
    import torch
    W = n - 300 if case == "short_window" else n
    H = n // 4 + 1 if case == "odd_hop" else n // 4
    L = W + 120 * H + 55
    pcm = noise_plus_sines(L + 1, seed=n + len(case))
    p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, store_polar=True, half_spectrum=True,
                     apply_windows=True, analysis_window=1, pad_mode=1)
    frames = ff.frame_count(p, L, 1)
    bins = n // 2 + 1
    big = torch.from_numpy(pcm).cuda()
    d = big[1:L + 1] if case == "pcm_offset" else big[:L]
    off = 1 if case == "plane_offset" else 0
    bufs = [torch.full((frames * bins + off,), float("nan"), device="cuda") for _ in range(3)]
    planes = [b[off:].view(frames, bins) for b in bufs]
    assert ff.stft_device(p, d, *planes, pad_mode=1) == frames
    torch.cuda.synchronize()
    re, im, mg = (x.cpu().numpy() for x in planes)
    x = d.cpu().numpy()
    ore, oim, omg = fp64.stft_planes(x, n, W, H, frames, fp64.build_window(1, W), False, 1, False, True, True)
    assert spectra_close(re, im, ore, oim, True, fp64.tolerance(n)), case
    assert fp64.rel_l2(mg, omg) <= fp64.tolerance(n), case


@pytest.mark.parametrize("mode", [(1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize("n", [2048, 4096])
def test_lines_2048_every_frame_vs_fp64(ff, mode, n):
    """The N=2048 and N=4096 line-aligned variants (16 values per lane: one wave per
    frame at N=2048, two waves with workgroup barriers between the passes at N=4096; the
    split's output ring sized to fit rather than a power of two) on 2^22 samples (8189 or
    4093 frames: whole tiles, some walked last frames first, and a rest block): every frame
    of all three planes against the FP64 oracle, and the same bits in both variants
    (window taps, or twiddles, in registers or in LDS)."""
    import torch
    reduce_mag, polar = mode
    W, H = n, n // 4
    pcm = noise_plus_sines((1 << 22) + 333, seed=31)
    p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, reduce_magnitude=bool(reduce_mag),
                     store_polar=bool(polar), half_spectrum=True, apply_windows=True, analysis_window=1, pad_mode=1)
    frames = ff.frame_count(p, pcm.size, 1)
    bins = p.bins
    d = torch.from_numpy(pcm).cuda()
    names = ff.variants(p)
    lines = [vi for vi, name in enumerate(names) if " lines" in name]
    assert len(lines) == 2, names
    ore, oim, omg = fp64.stft_planes(pcm, n, W, H, frames, fp64.build_window(1, W), False, 1, reduce_mag, polar, True)
    ref = None
    for vi in lines:
        assert ff.set_variant(p, vi)
        planes = [torch.full((frames, bins), float("nan"), device="cuda") for _ in range(3)]
        assert ff.stft_device(p, d, *planes, pad_mode=1) == frames
        torch.cuda.synchronize()
        got = [x.cpu().numpy() for x in planes]
        assert spectra_close(got[0], got[1], ore, oim, polar, fp64.tolerance(n)), names[vi]
        assert fp64.rel_l2(got[2], omg) <= fp64.tolerance(n), names[vi]
        if ref is None:
            ref = got
        else:
            for i in range(3):
                assert np.array_equal(got[i].view(np.uint32), ref[i].view(np.uint32)), (names[vi], i)
    ff.set_variant(p, -1)


@pytest.mark.parametrize("mode,samples", [((1, 0), 1 << 24), ((0, 0), 30000 * 256 + 768 + 100), ((0, 1), 26000 * 256 + 1000)])
def test_lines_variants_same_bits_at_scale(ff, mode, samples):
    """Every N=1024 line-aligned variant (tile order, stores per pass, register or LDS
    tables, workgroups per CU, small-batch tile sizes, persistent grids that walk several
    tiles per workgroup with the store pipeline running across tile edges) runs the same
    per-frame arithmetic, so at sizes where persistent workgroups take two or three tiles
    each, all three planes must equal the first variant's bit for bit; that one is checked
    against the FP64 oracle on a frame sample."""
    import torch
    reduce_mag, polar = mode
    n, W, H = 1024, 1024, 256
    pcm = noise_plus_sines(samples, seed=29)
    p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, reduce_magnitude=bool(reduce_mag),
                     store_polar=bool(polar), half_spectrum=True, apply_windows=True, analysis_window=1, pad_mode=1)
    frames = ff.frame_count(p, pcm.size, 1)
    bins = p.bins
    d = torch.from_numpy(pcm).cuda()
    names = ff.variants(p)
    lines = [vi for vi, name in enumerate(names) if " lines" in name]
    assert len(lines) >= 2 and lines[0] == 0
    ref = None
    for vi in lines:
        assert ff.set_variant(p, vi)
        planes = [torch.full((frames, bins), float("nan"), device="cuda") for _ in range(3)]
        assert ff.stft_device(p, d, *planes, pad_mode=1) == frames
        torch.cuda.synchronize()
        got = [x.cpu().numpy() for x in planes]
        if ref is None:
            ref = got
            sel = np.unique(np.concatenate([np.arange(64), np.linspace(0, frames - 1, 200).astype(int),
                                            np.arange(frames - 40, frames)]))
            win = fp64.build_window(1, W)
            x = np.concatenate([fp64.frame_signal(pcm[f * H:f * H + W], n, W, H, 1, win) for f in sel])
            ore, oim, omg = fp64.post_planes(fp64.fft64(x.astype(np.float64)), n, 1, reduce_mag, polar, True)
            assert spectra_close(got[0][sel], got[1][sel], ore, oim, polar, fp64.tolerance(n))
            assert fp64.rel_l2(got[2][sel], omg) <= fp64.tolerance(n)
        else:
            for i in range(3):
                assert np.array_equal(got[i].view(np.uint32), ref[i].view(np.uint32)), (names[vi], i)
    ff.set_variant(p, -1)


@pytest.mark.parametrize("n", [256, 1024, 4096])
def test_polar_phase_is_atan2_of_the_complex_planes(ff, n):
    """The polar planes' arg X (fast_atan2f) against atan2 of the complex planes of the
    same input and kernel: equal on the axes (signed zeros: atan2(+-0, +0) = +-0,
    atan2(+-0, -0) = +-pi, atan2(+-0, x < 0) = +-pi, atan2(y, +-0) = +-pi/2, in float)
    and within 4e-7 rad elsewhere (the bound fast_atan2f states), silent frames (X = 0)
    included; |X| equal bit for bit."""
    import torch
    W, H = n, n // 4
    pcm = noise_plus_sines(60 * H + W, seed=31)
    pcm[10 * H:20 * H + W] = 0.0  # whole frames of zeros
    d = torch.from_numpy(pcm).cuda()
    planes = {}
    for polar in (False, True):
        p = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, store_polar=polar, half_spectrum=True,
                         apply_windows=True, analysis_window=1, pad_mode=1)
        frames = ff.frame_count(p, pcm.size, 1)
        out = [torch.empty((frames, n // 2 + 1), device="cuda") for _ in range(3)]
        assert ff.stft_device(p, d, *out, pad_mode=1) == frames
        torch.cuda.synchronize()
        planes[polar] = [o.cpu().numpy() for o in out]
    re, im, mg = planes[False]
    m, ph, _ = planes[True]
    assert np.array_equal(m.view(np.uint32), mg.view(np.uint32))
    want = np.arctan2(im.astype(np.float64), re.astype(np.float64))
    axis = (re == 0) | (im == 0)
    assert axis.sum() > 0
    assert np.array_equal(ph[axis].view(np.uint32), want[axis].astype(np.float32).view(np.uint32))
    assert np.max(np.abs(ph.astype(np.float64) - want)) <= 4e-7

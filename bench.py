"""Benchmark of the batched FFT/STFT path on one MI355X.

Default workload (one "step" = one pass of the hot path over one batch): batched R2C STFT,
N = 1024, Hann window 1024, hop 256, half spectrum + magnitude, over 2^24 samples of
float32 PCM (noise + 8 sinusoids) -> 65533 frames x 513 bins, written as the three
output planes of fft_execute_batched. Inputs and outputs are resident in HBM before the
timed region; the step is one fft_execute_batched_device call (one kernel launch).

`value` is frames/s over the timed steps; `roofline` prices the kernel against HBM
(algorithmic bytes = 4*hop + 12*(N/2+1) per frame: each new input sample read once, each
of the three output planes written once); `cpu_baseline` times the oracle's restatement
of the reference CPU path on the same workload on this host's cores.

Other workloads (--workload): c2c1024 (64 x 1024-point C2C, Stockham); c2c_large
(forward + inverse round trip of 16 complex transforms of 2^22 points, or 2^--lg);
ragged (512 requests, lengths log-uniform in [2^12, 2^20], window from {256..4096}, hop
W/4, polar planes: one autotuned persistent launch per graph replay vs one plain launch
per request); stft_grad (differentiable STFT: 32 signals of 2^18 samples, window 2048, hop 512, polar
planes; forward kernel alone vs forward + adjoint kernel, and a whole torch autograd step
of the loss sum (|X| - |Y|)^2); stft_host (the default workload through the reference's host-pointer entry
fft_execute_batched: PCIe transfers of PCM and planes included, wall clock); istft (the
inverse of the default workload through fft_execute_complex_batched_device: 65533 frames x
513 complex bins -> Hermitian rebuild, 1024-point inverse, Hann synthesis window,
overlap-add with COLA normalisation into 65533*1024 output samples, zeros past the
overlap-add included, as the reference writes them); griffin_lim (fft_griffin_lim_device
on the magnitude spectrogram of 2^20 samples, N = 1024, hop 256, 4093 frames, 32
iterations; value = iterations/s from the 32-iteration call minus the 1-iteration call,
each call synchronised); phase (phase_infer_execute_device on the default workload's
|X| planes: mode 0, the linear running phase, timed with its roofline, and mode 1,
minimum phase, timed alongside); ft_pipeline (ft_apply_complex_pipeline_device on the
default workload's complex planes with the reference tool's default quantizer and 3x3
contrast, no resampling: the one-pass kernel); copy
(a plain 512 MiB device copy: the known-size kernel that calibrates the FETCH_SIZE and
WRITE_SIZE counters); write_probe (the default workload's reads and plane writes with no
transform, under several store forms and launch shapes: the floor memory sets);
large_probe (the same for the two passes of the 2^20-point C2C: their column and row
access patterns with no arithmetic); stream (fft_stream_push_pcm / push_spectrum latency
per push and throughput, host buffers).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # default: 1000 timed launches (about 85 ms) after 200 untimed ones for the headline
    # workload, whose clocks and caches are then in their steady state (200 after 20 read
    # ~1.3% slower); 200 after 20 for the others
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="stft", choices=["stft", "stft_host", "istft", "griffin_lim", "phase", "ft_pipeline",
                                                     "c2c1024", "c2c_large", "stft_grad", "ragged", "copy", "write_probe",
                                                     "large_probe", "stream"])
    ap.add_argument("--lg", type=int, default=22, help="log2 N of the c2c_large workload")
    ap.add_argument("--batch", type=int, default=64, help="transforms per launch of the c2c1024 workload")
    ap.add_argument("--n", type=int, default=1024,
                    help="transform size of the stft and c2c1024 workloads (variant studies; 1024 is the headline)")
    ap.add_argument("--polar", action="store_true", help="stft workload: polar planes instead of magnitude")
    ap.add_argument("--window", type=int, default=0,
                    help="ragged workload: every request takes this window (hop W/4) instead of the mix")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--variant", type=int, default=-1)
    args = ap.parse_args()
    headline = args.workload == "stft" and args.n == 1024 and not args.polar
    if args.steps is None:
        args.steps = 1000 if headline else 200
    if args.warmup is None:
        args.warmup = 200 if headline else 20
    return args


def read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if one exists."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_model() -> str:
    """The host CPU's model name (the CPU baseline's hardware)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def host_info() -> dict:
    return {"cpu_model": cpu_model(), "host_logical_cpus": os.cpu_count()}


def stft_rel_l2_fp64(pcm, planes, n, W, H, frames, win, reduce_magnitude, chunk=4096, polar=False):
    """Relative L2 of the device planes against the FP64 oracle over EVERY frame: framed
    float32 input (pcm*win rounded to float32, as the reference multiplies), float64 DFT,
    chunked so host memory stays small. planes = (re, im, mag) torch tensors, frames x
    bins; polar planes hold (|X|, arg X, |X|) and are compared as |X| e^(i arg X).
    Returns (complex rel-L2 of re + i im, rel-L2 of mag)."""
    import numpy as np
    from numpy.lib.stride_tricks import sliding_window_view
    bins = n // 2 + 1
    acc = np.zeros(4)
    for f0 in range(0, frames, chunk):
        f1 = min(frames, f0 + chunk)
        x = sliding_window_view(pcm[f0 * H:(f1 - 1) * H + W], W)[::H][:f1 - f0]
        x = (x * win[None, :W]).astype(np.float32) if win is not None else x.astype(np.float32)
        X = np.fft.rfft(x.astype(np.float64), n=n, axis=-1)[:, :bins]
        X[:, 0] = X[:, 0].real
        X[:, -1] = X[:, -1].real
        m = np.abs(X)
        re, im, mg = (p[f0:f1].cpu().numpy().astype(np.float64) for p in planes)
        want = m + 0j if reduce_magnitude else X
        got = re * np.exp(1j * im) if polar else re + 1j * im
        acc += [np.sum(np.abs(got - want) ** 2), np.sum(np.abs(want) ** 2),
                np.sum((mg - m) ** 2), np.sum(m ** 2)]
    return float((acc[0] / acc[1]) ** 0.5), float((acc[2] / acc[3]) ** 0.5)


def timed(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / 1000.0, wall


def cpu_baseline_stft(pcm, n, W, H, seconds):
    from oracle import fp64, ref
    taps = fp64.build_window(1, W)
    threads = ref.default_threads()
    runs = 0
    t0 = time.perf_counter()
    frames = 0
    while True:
        got, _, _, _ = ref.execute_batched(pcm, n, W, H, ctx_pad=0, pad_mode=0, kernel=2, transform=1,
                                           reduce_magnitude=True, half_spectrum=True, window_taps=taps,
                                           threads=threads)
        frames += got
        runs += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return dict({"value": frames / el, "unit": "frames/s", "cores": threads, "kind": "port",
                 "sample": f"{runs} full runs of the 2^24-sample workload (65533 frames each), Stockham code 2, "
                           f"restated reference CPU path, {threads} threads, {el:.1f} s"}, **host_info())


def stft_grid(ff, torch):
    """The headline STFT's planes (Hann 1024, hop 256, 2^24 samples) on the GPU, the input
    of the phase-inference and grid-pipeline workloads."""
    from fftfree_amd.workloads import noise_plus_sines
    n, H, L = 1024, 256, 1 << 24
    p = ff.make_plan(n, window=n, hop=H, kernel=2, transform=1, half_spectrum=True, apply_windows=True,
                     analysis_window=1, pad_mode=0)
    frames = ff.frame_count(p, L, 0)
    bins = n // 2 + 1
    re = torch.empty((frames, bins), device="cuda")
    im = torch.empty((frames, bins), device="cuda")
    mag = torch.empty((frames, bins), device="cuda")
    ff.stft_device(p, torch.from_numpy(noise_plus_sines(L, seed=2)).cuda(), re, im, mag, pad_mode=0)
    out_a, out_b = torch.empty_like(re), torch.empty_like(re)
    common = {"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
              "data": "synthetic: STFT planes (Hann 1024, hop 256) of 0.1*U[-1,1] noise + 8 sinusoids, seed 2"}
    return n, H, frames, bins, frames * bins, re, im, mag, out_a, out_b, common


def write_probe(args, torch):
    """The default workload's memory traffic with no transform: fft_probe_planes reads each
    32-frame tile's sample span (32*256 + 768 floats) and writes its region of the three
    planes (32*513 floats each) for the 2047 tiles of 65533 frames, under several store
    forms and launch shapes. The fastest full pattern is the floor memory alone sets for
    the STFT kernel; the single-plane and read-only rows show the parts."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "fftfree_amd", "libfft_probe.so"))
    lib.fft_probe_planes.restype = ctypes.c_int
    lib.fft_probe_planes.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_void_p]
    tiles, tile_hop, span4, tile4 = 2047, 32 * 256, (32 * 256 + 768) // 4, 32 * 513 // 4
    pcm = torch.rand(tiles * tile_hop + 4 * span4, device="cuda")
    planes = [torch.empty(tiles * tile4 * 4, device="cuda") for _ in range(3)]
    NT, INTER, NOR, NOW, BUF, POL = 1, 2, 4, 8, 16, 32
    cases = []
    for grid, wpc in ((768, 3), (1024, 4), (2047, 0), (512, 2)):
        for flags, what in ((BUF | NT, "buffer nt, plane after plane"), (BUF | NT | INTER, "buffer nt, planes interleaved"),
                            (NT, "global nt, plane after plane"), (0, "global, plane after plane")):
            cases.append((grid, wpc, flags, what, 3))
    # cache-policy operand of the buffer stores: bit 0 sc0, bit 1 nt, bit 4 sc1
    for grid, wpc in ((768, 3), (1024, 4), (512, 2)):
        for pol in (0, 1, 2, 3, 16, 17, 18, 19):
            cases.append((grid, wpc, BUF | INTER | POL | (pol << 8), f"buffer policy {pol}, planes interleaved", 3))
    cases.append((2047, 0, BUF | NT | NOR, "buffer nt, no reads", 3))
    cases.append((768, 3, BUF | NT | NOR, "buffer nt, no reads", 3))
    cases.append((2047, 0, BUF | NT, "buffer nt, plane 0 only", 1))
    cases.append((2047, 0, NOW, "reads only", 0))
    stream = torch.cuda.current_stream().cuda_stream
    rows = []
    for grid, wpc, flags, what, nplanes in cases:
        ptrs = [planes[i].data_ptr() if i < nplanes else None for i in range(3)]

        def step():
            rc = lib.fft_probe_planes(pcm.data_ptr(), tile_hop, span4, ptrs[0], ptrs[1], ptrs[2], tile4, tiles, grid, wpc,
                                      flags, stream)
            assert rc == 0, rc

        gpu_s, _ = timed(step, args.steps, args.warmup, torch)
        per = gpu_s / args.steps
        rd = 0 if flags & NOR else tiles * span4 * 16
        wr = 0 if flags & NOW else nplanes * tiles * tile4 * 16
        rows.append({"grid": grid, "wpc": wpc, "stores": what, "planes": nplanes, "us": per * 1e6,
                     "read_bytes": rd, "write_bytes": wr, "GBps": (rd + wr) / per / 1e9})
        print(f"probe grid={grid} wpc={wpc} {what} planes={nplanes}: {per * 1e6:.2f} us "
              f"{(rd + wr) / per / 1e9:.0f} GB/s", flush=True)
    full = [r for r in rows if r["planes"] == 3 and r["read_bytes"] > 0]
    best = min(full, key=lambda r: r["us"])
    alg = 470526940
    return {"metric": "HBM floor of the N=1024 STFT's memory pattern (probe, no transform)",
            "value": best["GBps"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": best["us"] / 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "uniform samples, constant plane words",
            "config": {"workload": "2047 tiles x (8960 samples read + 3 x 32 x 513 floats written)",
                       "best": best, "stft_algorithmic_bytes": alg,
                       "stft_floor_frac": alg / (best["us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, "probes": rows}}


def large_probe(args, torch):
    """The memory patterns of the 2^20-point two-pass transform with no arithmetic:
    fft_probe_large moves 16 transforms (128 MiB in, 128 MiB out) as pass A (columns read
    and written as runs of C values), pass B (rows read contiguously, written as runs of C)
    or a plain copy, in launches of `per` transforms on one stream or alternating over two,
    as the library runs its passes. The fastest pass-A and pass-B rows give the floor
    memory alone sets for the round trip (two passes per transform, forward and inverse)."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "fftfree_amd", "libfft_probe.so"))
    lib.fft_probe_large.restype = ctypes.c_int
    lib.fft_probe_large.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    B, n = 16, 1 << 20
    x = torch.zeros(B * n * 2, device="cuda")
    y = torch.zeros(B * n * 2, device="cuda")
    main_s = torch.cuda.current_stream()
    aux = torch.cuda.Stream()
    names = {0: "pass A (columns)", 1: "pass B (rows)", 2: "copy"}
    cases = []
    for mode in (0, 1, 2):
        for c, nts in ((8, 0), (8, 1), (16, 0), (16, 1)):
            for per, streams in ((4, 2), (4, 1), (16, 1)):
                for wpc in (0, 2):
                    cases.append((mode, c, nts, per, streams, wpc))
    rows = []
    for mode, c, nts, per, streams, wpc in cases:
        def step():
            aux.wait_stream(main_s)
            for i in range(B // per):
                s = aux if (streams == 2 and i % 2) else main_s
                off = i * per * n * 8
                rc = lib.fft_probe_large(x.data_ptr() + off, y.data_ptr() + off, per, mode, c, nts, wpc,
                                         s.cuda_stream)
                assert rc == 0, rc
            main_s.wait_stream(aux)

        gpu_s, _ = timed(step, args.steps, args.warmup, torch)
        us = gpu_s / args.steps * 1e6
        gbs = 2 * B * n * 8 / (us * 1e-6) / 1e9
        rows.append({"pattern": names[mode], "lines": c, "streaming_stores": bool(nts), "transforms_per_launch": per,
                     "streams": streams, "wpc": wpc, "us": us, "GBps": gbs})
        print(f"large probe {names[mode]} C={c} nts={nts} per={per} streams={streams} wpc={wpc}: {us:.1f} us "
              f"{gbs:.0f} GB/s", flush=True)
    best = {m: max((r for r in rows if r["pattern"] == names[m]), key=lambda r: r["GBps"]) for m in (0, 1, 2)}
    pass_bytes = 2 * B * n * 8
    floor_us = 2 * (pass_bytes / (best[0]["GBps"] * 1e9) + pass_bytes / (best[1]["GBps"] * 1e9)) * 1e6
    alg = 2 * pass_bytes
    return {"metric": "HBM floor of the 2^20 x 16 two-pass C2C's memory patterns (probe, no arithmetic)",
            "value": best[0]["GBps"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": best[0]["us"] / 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "zeros",
            "config": {"workload": "16 transforms of 2^20 complex64 (1024 x 1024), one pass per step",
                       "best_pass_a": best[0], "best_pass_b": best[1], "best_copy": best[2],
                       "round_trip_floor_us": floor_us,
                       "round_trip_floor_frac": alg / (floor_us * 1e-6) / 1e9 / HBM_PEAK_GBS, "probes": rows}}


def stream_bench(args, ff, np):
    """The streaming entries on the headline geometry (N = 1024, Hann 1024, hop 256): a
    forward context's fft_stream_push_pcm fed chunks of C samples (each push returns the
    C / 256 frames it completes: |X| plus re and im planes, host buffers), and an inverse
    context's fft_stream_push_spectrum fed F frames per push (unit-weight COLA overlap-add,
    256 samples out per frame). Every push is a synchronous host call (upload, kernel,
    download), so the wall clock per push is the latency a streaming caller sees."""
    import ctypes
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    n, W, H = 1024, 1024, 256
    bins = n // 2 + 1
    lib = ff.lib
    fwd = ff.make_plan(n, window=W, hop=H, kernel=2, radix=4, transform=1, reduce_magnitude=True,
                       half_spectrum=True, apply_windows=True, analysis_window=1)
    inv = ff.make_plan(n, window=W, hop=H, inverse=1, kernel=2, transform=2, half_spectrum=True, apply_windows=True,
                       analysis_window=1, synthesis_window=1, cola=1)
    rng = np.random.default_rng(5)
    rows = []
    for chunk in (256, 4096, 65536, 1 << 20):
        lib.fft_stream_reset(fwd.handle)
        x = (0.1 * rng.standard_normal(chunk)).astype(np.float32)
        cap = chunk // H + 4
        planes = [np.empty(cap * bins, np.float32) for _ in range(3)]
        pushes = max(8, min(2000, (1 << 23) // chunk))
        for _ in range(4 + W // chunk):  # fill the buffer past the first frame
            lib.fft_stream_push_pcm(fwd.handle, fp(x), chunk, fp(planes[0]), fp(planes[1]), fp(planes[2]), 0, 0)
        frames = 0
        t0 = time.perf_counter()
        for _ in range(pushes):
            frames += lib.fft_stream_push_pcm(fwd.handle, fp(x), chunk, fp(planes[0]), fp(planes[1]), fp(planes[2]),
                                              0, 0)
        el = time.perf_counter() - t0
        moved = pushes * chunk * 4 + frames * 3 * bins * 4
        rows.append({"entry": "fft_stream_push_pcm", "samples_per_push": chunk, "frames_per_push": frames / pushes,
                     "us_per_push": el / pushes * 1e6, "frames_per_s": frames / el, "host_GBps": moved / el / 1e9})
        print(f"stream pcm chunk={chunk}: {el / pushes * 1e6:.1f} us/push {frames / el:.3e} frames/s", flush=True)
    for fpp in (1, 16, 256, 4096):
        lib.fft_stream_reset(inv.handle)
        re = (rng.standard_normal(fpp * bins)).astype(np.float32)
        im = (rng.standard_normal(fpp * bins)).astype(np.float32)
        out = np.empty(fpp * H + n, np.float32)
        pushes = max(8, min(2000, 32768 // fpp))
        for _ in range(4):
            lib.fft_stream_push_spectrum(inv.handle, fp(re), fp(im), fpp, fp(out), 0, 0)
        samples = 0
        t0 = time.perf_counter()
        for _ in range(pushes):
            samples += lib.fft_stream_push_spectrum(inv.handle, fp(re), fp(im), fpp, fp(out), 0, 0)
        el = time.perf_counter() - t0
        moved = pushes * fpp * 2 * bins * 4 + samples * 4
        rows.append({"entry": "fft_stream_push_spectrum", "frames_per_push": fpp, "samples_per_push": samples / pushes,
                     "us_per_push": el / pushes * 1e6, "frames_per_s": pushes * fpp / el, "host_GBps": moved / el / 1e9})
        print(f"stream spectrum frames={fpp}: {el / pushes * 1e6:.1f} us/push {pushes * fpp / el:.3e} frames/s",
              flush=True)
    best = max((r for r in rows if r["entry"] == "fft_stream_push_pcm"), key=lambda r: r["frames_per_s"])
    one = next(r for r in rows if r["entry"] == "fft_stream_push_pcm" and r["samples_per_push"] == H)
    return {"metric": "streaming STFT push latency and throughput (host buffers), N=1024 hop 256",
            "value": best["frames_per_s"], "unit": "frames/s", "n_gpus": 1, "steps": None, "warmup": None,
            "ms_per_step": one["us_per_push"] / 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "0.1 * N(0,1) samples; N(0,1) spectra, seed 5",
            "config": {"workload": "fft_stream_push_pcm chunks of 256 .. 2^20 samples (magnitude + re/im planes); "
                                   "fft_stream_push_spectrum 1 .. 4096 frames (COLA overlap-add)",
                       "ms_per_step_is": "one push of one hop (256 samples -> 1 frame)", "pushes": rows},
            "roofline": {"bound": "pcie / launch latency", "achieved": None, "peak": None, "unit": None,
                         "frac": None, "traffic": None,
                         "note": "every push copies its samples in and its planes out over PCIe and waits for them; "
                                 "small pushes are bound by that round trip, large ones by the copies"}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; multi-GPU sharding is a design note, "
                                   "not a measured configuration", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from fftfree_amd import CffiFft
    from fftfree_amd.workloads import complex_normal, noise_plus_sines

    torch.cuda.set_device(0)
    ff = CffiFft()
    if args.workload == "stft":
        # N = 1024 is the headline; --n and --polar are for variant studies at other sizes
        n, L = args.n, 1 << 24
        W, H = n, n // 4
        p = ff.make_plan(n, window=W, hop=H, kernel=2, radix=4, transform=1, reduce_magnitude=not args.polar,
                         store_polar=args.polar, half_spectrum=True, apply_windows=True, analysis_window=1, pad_mode=0)
        if args.variant >= 0:
            assert ff.set_variant(p, args.variant)
        pcm = noise_plus_sines(L, seed=2)
        frames = ff.frame_count(p, L, 0)
        bins = n // 2 + 1
        d_pcm = torch.from_numpy(pcm).cuda()
        # three separate plane allocations (each 16-byte aligned, as a caller's buffers are)
        out = [torch.empty((frames, bins), device="cuda") for _ in range(3)]

        def step():
            ff.stft_device(p, d_pcm, out[0], out[1], out[2], pad_mode=0)

        gpu_s, wall = timed(step, args.steps, args.warmup, torch)
        per_launch = gpu_s / args.steps
        alg_bytes = frames * (4 * H + 12 * bins)
        achieved = alg_bytes / per_launch / 1e9
        headline = n == 1024 and not args.polar
        traffic = read_traffic("stft_r2c_n1024") if headline else None
        from oracle import fp64
        err_c, err_m = stft_rel_l2_fp64(pcm, out, n, W, H, frames, fp64.build_window(1, W), not args.polar,
                                        polar=args.polar)
        chosen = args.variant if args.variant >= 0 else ff.default_variant(p, frames)
        line = {
            "metric": ("STFT frames/s & %HBM peak at N=1024; ragged-batch and fwd+bwd vs plain, 1 GPU" if headline else
                       f"STFT frames/s at N={n}" + (" (polar planes)" if args.polar else "")),
            "value": frames * args.steps / gpu_s,
            "unit": "frames/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": per_launch * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic: 0.1*U[-1,1] noise + 8 sinusoids, seed 2",
            "config": {"workload": f"batched R2C STFT, Hann {W}, hop {H}, half spectrum + "
                                   f"{'polar' if args.polar else 'magnitude'}, 2^24 samples -> {frames} frames x "
                                   f"{bins} bins, three planes",
                       "n": n, "window": W, "hop": H, "samples": L, "frames": frames, "variant_index": chosen,
                       "variant": ff.variants(p)[chosen],
                       "rel_l2_fp64": {"planes_re_im": err_c, "plane_mag": err_m, "frames_checked": frames,
                                       "tolerance": fp64.tolerance(n)}},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                         "algorithmic_bytes_per_launch": alg_bytes},
            "wall_s": wall,
        }
        if not args.no_cpu:
            line["cpu_baseline"] = cpu_baseline_stft(pcm, n, W, H, args.cpu_seconds)
    elif args.workload == "istft":
        n, W, H, L = 1024, 1024, 256, 1 << 24
        fwd = ff.make_plan(n, window=W, hop=H, kernel=2, transform=1, half_spectrum=True, apply_windows=True,
                           analysis_window=1, pad_mode=0)
        inv = ff.make_plan(n, window=W, hop=H, inverse=1, kernel=2, transform=2, half_spectrum=True,
                           apply_windows=True, apply_ola=True, analysis_window=1, synthesis_window=1, cola=1)
        pcm = noise_plus_sines(L, seed=2)
        frames = ff.frame_count(fwd, L, 0)
        bins = n // 2 + 1
        d_pcm = torch.from_numpy(pcm).cuda()
        re = torch.empty((frames, bins), device="cuda")
        im = torch.empty((frames, bins), device="cuda")
        ff.stft_device(fwd, d_pcm, re, im, None, pad_mode=0)
        out = torch.empty(frames * n, device="cuda")

        def step():
            ff.complex_batched_device(inv, re, im, out)

        gpu_s, wall = timed(step, args.steps, args.warmup, torch)
        per_launch = gpu_s / args.steps
        span = (frames - 1) * H + n
        # interior samples (every frame present): the round trip returns the signal
        inner = slice(n, span - n)
        rt_err = float(torch.linalg.vector_norm(out[inner] - d_pcm[inner]) / torch.linalg.vector_norm(d_pcm[inner]))
        # FP64 oracle of the same inverse on frames [f0, f1): the samples whose covering
        # frames all lie in the slice, [(f0 + N/H - 1) H, f1 H), are exact there
        from oracle import fp64
        f0, f1 = frames // 2, frames // 2 + 4096
        hann = fp64.build_window(1, n)
        ref_out = fp64.complex_batched64(re[f0:f1].cpu().numpy(), im[f0:f1].cpu().numpy(), n, inverse=True,
                                         kernel=2, transform=2, half_spectrum=True, windows_enabled=True,
                                         analysis=hann, synthesis=hann, apply_ola=True, hop=H, cola_mode=1)
        lo, hi = (n // H - 1) * H, (f1 - f0) * H
        got = out[f0 * H + lo:f0 * H + hi].cpu().numpy()
        inv_err = fp64.rel_l2(got, ref_out[lo:hi])
        alg_bytes = frames * (8 * bins + 4 * n)
        achieved = alg_bytes / per_launch / 1e9
        line = {"metric": "inverse STFT frames/s (fft_execute_complex_batched, overlap-add + COLA), N=1024",
                "value": frames * args.steps / gpu_s, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": per_launch * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32",
                "data": "synthetic: planes of the forward STFT of 0.1*U[-1,1] noise + 8 sinusoids, seed 2",
                "config": {"workload": "inverse C2R half spectrum, Hann synthesis 1024, hop 256, overlap-add with "
                                       "COLA normalisation, 65533 frames x 513 bins -> 65533*1024 samples",
                           "n": n, "hop": H, "frames": frames, "interior_roundtrip_rel_l2": rt_err,
                           "rel_l2_fp64": {"output": inv_err, "frames_checked": f1 - f0,
                                           "tolerance": 2 * fp64.tolerance(n)}},
                "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic("istft_n1024"),
                             "algorithmic_bytes_per_launch": alg_bytes, "kernel": "istft_kernel"}}
    elif args.workload == "griffin_lim":
        n, H, L, iters = 1024, 256, 1 << 20, 32
        fwd = ff.make_plan(n, window=n, hop=H, kernel=0, transform=1, half_spectrum=True, pad_mode=1)
        inv = ff.make_plan(n, window=n, hop=H, inverse=1, kernel=0, transform=2, half_spectrum=True, pad_mode=1)
        pcm = torch.from_numpy(noise_plus_sines(L, seed=2)).cuda()
        frames = ff.frame_count(fwd, L, 1)
        bins = n // 2 + 1
        mag = torch.empty((frames, bins), device="cuda")
        ff.stft_device(ff.make_plan(n, window=n, hop=H, transform=1, reduce_magnitude=True, half_spectrum=True,
                                    pad_mode=1), pcm, mag, None, None, pad_mode=1)
        out_re, out_im = torch.empty_like(mag), torch.empty_like(mag)

        def call(k):
            return lambda: ff.griffin_lim_device(fwd, inv, mag, out_re, out_im, hop=H, iterations=k, seed=1)

        steps = max(1, args.steps // 20)

        def synced(fn, k):
            # each call synchronised: the host part of one call (the start's mt19937 words)
            # must not overlap the previous call's kernels
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(k):
                fn()
                torch.cuda.synchronize()
            return time.perf_counter() - t0

        t_full = synced(call(iters), steps)
        t_one = synced(call(1), steps)
        per_iter = (t_full - t_one) / steps / (iters - 1)
        ola_len = (frames - 1) * H + n
        # per iteration: inverse kernel reads |X| and the two forward planes and writes the
        # signal; the forward STFT reads the signal and writes two planes
        alg_bytes = frames * bins * 4 * 5 + ola_len * 4 * 2
        achieved = alg_bytes / per_iter / 1e9
        line = {"metric": "Griffin-Lim iterations/s (fft_griffin_lim_device), N=1024 hop 256, 4093 frames",
                "value": 1.0 / per_iter, "unit": "iterations/s", "n_gpus": 1, "steps": steps,
                "warmup": 1, "ms_per_step": t_full / steps * 1e3, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic: |STFT| of 0.1*U[-1,1] noise + 8 sinusoids, seed 2",
                "config": {"workload": "Griffin-Lim, inverse C2R + forward R2C half spectrum, no windows, pad 1, "
                                       f"{iters} iterations per call (ms_per_step), start included (the mt19937 start "
                                       "words are drawn on the host by a context's first call for a seed and "
                                       "frame count, about 0.7 ms, and reused from the device after that)",
                           "n": n, "hop": H, "frames": frames, "iterations": iters,
                           "ms_per_iteration": per_iter * 1e3, "ms_one_iteration_call": t_one / steps * 1e3,
                           "rel_l2_fp64": None,
                           "rel_l2_note": "iterative; tests/test_gpu_gl.py checks it against the oracle's loop"},
                "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                             "algorithmic_bytes_per_launch": alg_bytes,
                             "kernel": "one iteration: istft_kernel (Griffin-Lim mode) + stft_r2c_mf_kernel"}}
        if not args.no_cpu:
            from oracle import ref
            mag_h = mag.cpu().numpy()
            t0 = time.perf_counter()
            done = 0
            while True:
                ref.griffin_lim(mag_h, n, H, iterations=2, seed=1)
                done += 2
                el = time.perf_counter() - t0
                if el >= args.cpu_seconds:
                    break
            line["cpu_baseline"] = dict({"value": done / el, "unit": "iterations/s", "cores": ref.default_threads(),
                                         "kind": "port", "sample": f"{done} iterations of the restated reference "
                                                                   f"loop on the same spectrogram, {el:.1f} s"},
                                        **host_info())
    elif args.workload == "phase":
        n, H, frames, bins, cells, re, im, mag, out_a, out_b, common = stft_grid(ff, torch)
        lib = ff.lib
        handles = {m: lib.phase_init(n, H, 1, m, 0) for m in (0, 1)}
        assert all(handles.values())

        def run(m):
            return lambda: lib.phase_infer_execute_device(handles[m], mag.data_ptr(), frames, out_a.data_ptr(),
                                                          out_b.data_ptr(), torch.cuda.current_stream().cuda_stream)

        gpu_s, wall = timed(run(0), args.steps, args.warmup, torch)
        per_launch = gpu_s / args.steps
        gpu1_s, _ = timed(run(1), max(1, args.steps // 4), max(1, args.warmup // 4), torch)
        per1 = gpu1_s / max(1, args.steps // 4)
        # against the oracle's restatement of the reference loop (pinned to the literal
        # double running sum), frames [0, 4096) of the mode-0 output
        from oracle import fp64, ref
        run(0)()
        torch.cuda.synchronize()
        wre, wim = ref.phase_infer(mag[:4096].cpu().numpy(), n, H, half=True, mode=0)
        phase_err = fp64.rel_l2(out_a[:4096].cpu().numpy() + 1j * out_b[:4096].cpu().numpy(), wre + 1j * wim)
        alg_bytes = cells * (4 + 8)
        achieved = alg_bytes / per_launch / 1e9
        common.update(steps=args.steps, warmup=args.warmup)
        line = dict(common, **{
            "metric": "phase inference frames/s (phase_infer_execute_device, linear phase), N=1024 hop 256",
            "value": frames * args.steps / gpu_s, "unit": "frames/s", "ms_per_step": per_launch * 1e3,
            "config": {"workload": "phase_infer_execute mode 0 (linear running phase, double), 65533 frames x "
                                   "513 bins |X| -> (re, im); mode 1 (minimum phase: log, inverse DFT 1024, fold, "
                                   "forward DFT, exp, one launch, one wave per frame) timed alongside",
                       "n": n, "hop": H, "frames": frames, "mode1_ms_per_call": per1 * 1e3,
                       "mode1_frames_per_s": frames / per1,
                       "mode1_frac_of_hbm": alg_bytes / per1 / 1e9 / HBM_PEAK_GBS,
                       "rel_l2_oracle": {"mode0": phase_err, "frames_checked": 4096,
                                         "oracle": "oracle/cpu_ref.cpp ref_phase_linear"}},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes_per_launch": alg_bytes,
                         "kernel": "phase_linear_kernel"}})
        if not args.no_cpu:
            from oracle import ref
            sample = mag[:4096].cpu().numpy()
            t0 = time.perf_counter()
            done = 0
            while True:
                ref.phase_infer(sample, n, H, half=True, mode=0)
                done += sample.shape[0]
                el = time.perf_counter() - t0
                if el >= args.cpu_seconds:
                    break
            line["cpu_baseline"] = dict({"value": done / el, "unit": "frames/s", "cores": 1, "kind": "port",
                                         "sample": f"{done} frames (4096-frame slices of the same |X|) through the "
                                                   f"restated mode-0 loop, 1 thread (the reference's phase_infer.hpp loop runs on "
                                                   f"one thread: no worker pool), {el:.1f} s"}, **host_info())
        for h in handles.values():
            lib.phase_free(h)
    elif args.workload == "ft_pipeline":
        n, H, frames, bins, cells, re, im, mag, out_a, out_b, common = stft_grid(ff, torch)
        from fftfree_amd import ft_tools as ft
        spec = ft.grid_spec(freq_step=44100.0 / n)
        quant, contrast = ft.quantizer(), ft.contrast()

        def step():
            assert ft.apply_complex_pipeline_device(re, im, out_a, out_b, spec, None, quant, contrast)

        gpu_s, wall = timed(step, args.steps, args.warmup, torch)
        per_launch = gpu_s / args.steps
        alg_bytes = cells * 16
        achieved = alg_bytes / per_launch / 1e9
        # against the oracle on time rows [0, 2048): rows below 2047 see no slice edge
        # through the 3x3 contrast box
        from oracle import fp64, ref
        step()
        torch.cuda.synchronize()
        ok, wa, wb = ref.ft_pipeline(re[:2048].cpu().numpy(), im[:2048].cpu().numpy(), freq_step=44100.0 / n,
                                     quant=(55.0, 0.05, 0.5, 8.0), contrast=(3, 3, 0.0, 0.5, 1.0, 1))
        ft_err = fp64.rel_l2(out_a[:2047].cpu().numpy() + 1j * out_b[:2047].cpu().numpy(), wa[:2047] + 1j * wb[:2047])
        common.update(steps=args.steps, warmup=args.warmup)
        line = dict(common, **{
            "metric": "grid pipeline frames/s (ft_apply_complex_pipeline_device: harmonic quantizer + 3x3 local "
                      "contrast), 513 bins",
            "value": frames * args.steps / gpu_s, "unit": "frames/s", "ms_per_step": per_launch * 1e3,
            "config": {"workload": "ft pipeline on the complex STFT grid, 65533 time bins x 513 freq bins, "
                                   "quantizer 55 Hz / step 0.05 / mix 0.5 / decay 8, contrast 3x3 gain 0.5, "
                                   "freq step 44100/1024, no resampling (one pass)",
                       "frames": frames, "bins": bins, "cells": cells,
                       "rel_l2_oracle": {"planes": ft_err, "time_rows_checked": 2047,
                                         "oracle": "oracle/cpu_ref.cpp ref_ft_pipeline"}},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic("ft_fused_65533x513"),
                         "algorithmic_bytes_per_launch": alg_bytes, "kernel": "ft_fused_kernel<false>"}})
        if not args.no_cpu:
            from oracle import ref
            ra, rb = re[:2048].cpu().numpy(), im[:2048].cpu().numpy()
            t0 = time.perf_counter()
            done = 0
            while True:
                ref.ft_pipeline(ra, rb, freq_step=44100.0 / n, quant=(55.0, 0.05, 0.5, 8.0),
                                contrast=(3, 3, 0.0, 0.5, 1.0, 1))
                done += ra.shape[0]
                el = time.perf_counter() - t0
                if el >= args.cpu_seconds:
                    break
            line["cpu_baseline"] = dict({"value": done / el, "unit": "frames/s", "cores": 1, "kind": "port",
                                         "sample": f"{done} frames (2048 x 513 slices of the same grid) through the "
                                                   f"restated pipeline, 1 thread (the reference's ft_tools.hpp pipeline runs on one "
                                                   f"thread: no worker pool), {el:.1f} s"}, **host_info())
    elif args.workload == "c2c_large":
        n, B = 1 << args.lg, 16
        fwd = ff.make_plan(n, kernel=2, radix=4, transform=0, half_spectrum=False)
        inv = ff.make_plan(n, inverse=1, kernel=2, radix=4, transform=0, half_spectrum=False)
        if args.variant >= 0:
            assert ff.set_variant(fwd, args.variant) and ff.set_variant(inv, args.variant)
        x = torch.from_numpy(complex_normal(B, n, seed=1)).cuda()
        y = torch.empty_like(x)
        z = torch.empty_like(x)

        def step():
            ff.c2c_device(fwd, x, y)
            ff.c2c_device(inv, y, z)

        gpu_s, wall = timed(step, args.steps, args.warmup, torch)
        per_step = gpu_s / args.steps
        torch.cuda.synchronize()
        rt_err = float(torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x))
        # forward vs the FP64 DFT, every transform
        xs, ys = x.cpu().numpy(), y.cpu().numpy()
        num = den = 0.0
        for b in range(B):
            want = np.fft.fft(xs[b].astype(np.complex128))
            num += float(np.sum(np.abs(ys[b] - want) ** 2))
            den += float(np.sum(np.abs(want) ** 2))
        fwd_err = (num / den) ** 0.5
        from oracle import fp64
        # algorithmic bytes: each transform reads and writes B*N complex64 once; the two
        # passes per transform move twice that, which is not algorithmic (a two-pass
        # design caps at 50% of HBM peak)
        pass_bytes = 2 * 8 * n * B
        alg_bytes = 2 * pass_bytes  # forward + inverse per step
        achieved = alg_bytes / per_step / 1e9
        line = {"metric": f"C2C round trips/s, N=2^{args.lg} batch 16", "value": B * args.steps / gpu_s,
                "unit": "round trips/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": per_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32", "data": "synthetic complex64 N(0,1), seed 1",
                "config": {"workload": f"C2C forward + inverse N=2^{args.lg} x 16, two LDS passes per transform",
                           "n": n, "batch": B, "roundtrip_rel_l2": rt_err,
                           "rel_l2_fp64": {"forward": fwd_err, "transforms_checked": B,
                                           "tolerance": fp64.tolerance(n)},
                           "variant": ff.variants(fwd)[ff.default_variant(fwd, B)]},
                "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(f"c2c_large_2e{args.lg}"),
                             "algorithmic_bytes_per_launch": pass_bytes,
                             "algorithmic_bytes_per_step": alg_bytes,
                             "pass_traffic_GBs": 4 * pass_bytes / per_step / 1e9,
                             "note": "per step: forward + inverse, each a pass A (columns) and a pass B (rows) "
                                     "launch per chunk of ~32 MiB, chunks alternating over two streams; frac counts "
                                     "each transform's compulsory read + write once (cap 0.5 for two passes); "
                                     "pass_traffic_GBs is the rate of all four passes' reads and writes"}}
    elif args.workload == "stft_host":
        n, W, H, L = 1024, 1024, 256, 1 << 24
        p = ff.make_plan(n, window=W, hop=H, kernel=2, radix=4, transform=1, reduce_magnitude=True,
                         half_spectrum=True, apply_windows=True, analysis_window=1, pad_mode=0)
        pcm = noise_plus_sines(L, seed=2)
        frames = ff.frame_count(p, L, 0)
        bins = n // 2 + 1
        planes = np.empty((3, frames * bins), np.float32)
        fp = lambda a: a.ctypes.data_as(ff.lib.fft_execute.argtypes[1])

        def step():
            got = ff.lib.fft_execute_batched(p.handle, fp(pcm), L, fp(planes[0]), fp(planes[1]), fp(planes[2]), 0, 0,
                                             frames)
            assert got == frames

        for _ in range(args.warmup):
            step()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        wall = time.perf_counter() - t0
        per = wall / args.steps
        host_bytes = 4 * L + 3 * 4 * frames * bins
        # the same planes from the device entry (one launch over all frames), and every
        # frame against the FP64 oracle
        dev = [torch.empty((frames, bins), device="cuda") for _ in range(3)]
        ff.stft_device(p, torch.from_numpy(pcm).cuda(), dev[0], dev[1], dev[2], pad_mode=0)
        torch.cuda.synchronize()
        # the chunk launches may pick another kernel variant than one launch over all frames
        diff = max(float(np.max(np.abs(planes[i] - dev[i].cpu().numpy().reshape(-1)))) for i in range(3))
        scale = max(float(np.max(np.abs(planes[i]))) for i in range(3))
        from oracle import fp64
        rel = stft_rel_l2_fp64(pcm, [torch.from_numpy(planes[i].reshape(frames, bins)) for i in range(3)], n, W, H,
                               frames, fp64.build_window(1, W), True)
        line = {"metric": "STFT frames/s through the host-pointer entry (PCIe included)",
                "value": frames * args.steps / wall, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": per * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32", "data": "synthetic: 0.1*U[-1,1] noise + 8 sinusoids, seed 2",
                "config": {"workload": "fft_execute_batched on pageable host buffers: 64 MiB PCM in, "
                                       "3 x 134 MB planes out, N=1024 Hann hop 256 half + magnitude",
                           "host_GBs": host_bytes / per / 1e9, "host_bytes_per_call": host_bytes,
                           "max_abs_diff_vs_device_entry_rel": diff / scale,
                           "rel_l2_fp64": {"complex": rel[0], "magnitude": rel[1], "frames_checked": frames},
                           "pipeline": "pinned bounce slots (3 out, 2 in), ~24 MiB of planes per chunk, upload / "
                                       "kernel / download streams, host copies on a thread pool"}}
    elif args.workload == "stft_grad":
        from fftfree_amd.autograd import magnitude_loss, polar_plan
        n, W, H, B, L = 2048, 2048, 512, 32, 1 << 18
        p = polar_plan(n, W, H)
        x = torch.from_numpy(np.stack([noise_plus_sines(L, seed=1000 + i) for i in range(B)])).cuda()
        y = torch.from_numpy(np.stack([noise_plus_sines(L, seed=2000 + i) for i in range(B)])).cuda()
        frames = ff.frame_count(p, L, 0)
        bins = n // 2 + 1
        mag = torch.empty((B, frames, bins), device="cuda")
        phase = torch.empty_like(mag)
        ymag = torch.empty_like(mag)
        ff.stft_signals_device(p, y, ymag, None, None)
        gmag = torch.empty_like(mag)
        gx = torch.empty_like(x)

        def fwd():
            ff.stft_signals_device(p, x, mag, phase, None)

        def adj():  # the autograd backward: from the forward's saved planes
            ff.stft_adjoint_saved_device(p, L, mag, phase, gmag, None, gx)

        def adj_recompute():  # from the samples, recomputing each frame's spectrum
            ff.stft_adjoint_device(p, x, gmag, None, gx)

        fwd_s, _ = timed(fwd, args.steps, args.warmup, torch)
        torch.mul(mag - ymag, 2.0, out=gmag)
        adj_s, _ = timed(adj, args.steps, args.warmup, torch)
        rec_s, _ = timed(adj_recompute, args.steps, args.warmup, torch)
        xt = x.clone().requires_grad_(True)

        def train_step():
            xt.grad = None
            magnitude_loss(xt, ymag, n, W, H).backward()

        step_s, wall = timed(train_step, args.steps, args.warmup, torch)
        fwd_ms, adj_ms, rec_ms = fwd_s / args.steps * 1e3, adj_s / args.steps * 1e3, rec_s / args.steps * 1e3
        fwd_bytes = B * L * 4 + B * frames * bins * 8
        adj_bytes = B * frames * bins * 4 + B * L * 8
        # what the saved-plane adjoint reads and writes: the gradient plane, the saved |X| and
        # arg X, and dL/dx
        saved_bytes = B * frames * bins * 12 + B * L * 4
        # FP64: forward planes and the gradient of the timed adjoint call, two signals
        from oracle import fp64
        win = fp64.build_window(1, W)
        torch.mul(mag - ymag, 2.0, out=gmag)
        adj()
        torch.cuda.synchronize()
        fe = [0.0, 0.0, 0.0, 0.0]
        for sgl in (0, B - 1):
            xs = x[sgl].cpu().numpy()
            A, Bp = fp64.value_planes64(xs, n, W, H, frames, win, "polar")
            got_m, got_p = mag[sgl].cpu().numpy(), phase[sgl].cpu().numpy()
            fe[0] += float(np.sum(np.abs(got_m * np.exp(1j * got_p) - A * np.exp(1j * Bp)) ** 2))
            fe[1] += float(np.sum(A ** 2))
            want = fp64.stft_adjoint64(xs, n, W, H, frames, win, "magnitude", gmag[sgl].cpu().numpy())
            fe[2] += float(np.sum((gx[sgl].cpu().numpy() - want) ** 2))
            fe[3] += float(np.sum(want ** 2))
        fwd_err, grad_err = (fe[0] / fe[1]) ** 0.5, (fe[2] / fe[3]) ** 0.5
        line = {"metric": "differentiable STFT: (fwd + adjoint) / fwd time, and signals/s", "value": B / (step_s / args.steps),
                "unit": "signals/s (whole autograd step)", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": step_s / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32", "data": "synthetic: noise + 8 sinusoids, seeds 1000.. and 2000..",
                "config": {"workload": "32 signals x 2^18 samples, Hann 2048, hop 512, polar planes, 509 frames x 1025 bins",
                           "fwd_kernel_ms": fwd_ms, "adjoint_kernel_ms": adj_ms,
                           "adjoint": "fft_stft_adjoint_saved_device (the autograd backward)",
                           "fwd_plus_adjoint_over_fwd": (fwd_ms + adj_ms) / fwd_ms,
                           "adjoint_recompute_kernel_ms": rec_ms,
                           "fwd_plus_adjoint_recompute_over_fwd": (fwd_ms + rec_ms) / fwd_ms,
                           "rel_l2_fp64": {"forward_polar": fwd_err, "gradient": grad_err, "signals_checked": 2,
                                           "tolerance_forward": fp64.tolerance(n),
                                           "tolerance_gradient": 4 * fp64.tolerance(n)}},
                "roofline": {"bound": "hbm", "achieved": saved_bytes / (adj_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": saved_bytes / (adj_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "traffic": read_traffic("stft_grad_adjoint_n2048"), "algorithmic_bytes_per_launch": saved_bytes,
                             "kernel": "stft_adjoint_kernel from saved planes + stft_halo_kernel",
                             "recompute": {"achieved": adj_bytes / (rec_ms * 1e-3) / 1e9,
                                           "frac": adj_bytes / (rec_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                           "algorithmic_bytes_per_launch": adj_bytes},
                             "forward": {"achieved": fwd_bytes / (fwd_ms * 1e-3) / 1e9,
                                         "frac": fwd_bytes / (fwd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                         "algorithmic_bytes_per_launch": fwd_bytes,
                                         "traffic": read_traffic("stft_grad_forward_n2048")}}}
    elif args.workload == "ragged":
        from fftfree_amd.cffi_fft import RaggedBatch
        from fftfree_amd.workloads import ragged_requests
        reqs = ragged_requests(512, seed=5)
        if args.window:  # one window size: the first 160 lengths of the same distribution
            reqs = [(L, args.window, args.window // 4) for L, W, H in reqs[:160]]
        pcms = [torch.from_numpy(noise_plus_sines(L, seed=5000 + i)).cuda() for i, (L, W, H) in enumerate(reqs)]
        rb = RaggedBatch(ff, [(pcm, W, H) for pcm, (L, W, H) in zip(pcms, reqs)])
        t0 = time.perf_counter()
        rb.autotune(3)
        tune_s = time.perf_counter() - t0
        frames = sum(rb.frames)
        alg_bytes = sum(4 * L + 12 * f * (W // 2 + 1) for (L, W, H), f in zip(reqs, rb.frames))
        graph_s, wall = timed(rb.execute, args.steps, args.warmup, torch)
        plain_s, _ = timed(rb.execute_plain, args.steps, args.warmup, torch)
        per, per_plain = graph_s / args.steps, plain_s / args.steps
        # FP64 of every 8th request's polar planes, all their frames (graph output)
        from oracle import fp64
        rb.execute()
        torch.cuda.synchronize()
        acc = [0.0, 0.0]
        worst = 0.0
        for i in range(0, len(reqs), 8):
            L, W, H = reqs[i]
            f = rb.frames[i]
            mg, ph = (t[:f].cpu().numpy() for t in (rb.outputs[i][0], rb.outputs[i][1]))
            A, Bp = fp64.value_planes64(pcms[i].cpu().numpy(), W, W, H, f, fp64.build_window(1, W), "polar")
            e = float(np.sum(np.abs(mg * np.exp(1j * ph) - A * np.exp(1j * Bp)) ** 2))
            d = float(np.sum(A ** 2))
            acc[0] += e
            acc[1] += d
            worst = max(worst, (e / d) ** 0.5 / fp64.tolerance(W))
        ragged_err = (acc[0] / acc[1]) ** 0.5
        line = {"metric": "ragged batched STFT: frames/s in one graph launch, speedup vs per-request launches",
                "value": frames / per, "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": per * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32", "data": "synthetic: noise + 8 sinusoids per request, seeds 5000..",
                "config": {"workload": (f"160 requests, L log-uniform [2^12, 2^20], W = {args.window}, hop W/4, "
                                        if args.window else
                                        "512 requests, L log-uniform [2^12, 2^20], W in {256..4096}, hop W/4, ")
                                       + "Hann, R2C half spectrum, polar planes, pad_mode 1",
                           "frames": frames, "samples": sum(L for L, W, H in reqs),
                           "plain_ms_per_step": per_plain * 1e3, "speedup_vs_plain": per_plain / per,
                           "autotune_s": tune_s, "plan": rb.describe(),
                           "rel_l2_fp64": {"planes_polar": ragged_err, "requests_checked": len(range(0, len(reqs), 8)),
                                           "worst_request_over_tolerance": worst,
                                           "tolerance": "2e-7 * log2(N) per request"}},
                "roofline": {"bound": "hbm", "achieved": alg_bytes / per / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": alg_bytes / per / 1e9 / HBM_PEAK_GBS,
                             "traffic": None if args.window else read_traffic("stft_ragged_512"),
                             "algorithmic_bytes_per_launch": alg_bytes,
                             "kernel": "one graph: a stft_ragged_lines_kernel or stft_ragged_tile_kernel launch "
                                       "per window size"}}
        rb.close()
    elif args.workload == "write_probe":
        line = write_probe(args, torch)
    elif args.workload == "large_probe":
        line = large_probe(args, torch)
    elif args.workload == "stream":
        line = stream_bench(args, ff, np)
    elif args.workload == "copy":
        nbytes = 512 << 20
        x = torch.ones(nbytes // 4, device="cuda")
        y = torch.empty_like(x)

        def step():
            torch.mul(x, 1.0, out=y)

        gpu_s, wall = timed(step, args.steps, args.warmup, torch)
        per = gpu_s / args.steps
        line = {"metric": "device copy GB/s (counter calibration)", "value": 2 * nbytes / per / 1e9, "unit": "GB/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": per * 1e3,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "ones",
                "config": {"workload": "elementwise copy, 512 MiB read + 512 MiB written per launch"}}
    else:
        n, B = args.n, args.batch
        p = ff.make_plan(n, kernel=2, radix=4, transform=0, half_spectrum=False)
        if args.variant >= 0:
            assert ff.set_variant(p, args.variant)
        x = torch.from_numpy(complex_normal(B, n, seed=1)).cuda()
        y = torch.empty_like(x)

        def step():
            ff.c2c_device(p, x, y)

        gpu_s, wall = timed(step, args.steps, args.warmup, torch)
        per_launch = gpu_s / args.steps
        alg_bytes = 2 * 8 * n * B
        achieved = alg_bytes / per_launch / 1e9
        from oracle import fp64
        fwd_err = fp64.rel_l2(y.cpu().numpy(), fp64.fft64(x.cpu().numpy()))
        line = {"metric": f"C2C transforms/s, N={n} batch {B}", "value": B * args.steps / gpu_s,
                "unit": "transforms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": per_launch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32", "data": "synthetic complex64 N(0,1), seed 1",
                "config": {"workload": f"C2C forward N={n} x {B}, Stockham",
                           "variant": ff.variants(p)[args.variant if args.variant >= 0 else ff.default_variant(p, B)],
                           "rel_l2_fp64": {"forward": fwd_err, "transforms_checked": B,
                                           "tolerance": fp64.tolerance(n)}},
                "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(f"c2c_n{n}_b{B}"),
                             "algorithmic_bytes_per_launch": alg_bytes, "kernel": "c2c_kernel"}}
    print(json.dumps(line), flush=True)
    ff.close()


if __name__ == "__main__":
    main()

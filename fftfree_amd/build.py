"""Builds libfft_cffi.so (gfx950) in-tree with hipcc.

Compile flags: -O3, gfx950 only, and packed FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops) for every HIP translation unit.
"""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libfft_cffi.so"
# bandwidth probes for bench.py's write_probe workload (measurement only, not the library)
PROBE_SRC = "probe.hip"
PROBE_LIB = PKG / "libfft_probe.so"
# the slowest to compile first (the pool starts them first)
SOURCES = ["fft_stft_lines.hip", "fft_stft_lines_exch.hip", "fft_stft_lines_order.hip",
           "fft_stft_lines_2048.hip", "fft_stft_lines_4096.hip",
           "fft_stft_mf.hip", "fft_stft.hip", "fft_ragged.hip", "fft_adjoint.hip", "fft_large.hip", "fft_kernels.hip",
           "fft_istft.hip", "fft_phase.hip", "ft_tools.hip", "plan.cpp", "fft_cffi.cpp", "ft_cffi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
HIP_ONLY = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: Path, obj: Path, verbose: bool) -> None:
    cmd = [HIPCC, *COMMON]
    if src.suffix == ".hip":
        cmd += HIP_ONLY
    cmd += ["-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature; that is harmless
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{msg}")
    if msg.strip() and verbose:
        print(msg)


def build(verbose: bool = False) -> Path:
    out = PKG / "_build"
    out.mkdir(exist_ok=True)
    objs = [out / (name + ".o") for name in SOURCES]
    probe_obj = out / (PROBE_SRC + ".o")
    # translation units are independent: compile them side by side
    with ThreadPoolExecutor(max_workers=min(len(SOURCES) + 1, os.cpu_count() or 1)) as pool:
        jobs = [pool.submit(_compile, CSRC / name, obj, verbose) for name, obj in zip(SOURCES, objs)]
        jobs.append(pool.submit(_compile, CSRC / PROBE_SRC, probe_obj, verbose))
        for f in jobs:
            f.result()
    _link([str(probe_obj)], PROBE_LIB, verbose)
    _link([str(o) for o in objs], LIB, verbose)
    return LIB


def _link(objs, lib: Path, verbose: bool) -> None:
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    tmp.replace(lib)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

"""Builds tests/abi/caller: the C++ caller of include/fft_cffi.h (test infrastructure).

g++ compiles caller.cpp against the header and links it to fftfree_amd/libfft_cffi.so
with an $ORIGIN-relative run path, so the binary finds the library wherever the tree is.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
SRC = HERE / "caller.cpp"
OUT = HERE / "caller"


def command(out: Path = OUT):
    return ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"), str(SRC),
            "-L", str(ROOT / "fftfree_amd"), "-lfft_cffi", "-Wl,-rpath,$ORIGIN/../../fftfree_amd", "-o", str(out)]


def build(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

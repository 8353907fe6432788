"""Build recipe of the oracle's C++ restatement (libcpu_ref.so). TEST INFRASTRUCTURE ONLY.

The reference's own compile flags are -O3 -ffast-math -fno-math-errno -march=native
(CMakeLists.txt:16, 225, 237-239). -march=native is replaced by x86-64-v3 (AVX2/FMA)
because this library is built once here and then run on other hosts' CPUs.

The reference library itself is not built: it needs Eigen 3.4.0 fetched by CMake
FetchContent (CMakeLists.txt:105-111), which is not available offline. One part of the
path is header-only with standard includes: the grid pipeline's templates in
ft_tools.hpp. When the reference tree is present, build_ref() compiles that header, from
where it lies, with ref_ft_driver.cpp into oracle/_ref/libft_ref.so, with the flags above.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_ref.cpp"
OUT = HERE / "libcpu_ref.so"
FLAGS = ["-O3", "-ffast-math", "-fno-math-errno", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-pthread"]


REFERENCE = Path("/root/reference")
REF_OUT = HERE / "_ref" / "libft_ref.so"


def build_ref(verbose: bool = False):
    """oracle/_ref/libft_ref.so from the reference's ft_tools.hpp, or None without the tree."""
    if not (REFERENCE / "ft_tools.hpp").exists():
        return None
    REF_OUT.parent.mkdir(exist_ok=True)
    tmp = REF_OUT.with_suffix(".so.tmp")
    cmd = ["g++", *FLAGS, "-I", str(REFERENCE), str(HERE / "ref_ft_driver.cpp"), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    tmp.replace(REF_OUT)
    return REF_OUT


def build(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["g++", *FLAGS, str(SRC), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    tmp.replace(OUT)
    build_ref(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

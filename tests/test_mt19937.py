"""The library's block Mersenne Twister (fftfree_amd/csrc/mt19937.hpp, the Griffin-Lim
start) against std::mt19937 and numpy's MT19937: the same 32-bit outputs for several
seeds, across chunk boundaries of the fill calls. CPU only (compiled with g++ here)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = r"""
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "mt19937.hpp"
int main(int argc, char** argv) {
  const unsigned seed = (unsigned)std::strtoul(argv[1], nullptr, 10);
  const size_t n = 100000;
  std::vector<uint32_t> a(n), b(n);
  fftfree::Mt19937Block blk(seed);
  size_t done = 0, step = 1;
  while (done < n) {  // uneven chunks: 1, 2, 4, ... words
    const size_t take = std::min(step, n - done);
    blk.fill(a.data() + done, take);
    done += take;
    step = step * 2 + 1;
  }
  std::mt19937 ref(seed);
  for (size_t i = 0; i < n; ++i) b[i] = (uint32_t)ref();
  for (size_t i = 0; i < n; ++i)
    if (a[i] != b[i]) { std::printf("mismatch %zu\n", i); return 1; }
  std::fwrite(a.data(), 4, n, stdout);
  return 0;
}
"""


@pytest.fixture(scope="module")
def prog():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "mt.cpp")
    exe = os.path.join(d, "mt")
    with open(src, "w") as f:
        f.write(PROG)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fftfree_amd", "csrc"), src, "-o", exe],
                   check=True)
    return exe


@pytest.mark.parametrize("seed", [0x47d7bfad, 1, 99, 4294967295])
def test_block_mt19937_matches_std_and_numpy(prog, seed):
    out = subprocess.run([prog, str(seed)], capture_output=True, check=True).stdout
    words = np.frombuffer(out, dtype=np.uint32)
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    assert np.array_equal(words, bg.random_raw(words.size).astype(np.uint32))

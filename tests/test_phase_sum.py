"""The closed-form running phase (fftfree_amd/csrc/phase_sum.hpp) against the literal
double loop phi += d, value for value. CPU only (compiled with g++ without fast-math,
the same arithmetic the GPU's double units perform).

Cases: the linear-phase increments 2 pi kk hop / N of the reference's phase inference
(phase_infer.hpp:47-51) for several N and hops, every frame count up to 300 and sampled
counts up to 70000; increments with few significant bits, where d / u falls half-way
between integers and round-half-even depends on phi (d = k * 2^-j, 0.5, 3, 0.1, ...);
increments below half a unit of the sum (phi stops moving); negative increments; long
runs to 2^21 steps. The segment walk (running_segment) that the phase kernel uses to
rotate (cos, sin) instead of evaluating them per frame visits exactly the same sums,
and its rotated values stay within 1e-12 of cos/sin of them."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = r"""
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include "phase_sum.hpp"

static long checked = 0;

// compare the closed form with the loop at every count in `at` (ascending)
static int check(double d, const std::vector<uint64_t>& at) {
  double phi = 0.0;
  uint64_t k = 0;
  for (uint64_t f : at) {
    while (k < f) { phi += d; ++k; }
    const double got = fftfree::running_sum(d, f);
    ++checked;
    if (std::memcmp(&got, &phi, sizeof(double)) != 0) {
      std::printf("MISMATCH d=%.17g f=%llu got=%.17g want=%.17g\n", d, (unsigned long long)f, got, phi);
      return 1;
    }
  }
  return 0;
}

int main() {
  std::mt19937_64 rng(12345);
  std::vector<uint64_t> counts;
  for (uint64_t f = 0; f <= 300; ++f) counts.push_back(f);
  for (uint64_t f = 301; f < 70000; f += 1 + rng() % 997) counts.push_back(f);
  counts.push_back(65533);
  counts.push_back(70000);
  std::sort(counts.begin(), counts.end());
  counts.erase(std::unique(counts.begin(), counts.end()), counts.end());
  int bad = 0;
  // linear-phase increments
  const int Ns[] = {256, 512, 1000, 1024, 2048, 4096};
  for (int N : Ns) {
    const int hops[] = {1, 64, 100, N / 4, N / 2, N};
    for (int hop : hops) {
      for (int kk = 0; kk <= N / 2; kk += (kk < 40 ? 1 : 1 + (int)(rng() % 97))) {
        const double d = 6.28318530717958647692 * (double)kk * (double)hop / (double)N;
        bad |= check(d, counts);
      }
    }
  }
  // few significant bits: frequent half-way cases
  for (int j = 0; j < 40; ++j)
    for (int k = 1; k < 40; k += 2) bad |= check(std::ldexp((double)k, -j), counts);
  const double misc[] = {0.5, 1.0, 3.0, 0.1, 0.2, 0.3, 1e-300, 1e-17, 5e-324, 1e300, 7.0 / 3.0,
                         3.141592653589793, 1.5707963267948966, 6.283185307179586, -0.1, -2.5, -3.141592653589793};
  for (double d : misc) bad |= check(d, counts);
  // random increments over many exponents
  std::uniform_real_distribution<double> mant(1.0, 2.0);
  for (int i = 0; i < 400; ++i) {
    const double d = std::ldexp(mant(rng), (int)(rng() % 80) - 40) * ((rng() & 1) ? 1.0 : -1.0);
    bad |= check(d, counts);
  }
  // long runs
  std::vector<uint64_t> lng;
  for (uint64_t f = 1; f <= (1u << 21); f = f * 3 + 1) lng.push_back(f);
  lng.push_back(1u << 21);
  const double long_d[] = {0.1, 2.4543692606170259, 804.24771931898704, std::ldexp(3.0, -7)};
  for (double d : long_d) bad |= check(d, lng);
  // segments: the kernel's walk over a block of frames (segment, rotate, literal step)
  // must visit exactly the running sums, and the rotated (cos, sin) stay within 1e-12
  long seg_checked = 0;
  auto walk = [&](double d, uint64_t f0, int count) {
    double phi = fftfree::running_sum(d, f0);
    double want = phi;
    int i = 0;
    double worst = 0.0;
    while (i < count) {
      double step; uint64_t run;
      fftfree::running_segment(phi, d, &step, &run);
      double c = std::cos(phi), s = std::sin(phi);
      const double cd = std::cos(step), sd = std::sin(step);
      for (;;) {
        if (std::memcmp(&phi, &want, sizeof(double)) != 0) {
          std::printf("SEGMENT d=%.17g f=%llu got=%.17g want=%.17g\n", d, (unsigned long long)(f0 + i), phi, want);
          return 1;
        }
        worst = std::fmax(worst, std::fmax(std::fabs(c - std::cos(phi)), std::fabs(s - std::sin(phi))));
        ++seg_checked;
        ++i;
        want += d;  // the literal next running sum
        if (i >= count) break;
        if (run == 0) { phi = phi + d; break; }
        const double c2 = c * cd - s * sd;
        s = s * cd + c * sd;
        c = c2;
        phi += step;
        --run;
      }
    }
    if (worst > 1e-12) { std::printf("ROTATION d=%.17g f0=%llu err=%.3g\n", d, (unsigned long long)f0, worst); return 1; }
    return 0;
  };
  for (int N : Ns) {
    const int hops[] = {1, 100, N / 4, N};
    for (int hop : hops)
      for (int kk = 0; kk <= N / 2; kk += 1 + (int)(rng() % 61)) {
        const double d = 6.28318530717958647692 * (double)kk * (double)hop / (double)N;
        for (uint64_t f0 = 0; f0 < 70000; f0 += 64 * (1 + rng() % 40)) bad |= walk(d, f0, 64);
      }
  }
  for (int j = 0; j < 30; ++j)
    for (int k = 1; k < 20; k += 2) {
      const double d = std::ldexp((double)k, -j);
      for (uint64_t f0 = 0; f0 < 5000; f0 += 64) bad |= walk(d, f0, 64);
    }
  std::printf("checked %ld segments-checked %ld\n", checked, seg_checked);
  return bad;
}
"""


@pytest.fixture(scope="module")
def prog():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "t.cpp")
        exe = os.path.join(tmp, "t")
        with open(src, "w") as fh:
            fh.write(PROG.replace("#include <vector>", "#include <vector>\n#include <algorithm>"))
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fftfree_amd", "csrc"), src, "-o", exe],
                       check=True)
        yield exe


def test_running_sum_matches_loop(prog):
    r = subprocess.run([prog], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "checked" in r.stdout

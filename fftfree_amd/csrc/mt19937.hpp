// mt19937.hpp -- the 32-bit Mersenne Twister (the generator std::mt19937 names), producing
// whole 624-word twists at a time so that the twist and the tempering run as vector
// loops. Output i equals the i-th value of std::mt19937 with the same seed (init_genrand
// seeding, Matsumoto and Nishimura 1998).
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace fftfree {

class Mt19937Block {
 public:
  explicit Mt19937Block(uint32_t seed) {
    mt_[0] = seed;
    for (uint32_t i = 1; i < kN; ++i) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + i;
  }

  // The next n outputs into out.
  void fill(uint32_t* out, size_t n) {
    size_t done = 0;
    while (done < n) {
      if (pos_ == kN) {
        twist();
        pos_ = 0;
      }
      size_t take = kN - pos_;
      if (take > n - done) take = n - done;
      temper(out + done, pos_, take);
      pos_ += take;
      done += take;
    }
  }

 private:
  static constexpr uint32_t kN = 624, kM = 397;
  static constexpr uint32_t kUpper = 0x80000000u, kLower = 0x7fffffffu, kA = 0x9908b0dfu;

  static uint32_t step(uint32_t cur, uint32_t next, uint32_t far) {
    const uint32_t y = (cur & kUpper) | (next & kLower);
    return far ^ (y >> 1) ^ ((0u - (y & 1u)) & kA);
  }

  void twist() {
    for (uint32_t i = 0; i < kN - kM; ++i) mt_[i] = step(mt_[i], mt_[i + 1], mt_[i + kM]);
    for (uint32_t i = kN - kM; i < kN - 1; ++i) mt_[i] = step(mt_[i], mt_[i + 1], mt_[i + kM - kN]);
    mt_[kN - 1] = step(mt_[kN - 1], mt_[0], mt_[kM - 1]);
  }

  void temper(uint32_t* out, uint32_t from, size_t count) const {
    for (size_t j = 0; j < count; ++j) {
      uint32_t y = mt_[from + j];
      y ^= y >> 11;
      y ^= (y << 7) & 0x9d2c5680u;
      y ^= (y << 15) & 0xefc60000u;
      y ^= y >> 18;
      out[j] = y;
    }
  }

  uint32_t mt_[kN];
  uint32_t pos_ = kN;
};

}  // namespace fftfree

// fft_stft_tile.hpp -- the fused real-input STFT of one tile of frames, as a device
// routine shared by the single-request kernels and the ragged persistent kernel.
#pragma once

#include "fft_device.hpp"
#include "fft_kernels.hpp"

namespace fftfree {
namespace dev {

// Cooperative store of one contiguous run of `count` plane values (frame-major,
// `bins` per frame) from LDS staging sA/sB laid out [frame][0..L]. For a full spectrum
// (bins = 2L) bins above L read their mirror N-k, conjugated (B negated).
template <int L, bool HALF, int MODE, int WHICH>
__device__ __forceinline__ void store_plane_run(float* __restrict__ dst, const float* sA, const float* sB,
                                                int count, int tid, int nthreads) {
  if (!dst) return;
  constexpr int BINS = HALF ? L + 1 : 2 * L;
  constexpr bool kB = WHICH != 0 && mode_has_b<MODE>();
  constexpr bool kZero = WHICH == 1 && !mode_has_b<MODE>();
  auto value = [&](int i) -> float {
    if constexpr (kZero) {
      return 0.0f;
    } else if constexpr (HALF) {
      return plane_value<MODE, WHICH>(sA[i], kB ? sB[i] : 0.0f);
    } else {
      const int f = i / BINS;
      const int k = i - f * BINS;
      const int j = f * (L + 1) + (k <= L ? k : 2 * L - k);
      const float b = kB ? (k <= L ? sB[j] : -sB[j]) : 0.0f;
      return plane_value<MODE, WHICH>(sA[j], b);
    }
  };
  if ((((uintptr_t)dst) & 15) == 0) {
    const int n4 = count >> 2;
    for (int i = tid; i < n4; i += nthreads) {
      f32x4 o;
      if constexpr (kZero) {
        o = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      } else if constexpr (HALF) {
        const f32x4 a4 = reinterpret_cast<const f32x4*>(sA)[i];
        f32x4 b4 = {0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (kB) b4 = reinterpret_cast<const f32x4*>(sB)[i];
        o.x = plane_value<MODE, WHICH>(a4.x, b4.x);
        o.y = plane_value<MODE, WHICH>(a4.y, b4.y);
        o.z = plane_value<MODE, WHICH>(a4.z, b4.z);
        o.w = plane_value<MODE, WHICH>(a4.w, b4.w);
      } else {
        o.x = value(4 * i);
        o.y = value(4 * i + 1);
        o.z = value(4 * i + 2);
        o.w = value(4 * i + 3);
      }
      __builtin_nontemporal_store(o, reinterpret_cast<f32x4*>(dst) + i);  // written once: streaming
    }
    for (int i = (n4 << 2) + tid; i < count; i += nthreads) __builtin_nontemporal_store(value(i), dst + i);
  } else {
    for (int i = tid; i < count; i += nthreads) __builtin_nontemporal_store(value(i), dst + i);
  }
}

// Fused real-input STFT. One workgroup owns FPW consecutive frames (FPW*T threads).
//  1. every global load is issued up front: the frame's samples times the window
//     (z[n] = x[2n] + i x[2n+1]: the real DFT of size N = 2L is computed as a complex
//     DFT of size L), the pass twiddles and the split twiddles, all into registers;
//  2. Stockham passes through padded LDS (ping-pong pair, or one buffer when IP);
//  3. split, two bins per thread step: with e = (Z_p + conj Z_{L-p})/2 and
//     o = -i (Z_p - conj Z_{L-p})/2, X_p = e + w^p o and X_{L-p} = conj(e - w^p o);
//     inverse plans scale by 1/N; bins 0 and L are real (imag +0);
//  4. post-transform values are staged in LDS [frame][bin] and each output plane is
//     written as one contiguous run of FPW*bins floats (16-byte stores when aligned).
// Stores are the last memory operations of every wave: on CDNA the vector-memory
// counter also counts stores, so a load issued after a store would wait for it.
template <int L, int E, int FPW, int IP, int MODE, int HALF, int DIR, int... Rs>
struct StftTile {
  static constexpr int T = L / E;
  static constexpr int NT = FPW * T;  // threads
  static constexpr int LP = padded_len<L>();
  static constexpr int NBUF = IP ? 1 : 2;
  static constexpr int STAGE = (mode_has_b<MODE>() ? 2 : 1) * FPW * (L + 1);
  static constexpr bool kStageInLds = STAGE <= 2 * FPW * LP;
  static constexpr int kLds = NBUF * FPW * LP;              // float2 of LDS (transform buffers)
  static constexpr int kExtra = kStageInLds ? 0 : STAGE;    // floats of extra staging LDS

  // Tile `tile` (frames tile*FPW ..) of the STFT described by `a`; threads 0..NT-1.
  __device__ static __forceinline__ void run(const StftArgs& a, size_t tile, float2* lds, float* stage_extra) {
    constexpr int N = 2 * L;
    constexpr int NP = sizeof...(Rs);
    constexpr int R0 = FirstOf<Rs...>::value;
    constexpr int Rl = LastOf<Rs...>::value;
    constexpr int J = (L / 2 + T) / T;  // split steps: pairs p = t + j*T, p <= L/2
    // staging ([frame][0..L]) goes to the ping-pong buffer not holding Z, or (in place)
    // over the one buffer
    static_assert(product<Rs...>() == L, "radices must multiply to L");
    static_assert((T & (T - 1)) == 0, "threads per transform must be a power of two");

    const int tid = threadIdx.x;
    const int tr = tid / T;
    const int t = tid & (T - 1);
    const size_t f = tile * FPW + tr;
    const size_t start = (f + a.frame_offset) * (size_t)a.hop;
    size_t cnt = 0;
    if (f < a.frames && start < a.pcm_len) {
      const size_t rem = a.pcm_len - start;
      cnt = rem < (size_t)a.window ? rem : (size_t)a.window;
    }
    const float* __restrict__ src = a.pcm + (cnt ? start : 0);
    const float* __restrict__ win = a.win;

    // 1. loads
    float2 v[E];
    if (cnt == (size_t)N && ((((uintptr_t)src) & 7) == 0)) {
      const float2* __restrict__ s2 = reinterpret_cast<const float2*>(src);
      const float2* __restrict__ w2 = reinterpret_cast<const float2*>(win);
#pragma unroll
      for (int q = 0; q < E / R0; ++q)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const int n = t + q * T + r * (L / R0);
          float2 x = s2[n];
          if (w2) {
            const float2 w = w2[n];
            x.x *= w.x;
            x.y *= w.y;
          }
          v[q * R0 + r] = x;
        }
    } else {
#pragma unroll
      for (int q = 0; q < E / R0; ++q)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const int n = t + q * T + r * (L / R0);
          v[q * R0 + r] = make_float2(frame_sample(src, win, cnt, 2 * n), frame_sample(src, win, cnt, 2 * n + 1));
        }
    }
    float2 twr[E * NP];
    StockhamPlan<L, E, DIR, Rs...>::load_twiddles(twr, t, a.tw, 2);
    float2 wsp[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int p = t + j * T;
      wsp[j] = load_tw<DIR>(a.tw, p <= L / 2 ? p : 0);
    }

    // 2. passes
    float2* Zb;
    float2* buf0 = lds + (size_t)tr * LP;
    float2* buf1 = lds + (size_t)((NBUF - 1) * FPW + tr) * LP;
    StockhamPlan<L, E, DIR, Rs...>::template run<IP != 0>(v, t, buf0, buf1, twr, Zb);
    store_natural<L, E, Rl>(v, t, Zb);
    __syncthreads();

    // 3. split + post-transform values into registers
    constexpr float scale = DIR > 0 ? 1.0f / (float)N : 1.0f;
    float Av[2 * J], Bv[2 * J];
    const float2* Z = Zb;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int p = t + j * T;
      const float2 zp = Z[pidx(p & (L - 1))];
      const float2 zc = c_conj(Z[pidx((L - p) & (L - 1))]);
      const float2 e = c_scale(c_add(zp, zc), 0.5f);
      const float2 d = c_sub(zp, zc);
      const float2 wo = c_mul(wsp[j], make_float2(0.5f * d.y, -0.5f * d.x));
      float2 X1 = c_add(e, wo);
      float2 X2 = c_conj(c_sub(e, wo));
      if (DIR > 0) {
        X1 = c_scale(X1, scale);
        X2 = c_scale(X2, scale);
      }
      if (p == 0) {  // bins 0 and L
        X1.y = 0.0f;
        X2.y = 0.0f;
      }
      stage_vals<MODE>(X1, Av[2 * j], Bv[2 * j]);
      stage_vals<MODE>(X2, Av[2 * j + 1], Bv[2 * j + 1]);
    }
    float* sA = kStageInLds ? reinterpret_cast<float*>(lds + (size_t)((NP & (NBUF - 1)) ? FPW * LP : 0)) : stage_extra;
    if constexpr (IP) {
      if (kStageInLds) __syncthreads();  // Z is read; the one buffer becomes staging
      sA = kStageInLds ? reinterpret_cast<float*>(lds) : stage_extra;
    }
    float* sB = sA + FPW * (L + 1);
    float* rowA = sA + tr * (L + 1);
    float* rowB = sB + tr * (L + 1);
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int p = t + j * T;
      if (p <= L / 2) {
        rowA[p] = Av[2 * j];
        rowA[L - p] = Av[2 * j + 1];
        if constexpr (mode_has_b<MODE>()) {
          rowB[p] = Bv[2 * j];
          rowB[L - p] = Bv[2 * j + 1];
        }
      }
    }
    __syncthreads();

    // 4. plane stores
    const size_t f0 = tile * FPW;
    if (f0 < a.frames) {
      const size_t left = a.frames - f0;
      constexpr int BINS = HALF ? L + 1 : N;
      const int count = (int)((left < (size_t)FPW ? left : (size_t)FPW) * BINS);
      const size_t base = f0 * (size_t)BINS;
      store_plane_run<L, HALF != 0, MODE, 0>(a.o_re ? a.o_re + base : nullptr, sA, sB, count, tid, NT);
      store_plane_run<L, HALF != 0, MODE, 1>(a.o_im ? a.o_im + base : nullptr, sA, sB, count, tid, NT);
      store_plane_run<L, HALF != 0, MODE, 2>(a.o_mag ? a.o_mag + base : nullptr, sA, sB, count, tid, NT);
    }
  }
};

}  // namespace dev
}  // namespace fftfree

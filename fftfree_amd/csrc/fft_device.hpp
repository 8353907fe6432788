// fft_device.hpp -- gfx950 device building blocks of the batched FFT/STFT engine.
//
// Complex values are float2 (x = real, y = imag). A transform of complex length L is
// held by T = L/E threads, each owning E values in registers. Every pass is a Stockham
// autosort radix-R step (natural-order output, no bit reversal): butterfly b reads
// src[b + r*L/R], multiplies by exp(DIR*2*pi*i*(b mod Ns)*r/(Ns*R)), runs an R-point DFT
// in registers and writes dst[(b - b mod Ns)*R + (b mod Ns) + r*Ns]. Data moves between
// passes through LDS ping-pong buffers, one barrier per pass. DIR = -1 is the forward
// transform, +1 the inverse (unscaled here; callers apply 1/N).
//
// The reference computes the same DFT with lgN radix-2 stages (eigen_fft.hpp:2902-3057
// for Cooley-Tukey, 3472-3532 for Stockham); radix-8/16 passes change only the order of
// float operations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fftfree {
namespace dev {

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
// a * b with fused multiply-adds: (ax*bx - ay*by, ax*by + ay*bx)
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// a * (DIR * i)
template <int DIR>
__device__ __forceinline__ float2 c_mul_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// cos(2*pi*k/16) and sin(2*pi*k/16), k = 0..15, correctly rounded to float.
__device__ constexpr float kCos16[16] = {
    1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
    0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
    -1.0f, -0.92387953251128675613f, -0.70710678118654752440f, -0.38268343236508977173f,
    0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f};
__device__ constexpr float kSin16[16] = {
    0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f,
    1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
    0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
    -1.0f, -0.92387953251128675613f, -0.70710678118654752440f, -0.38268343236508977173f};

// x * exp(DIR * 2*pi*i * K / R) for a compile-time K, R <= 16 (R a power of two).
template <int R, int K, int DIR>
__device__ __forceinline__ float2 c_mul_root(float2 x) {
  constexpr int k16 = (K * (16 / R)) & 15;
  if constexpr (k16 == 0) {
    return x;
  } else if constexpr (k16 == 8) {
    return make_float2(-x.x, -x.y);
  } else if constexpr (k16 == 4) {
    return c_mul_i<DIR>(x);
  } else if constexpr (k16 == 12) {
    return c_mul_i<-DIR>(x);
  } else {
    constexpr float c = kCos16[k16];
    constexpr float s = (DIR < 0 ? -1.0f : 1.0f) * kSin16[k16];
    return make_float2(__builtin_fmaf(x.x, c, -x.y * s), __builtin_fmaf(x.x, s, x.y * c));
  }
}

// In-register R-point DFT, natural order in and out (recursive radix-2 split on
// compile-time indices so every register index and twiddle constant folds).
template <int R, int DIR, int STRIDE = 1, int OFF = 0>
struct RegDFT;

template <int DIR, int STRIDE, int OFF>
struct RegDFT<1, DIR, STRIDE, OFF> {
  __device__ static __forceinline__ void run(const float2* in, float2* out) { out[0] = in[OFF]; }
};

template <int R, int DIR, int STRIDE, int OFF>
struct RegDFT {
  template <int K>
  __device__ static __forceinline__ void combine(const float2* e, const float2* o, float2* out) {
    if constexpr (K < R / 2) {
      const float2 t = c_mul_root<R, K, DIR>(o[K]);
      out[K] = c_add(e[K], t);
      out[K + R / 2] = c_sub(e[K], t);
      combine<K + 1>(e, o, out);
    }
  }
  __device__ static __forceinline__ void run(const float2* in, float2* out) {
    float2 e[R / 2], o[R / 2];
    RegDFT<R / 2, DIR, STRIDE * 2, OFF>::run(in, e);
    RegDFT<R / 2, DIR, STRIDE * 2, OFF + STRIDE>::run(in, o);
    combine<0>(e, o, out);
  }
};

template <int R, int DIR>
__device__ __forceinline__ void reg_dft(float2* v) {
  float2 t[R];
#pragma unroll
  for (int i = 0; i < R; ++i) t[i] = v[i];
  RegDFT<R, DIR>::run(t, v);
}

// Twiddle exp(DIR*2*pi*i*m/Lt) from a forward table tw[m'] = exp(-2*pi*i*m'/N) where
// Lt divides N: m' = m * (N/Lt) = m * stride. The inverse direction conjugates.
template <int DIR>
__device__ __forceinline__ float2 load_tw(const float2* __restrict__ tw, int idx) {
  const float2 w = tw[idx];
  return DIR < 0 ? w : make_float2(w.x, -w.y);
}

// Radix-list helpers.
template <int R0, int... Rs>
struct FirstOf {
  static constexpr int value = R0;
};

template <int R0, int... Rs>
struct LastOf {
  static constexpr int value = LastOf<Rs...>::value;
};
template <int R0>
struct LastOf<R0> {
  static constexpr int value = R0;
};

template <int... Rs>
constexpr int product() {
  return (1 * ... * Rs);
}

// LDS layout: one float2 of padding per 16 keeps the autosort scatter stores (stride R
// per lane) from hitting the same banks; index i lives at pidx(i).
constexpr int kPadShift = 4;
__device__ __forceinline__ int pidx(int i) { return i + (i >> kPadShift); }
template <int L>
constexpr int padded_len() {
  return L + (L >> kPadShift);
}

// The passes of one transform of length L held by T = L/E threads, E values per thread.
// Pass P has radix R and input span Ns; butterfly b = t + q*T (q < E/R) owns elements
// v[q*R + r] = src[b + r*L/R]. Twiddles exp(DIR 2 pi i (b mod Ns) r / (Ns R)) come from
// registers preloaded by load_twiddles (slot P*E + q*(R-1) + r-1), so every global
// load of a kernel can be issued up front.
template <int L, int E, int DIR, int Ns, int P, int R, int... Rest>
struct Stockham {
  static constexpr int T = L / E;
  static constexpr bool kLast = sizeof...(Rest) == 0;

  // tw is a forward table exp(-2 pi i m / (L*stride)).
  __device__ static __forceinline__ void load_twiddles(float2* twr, int t, const float2* __restrict__ tw,
                                                       int stride) {
    if constexpr (Ns > 1) {
#pragma unroll
      for (int q = 0; q < E / R; ++q) {
        const int b = t + q * T;
        const int base = (b & (Ns - 1)) * (L / R / Ns) * stride;
#pragma unroll
        for (int r = 1; r < R; ++r) twr[P * E + q * (R - 1) + r - 1] = load_tw<DIR>(tw, base * r);
      }
    }
    if constexpr (!kLast) Stockham<L, E, DIR, Ns * R, P + 1, Rest...>::load_twiddles(twr, t, tw, stride);
  }

  // Runs this pass and the rest. Between passes data goes through the LDS ping-pong
  // pair (a, b) (padded layout), one barrier each. On return v holds the result in
  // last-pass order v[q*Rl + r] = X[t + q*T + r*L/Rl] (consecutive threads own
  // consecutive outputs) and `free_buf` is the buffer nobody reads any more.
  // INPLACE: a == b, one LDS buffer; a second barrier after each gather keeps the next
  // scatter from overwriting values other threads have not read yet.
  template <bool INPLACE = false>
  __device__ static __forceinline__ void run(float2* v, int t, float2* a, float2* b, const float2* twr,
                                             float2*& free_buf) {
#pragma unroll
    for (int q = 0; q < E / R; ++q) {
      float2* u = v + q * R;
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) u[r] = c_mul(u[r], twr[P * E + q * (R - 1) + r - 1]);
      }
      reg_dft<R, DIR>(u);
    }
    if constexpr (kLast) {
      free_buf = a;
    } else {
#pragma unroll
      for (int q = 0; q < E / R; ++q) {
        const int bb = t + q * T;
        const int k = bb & (Ns - 1);
        const int o = (bb - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) a[pidx(o + r * Ns)] = v[q * R + r];
      }
      __syncthreads();
      constexpr int Rn = FirstOf<Rest...>::value;
#pragma unroll
      for (int q = 0; q < E / Rn; ++q) {
        const int bb = t + q * T;
#pragma unroll
        for (int r = 0; r < Rn; ++r) v[q * Rn + r] = a[pidx(bb + r * (L / Rn))];
      }
      if constexpr (INPLACE) __syncthreads();
      Stockham<L, E, DIR, Ns * R, P + 1, Rest...>::template run<INPLACE>(v, t, b, a, twr, free_buf);
    }
  }
};

template <int L, int E, int DIR, int... Rs>
using StockhamPlan = Stockham<L, E, DIR, 1, 0, Rs...>;

// Stores last-pass-ordered registers to LDS in natural order (padded layout).
template <int L, int E, int Rl>
__device__ __forceinline__ void store_natural(const float2* v, int t, float2* dst) {
  constexpr int T = L / E;
#pragma unroll
  for (int q = 0; q < E / Rl; ++q) {
#pragma unroll
    for (int r = 0; r < Rl; ++r) dst[pidx(t + q * T + r * (L / Rl))] = v[q * Rl + r];
  }
}

// XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin, so remap the
// linear workgroup id so that each XCD walks a contiguous run of tiles (neighbouring
// frames share input samples and land in the same L2). The map is a bijection on
// [0, G) for any G, so correctness does not depend on the dispatch order.
__device__ __forceinline__ int xcd_tile(int i, int G) {
  const int q = G >> 3, rem = G & 7;
  const int x = i & 7, s = i >> 3;
  return x * q + (x < rem ? x : rem) + s;
}

// Launch order in groups of G: the workgroups dispatched to one XCD in a round of 8*G
// consecutive ids (ids i with i % 8 == x) take G consecutive tiles, so neighbouring tiles
// run close together in time and on one XCD (one L2). Block 0 keeps tile 0; a last
// partial round maps to itself, so the map is a bijection on [0, G_total) for any size.
template <int G>
__device__ __forceinline__ int xcd_group_tile(int i, int total) {
  if constexpr (G == 1) {
    return i;
  } else {
    constexpr int S = 8 * G;
    const int base = i - i % S;
    if (base + S > total) return i;
    const int k = i - base;
    return base + (k & 7) * G + (k >> 3);
  }
}

// Output modes of the post-transform hook (eigen_fft.hpp:2535-2596) + output copy
// (fft_cffi.cpp:811-825): which value lands in each of the three planes.
enum OutMode : int {
  kOutComplex = 0,    // (re, im, |z|)
  kOutC2RClear = 1,   // (re, 0, |re|)
  kOutMagnitude = 2,  // (|z|, 0, |z|)
  kOutPolar = 3,      // (|z|, arg z, |z|)
};

// atan2(y, x) for finite x, y with the signed-zero results of atan2 (atan2(+-0, +0) =
// +-0, atan2(+-0, -0) = +-pi, atan2(+-0, x < 0) = +-pi exactly): a = min/max of |x|, |y|
// in [0, 1], atan(a) = a * P(a^2) with P of degree 8, then pi/2 - r when |y| > |x|,
// pi - r when x is negative, and the sign of y. P is a minimax fit of atan on [0, 1]
// (iteratively reweighted least squares on Chebyshev nodes) whose float evaluation is
// within 9.2e-8 of atan; the reciprocal, the rounding of pi - r and pi/2 - r and the
// representation error of float(pi) bring the result to within 4e-7 of atan2 (3.6e-7
// worst case in a float emulation over 6M arguments, rms 7e-8), exact on the axes.
// About 20 instructions, against 45 for atan2f.
__device__ __forceinline__ float fast_atan2f(float y, float x) {
  const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
  const float mx = __builtin_fmaxf(ax, ay), mn = __builtin_fminf(ax, ay);
  const float a = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
  const float t = a * a;
  float p = 0.0024562347680330276f;
  p = __builtin_fmaf(p, t, -0.01439903024584055f);
  p = __builtin_fmaf(p, t, 0.03977661952376366f);
  p = __builtin_fmaf(p, t, -0.07234367728233337f);
  p = __builtin_fmaf(p, t, 0.10498644411563873f);
  p = __builtin_fmaf(p, t, -0.1416112184524536f);
  p = __builtin_fmaf(p, t, 0.1998588591814041f);
  p = __builtin_fmaf(p, t, -0.33332595229148865f);
  p = __builtin_fmaf(p, t, 0.9999998807907104f);
  float r = p * a;
  if (ay > ax) r = 1.57079637050628662f - r;              // float(pi/2)
  if (__builtin_signbit(x)) r = 3.14159274101257324f - r;  // float(pi)
  return __builtin_copysignf(r, y);
}

__device__ __forceinline__ void store_planes(float2 z, int mode, size_t idx, float* __restrict__ o_re,
                                             float* __restrict__ o_im, float* __restrict__ o_mag) {
  float a, b, c;
  if (mode == kOutComplex) {
    a = z.x;
    b = z.y;
    c = __builtin_sqrtf(__builtin_fmaf(z.x, z.x, z.y * z.y));
  } else if (mode == kOutC2RClear) {
    a = z.x;
    b = 0.0f;
    c = __builtin_fabsf(z.x);
  } else {
    a = __builtin_sqrtf(__builtin_fmaf(z.x, z.x, z.y * z.y));
    b = (mode == kOutPolar) ? fast_atan2f(z.y, z.x) : 0.0f;
    c = a;
  }
  if (o_re) o_re[idx] = a;
  if (o_im) o_im[idx] = b;
  if (o_mag) o_mag[idx] = c;
}

// Reads framed, windowed real samples x[m] = pcm[start+m] * win[m] for m < cnt, else 0
// (fft_cffi.cpp:766-791; the product is a single float multiply as in the reference).
__device__ __forceinline__ float frame_sample(const float* __restrict__ src, const float* __restrict__ win,
                                             size_t cnt, int m) {
  if ((size_t)m >= cnt) return 0.0f;
  const float x = src[m];
  return win ? x * win[m] : x;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// The pair (A, B) staged for one bin of the post-transformed value, per output mode:
// complex (re, im), C2R-clear (re, -), magnitude (|z|, -), polar (|z|, arg z).
// |z| uses the hardware square root (within 1 ulp); arg z is fast_atan2f.
template <int MODE>
__device__ __forceinline__ void stage_vals(float2 X, float& A, float& B) {
  if constexpr (MODE == kOutComplex || MODE == kOutC2RClear) {
    A = X.x;
    B = X.y;
  } else {
    A = __builtin_amdgcn_sqrtf(__builtin_fmaf(X.x, X.x, X.y * X.y));
    B = (MODE == kOutPolar) ? fast_atan2f(X.y, X.x) : 0.0f;
  }
}

template <int MODE>
constexpr bool mode_has_b() {
  return MODE == kOutComplex || MODE == kOutPolar;
}

// Plane WHICH (0 re, 1 im, 2 mag) from the staged pair: the output copy of
// fft_cffi.cpp:811-825 (mag plane = |v|, or re when polar) applied to the value.
template <int MODE, int WHICH>
__device__ __forceinline__ float plane_value(float A, float B) {
  if constexpr (WHICH == 0) return A;
  if constexpr (WHICH == 1) return mode_has_b<MODE>() ? B : 0.0f;
  if constexpr (MODE == kOutComplex) return __builtin_amdgcn_sqrtf(__builtin_fmaf(A, A, B * B));
  if constexpr (MODE == kOutC2RClear) return __builtin_fabsf(A);
  return A;
}

// LDS exchanges between passes stay inside one wave, so they are ordered by the wave's
// in-order LDS queue plus a compiler fence; no workgroup barrier is needed.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exchange barrier of a transform held by T threads: a wave fence when it is one wave,
// else a workgroup barrier (every wave of the workgroup runs the same pass schedule).
template <int T>
__device__ __forceinline__ void pass_sync() {
  if constexpr (T <= 64)
    wave_lds_fence();
  else
    __syncthreads();
}

template <int L, int E, int DIR, int Ns, int P, int R, int... Rest>
struct WaveStockham {
  static constexpr int T = L / E;
  static constexpr bool kLast = sizeof...(Rest) == 0;
  // Runs this pass and the rest in one wave; result in last-pass order in v. hook(P) runs
  // after pass P's arithmetic (used to slot in independent stores).
  template <class Hook>
  __device__ static __forceinline__ void run(float2* v, int t, float2* buf, const float2* twr, Hook&& hook) {
#pragma unroll
    for (int q = 0; q < E / R; ++q) {
      float2* u = v + q * R;
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) u[r] = c_mul(u[r], twr[P * E + q * (R - 1) + r - 1]);
      }
      reg_dft<R, DIR>(u);
    }
    hook(P);
    if constexpr (!kLast) {
#pragma unroll
      for (int q = 0; q < E / R; ++q) {
        const int bb = t + q * T;
        const int k = bb & (Ns - 1);
        const int o = (bb - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[pidx(o + r * Ns)] = v[q * R + r];
      }
      pass_sync<T>();
      constexpr int Rn = FirstOf<Rest...>::value;
#pragma unroll
      for (int q = 0; q < E / Rn; ++q) {
        const int bb = t + q * T;
#pragma unroll
        for (int r = 0; r < Rn; ++r) v[q * Rn + r] = buf[pidx(bb + r * (L / Rn))];
      }
      pass_sync<T>();
      WaveStockham<L, E, DIR, Ns * R, P + 1, Rest...>::run(v, t, buf, twr, hook);
    }
  }
};

// Cross-lane exchanges in registers (gfx950 DPP and permlane swaps; no LDS). xlane_swap<B>
// exchanges a register pair across lane bit B: afterwards a lane with bit B clear holds
// (a, the partner's a) and a lane with it set (the partner's b, b), the partner being
// lane ^ (1 << B). Lane bits 5 and 4 are one v_permlane32/16_swap per pair; bits 3 and 2
// are two DPP row rotations that write only the lanes their bank mask selects (banks are
// lane bits 2..3); bits 1 and 0 are DPP quad permutations and a select.
template <int B>
__device__ __forceinline__ void xlane_swap(float& a, float& b, int lane) {
  const int ia = __builtin_bit_cast(int, a), ib = __builtin_bit_cast(int, b);
  int na, nb;
  if constexpr (B == 5 || B == 4) {
    const auto r = B == 5 ? __builtin_amdgcn_permlane32_swap((unsigned)ia, (unsigned)ib, false, false)
                          : __builtin_amdgcn_permlane16_swap((unsigned)ia, (unsigned)ib, false, false);
    na = (int)r[0];
    nb = (int)r[1];
  } else if constexpr (B == 3) {
    // row_ror:8 reads lane ^ 8 inside each 16-lane row; banks 2, 3 have bit 3 set
    na = __builtin_amdgcn_update_dpp(ia, ib, 0x128, 0xf, 0xc, false);
    nb = __builtin_amdgcn_update_dpp(ib, ia, 0x128, 0xf, 0x3, false);
  } else if constexpr (B == 2) {
    // row_ror:n reads lane (i - n) mod 16: lanes with bit 2 set (banks 1, 3) read i - 4
    // (row_ror:4), the others i + 4 (row_ror:12)
    na = __builtin_amdgcn_update_dpp(ia, ib, 0x124, 0xf, 0xa, false);
    nb = __builtin_amdgcn_update_dpp(ib, ia, 0x12c, 0xf, 0x5, false);
  } else {
    static_assert(B == 0 || B == 1, "lane bits 0..5");
    constexpr int ctrl = B == 0 ? 0xb1 : 0x4e;  // quad_perm [1,0,3,2] / [2,3,0,1]
    const int pa = __builtin_amdgcn_mov_dpp(ia, ctrl, 0xf, 0xf, false);
    const int pb = __builtin_amdgcn_mov_dpp(ib, ctrl, 0xf, 0xf, false);
    const bool set = (lane >> B) & 1;
    na = set ? pb : ia;
    nb = set ? ib : pa;
  }
  a = __builtin_bit_cast(float, na);
  b = __builtin_bit_cast(float, nb);
}

// One stage of a lane-digit / register-index transpose over v[0..E): bit j of the
// register index trades places with lane bit B.
template <int B, int J, int E = 8>
__device__ __forceinline__ void xlane_stage8(float2* v, int lane) {
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (!(e & (1 << J))) {
      xlane_swap<B>(v[e].x, v[e | (1 << J)].x, lane);
      xlane_swap<B>(v[e].y, v[e | (1 << J)].y, lane);
    }
}
// Transposes the 3-bit lane digit at lane bits S..S+2 with the index of v[0..8): lane
// digit d, register e afterwards holds what lane digit e, register d held.
template <int S>
__device__ __forceinline__ void xlane_transpose8(float2* v, int lane) {
  xlane_stage8<S, 0>(v, lane);
  xlane_stage8<S + 1, 1>(v, lane);
  xlane_stage8<S + 2, 2>(v, lane);
}

// The three radix-8 passes of a 512-point transform held by one wave (E = 8 values per
// lane), with the exchanges between passes done in registers: the same butterflies,
// twiddles and operation order as WaveStockham<512, 8, DIR, 1, 0, 8, 8, 8>, so the same
// bits, but butterfly b of passes 1 and 2 runs on lane (b >> 3) | ((b & 7) << 3) (the
// lane's 3-bit digits rotated), which makes each exchange one transpose of a lane digit
// with the register index:
//  * after pass 0, lane t register r holds element 8t + r; pass 1's butterfly
//    b = b0 + 8 b1 needs elements b + 64r, held by lane b1 + 8r register b0: transpose
//    lane bits 3..5 with the register index;
//  * after pass 1, lane b1 + 8 b0 register r holds element b0 + 8r + 64 b1; pass 2's
//    butterfly c = c0 + 8 c1 needs elements c + 64r, held by lane r + 8 c0 register c1:
//    transpose lane bits 0..2 with the register index.
// Twiddles: slots as load_twiddles fills them for thread index rot_lane(lane). On return
// lane l register r holds X[rot_lane(l) + 64 r] (store_natural_rot puts it in LDS).
// LO_LDS: the second exchange (lane bits 0..2, the dearer one in registers: DPP quad
// permutations need a select each) goes through LDS as in WaveStockham instead, pass 2's
// butterfly c runs on lane c (twiddle slots for thread index lane), and the result is in
// WaveStockham's order (store_natural).
__device__ __forceinline__ int rot_lane(int lane) { return (lane >> 3) | ((lane & 7) << 3); }
template <int DIR, bool LO_LDS = false>
struct WaveRegs888 {
  template <class Hook>
  __device__ static __forceinline__ void run(float2* v, int lane, float2* buf, const float2* twr, Hook&& hook) {
    reg_dft<8, DIR>(v);
    hook(0);
    xlane_transpose8<3>(v, lane);
#pragma unroll
    for (int r = 1; r < 8; ++r) v[r] = c_mul(v[r], twr[8 + r - 1]);
    reg_dft<8, DIR>(v);
    hook(1);
    if constexpr (LO_LDS) {
      // pass 1's butterfly b = rot_lane(lane) wrote elements (b & 7) + 8r + 64 (b >> 3)
      const int b = rot_lane(lane);
#pragma unroll
      for (int r = 0; r < 8; ++r) buf[pidx((b & 7) + 8 * r + 64 * (b >> 3))] = v[r];
      wave_lds_fence();
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = buf[pidx(lane + 64 * r)];
      wave_lds_fence();
    } else {
      xlane_transpose8<0>(v, lane);
    }
#pragma unroll
    for (int r = 1; r < 8; ++r) v[r] = c_mul(v[r], twr[16 + r - 1]);
    reg_dft<8, DIR>(v);
    hook(2);
  }
};
// The passes of a 1024-point transform held by one wave (E = 16 values per lane), radices
// 16, 16, 4, with the exchanges in registers: the butterflies, twiddles and operation
// order of WaveStockham<1024, 16, DIR, 1, 0, 16, 16, 4> on permuted lanes.
//  * after pass 0 lane t register r holds element 16t + r; pass 1's butterfly b (elements
//    b + 64r) runs on lane (b >> 4) | ((b & 15) << 2): transpose lane bits 2..5 with the
//    register index (one permlane swap or two bank-masked DPP moves per pair);
//  * after pass 1 lane l register r holds element (l >> 2) + 16r + 256(l & 3); pass 2's
//    butterfly c (elements c + 256r, four per lane in registers 4q .. 4q+3) is
//    c = (l >> 2) + 16(4q + (l & 3)): transpose lane bits 0..1 with register bits 0..1.
// twr: pass 1's slots 16 + r - 1 for butterfly c1(lane), pass 2's 32 + 3q + r - 1 for
// c2(lane, q) (regx1024_pass1 / regx1024_pass2). On return lane l register 4q + r holds
// X[c2(l, q) + 256 r].
__device__ __forceinline__ int regx1024_pass1(int lane) { return (lane >> 2) + 16 * (lane & 3); }
__device__ __forceinline__ int regx1024_pass2(int lane, int q) { return (lane >> 2) + 16 * (4 * q + (lane & 3)); }
template <int DIR>
struct WaveRegs16x16x4 {
  template <class Hook>
  __device__ static __forceinline__ void run(float2* v, int lane, const float2* twr, Hook&& hook) {
    reg_dft<16, DIR>(v);
    hook(0);
    xlane_stage8<2, 0, 16>(v, lane);
    xlane_stage8<3, 1, 16>(v, lane);
    xlane_stage8<4, 2, 16>(v, lane);
    xlane_stage8<5, 3, 16>(v, lane);
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = c_mul(v[r], twr[16 + r - 1]);
    reg_dft<16, DIR>(v);
    hook(1);
    xlane_stage8<0, 0, 16>(v, lane);
    xlane_stage8<1, 1, 16>(v, lane);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int r = 1; r < 4; ++r) v[4 * q + r] = c_mul(v[4 * q + r], twr[32 + 3 * q + r - 1]);
      reg_dft<4, DIR>(v + 4 * q);
    }
    hook(2);
  }
  // the pass twiddles for this lane's butterflies; tw is exp(-2 pi i m / (1024 * stride))
  __device__ static __forceinline__ void load_twiddles(float2* twr, int lane, const float2* __restrict__ tw,
                                                       int stride) {
    const int b1 = regx1024_pass1(lane) & 15;  // pass 1: span 16, twiddle step L/R/Ns = 4
#pragma unroll
    for (int r = 1; r < 16; ++r) twr[16 + r - 1] = load_tw<DIR>(tw, b1 * 4 * stride * r);
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // pass 2: span 256, twiddle step 1
      const int b2 = regx1024_pass2(lane, q);
#pragma unroll
      for (int r = 1; r < 4; ++r) twr[32 + 3 * q + r - 1] = load_tw<DIR>(tw, b2 * stride * r);
    }
  }
};

__device__ __forceinline__ void store_natural_rot(const float2* v, int lane, float2* dst) {
  const int c = rot_lane(lane);
#pragma unroll
  for (int r = 0; r < 8; ++r) dst[pidx(c + 64 * r)] = v[r];
}

}  // namespace dev
}  // namespace fftfree

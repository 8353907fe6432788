// fft_adjoint.hip -- batched STFT of equal-length signals, and its adjoint (the gradient
// of a loss on the output planes with respect to the samples) for the differentiable path.
//
// Forward (stft_signals_kernel): for every frame f of every signal,
//   X_fk = sum_{n<W} w[n] x[fH+n] e^{-2 pi i kn/N},  k = 0..N/2,
// written as a R2C half-spectrum context writes its planes (post hook
// eigen_fft.hpp:2535-2596, output copy fft_cffi.cpp:811-825), in [signal][frame][bin]
// order. Frame count per signal: the reference rule (fft_cffi.cpp:715-750).
//
// Adjoint (stft_adjoint_kernel): with upstream gradients Ga, Gb of the two value planes
// (re/im for complex output, |X|/arg X for polar, |X| for magnitude),
//   dL/dx[j] = sum_{f: 0 <= j-fH < W} w[j-fH] r_f[j-fH],
//   r_f[n]   = Re sum_{k=0}^{N/2} V_fk e^{+2 pi i kn/N},
//   V = Ga + i Gb (complex), Ga X/|X| + i Gb X/|X|^2 (polar), Ga X/|X| (magnitude),
// V = 0 where |X| = 0, and no imaginary-part terms at bins 0 and N/2, whose imaginary
// parts the forward pins to 0. Each frame's spectrum is recomputed from the samples;
// nothing is stored by the forward pass. r_f is a real inverse DFT of a half spectrum:
// with L = N/2, e = e^{+2 pi i p/N},
//   Y_p = (V_p + conj V_{L-p})/2 + i e (V_p - conj V_{L-p})/2   (0 < p < L),
//   Y_0 = (Re V_0 + Re V_L) + i (Re V_0 - Re V_L),
// the unscaled L-point inverse y = IDFT(Y) gives r[2m] = Re y[m], r[2m+1] = Im y[m]
// (the mirror image of the forward split). The inverse runs as conj(DFT(conj Y)) on the
// forward engine with the forward twiddles.
//
// Overlap-add is a gather in a fixed frame order, so the result is deterministic. A
// workgroup owns a run of consecutive frame positions of one signal; each of its four
// waves transforms one frame per step into a slot of an 8-slot LDS ring, and after each
// step the workgroup writes the samples that the last four frames complete, each as the
// ascending-frame sum of the (at most five) windowed r_f that cover it. The four frames
// before its run are recomputed as a warm-up instead of exchanged with the neighbouring
// workgroup, so every sample is written exactly once.
#include "fft_device.hpp"
#include "fft_kernels.hpp"

namespace fftfree {
using namespace dev;

namespace {

constexpr int kWaves = 4;       // frames in flight per workgroup (one per wave)
constexpr int kRing = 8;        // LDS ring slots (4 current + 4 previous frames)
constexpr int kFwdTile = 32;    // frames per forward workgroup
constexpr int kAdjSteps = 8;    // adjoint steps per workgroup (kWaves frame positions each)

// Upstream-gradient spectrum V for one bin (see the header comment).
template <int MODE>
__device__ __forceinline__ float2 grad_spectrum(float2 X, float ga, float gb, bool real_bin) {
  if constexpr (MODE == kOutComplex) {
    return make_float2(ga, real_bin ? 0.0f : gb);
  } else {
    const float m2 = __builtin_fmaf(X.x, X.x, X.y * X.y);
    if (m2 == 0.0f) return make_float2(0.0f, 0.0f);
    const float inv = __builtin_amdgcn_rsqf(m2);  // 1/|X|
    float2 V = c_scale(X, ga * inv);
    if (MODE == kOutPolar && !real_bin) {
      const float s = gb * inv * inv;  // i Gb X/|X|^2
      V.x -= s * X.y;
      V.y += s * X.x;
    }
    return V;
  }
}

// One frame's samples, windowed, in first-pass order: z[m] = x[2m] w[2m] + i x[2m+1] w[2m+1].
template <int L, int E, int R0>
__device__ __forceinline__ void load_frame(float2* v, int t, const float* __restrict__ x, size_t len, size_t start,
                                           int window, const float* __restrict__ win) {
  constexpr int T = L / E;
  constexpr int N = 2 * L;
  const size_t rem = len - start;
  const size_t cnt = rem < (size_t)window ? rem : (size_t)window;
  const float* __restrict__ src = x + start;
  if (cnt == (size_t)N && ((((uintptr_t)src) & 7) == 0)) {
    const float2* __restrict__ s2 = reinterpret_cast<const float2*>(src);
    const float2* __restrict__ w2 = reinterpret_cast<const float2*>(win);
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int m = t + q * T + r * (L / R0);
        float2 z = s2[m];
        if (win) {
          const float2 w = w2[m];
          z.x *= w.x;
          z.y *= w.y;
        }
        v[q * R0 + r] = z;
      }
  } else {
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int m = t + q * T + r * (L / R0);
        v[q * R0 + r] = make_float2(frame_sample(src, win, cnt, 2 * m), frame_sample(src, win, cnt, 2 * m + 1));
      }
  }
}

// Forward real DFT of one frame by one wave: on return buf holds Z (natural order) and
// X1[j] = X_p, X2[j] = X_{L-p} for p = t + j*T (j < JF), Xm = X_{L/2}.
template <int L, int E, int... Rs>
__device__ __forceinline__ void frame_spectrum(float2* v, int t, float2* buf, const float2* twr, const float2* wsp,
                                               float2 wmid, float2* X1, float2* X2, float2& Xm) {
  constexpr int T = L / E;
  constexpr int JF = (L / 2) / T;
  constexpr int Rl = LastOf<Rs...>::value;
  WaveStockham<L, E, -1, 1, 0, Rs...>::run(v, t, buf, twr, [](int) {});
  store_natural<L, E, Rl>(v, t, buf);
  wave_lds_fence();
#pragma unroll
  for (int j = 0; j < JF; ++j) {
    const int p = t + j * T;
    const float2 zp = buf[pidx(p)];
    const float2 zc = c_conj(buf[pidx((L - p) & (L - 1))]);
    const float2 e = c_scale(c_add(zp, zc), 0.5f);
    const float2 d = c_sub(zp, zc);
    const float2 wo = c_mul(wsp[j], make_float2(0.5f * d.y, -0.5f * d.x));
    X1[j] = c_add(e, wo);
    X2[j] = c_conj(c_sub(e, wo));
    if (p == 0) {  // bins 0 and L are real
      X1[j].y = 0.0f;
      X2[j].y = 0.0f;
    }
  }
  const float2 zm = buf[pidx(L / 2)];
  const float2 zc = c_conj(zm);
  const float2 e = c_scale(c_add(zm, zc), 0.5f);
  const float2 d = c_sub(zm, zc);
  Xm = c_add(e, c_mul(wmid, make_float2(0.5f * d.y, -0.5f * d.x)));
}

template <int L, int E, int MODE, int... Rs>
__global__ void __launch_bounds__(256) stft_signals_kernel(SignalArgs a) {
  constexpr int T = L / E;
  constexpr int NP = sizeof...(Rs);
  constexpr int LP = padded_len<L>();
  constexpr int JF = (L / 2) / T;
  constexpr int BINS = L + 1;
  static_assert(T == 64, "one transform per wave");
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 lds[kWaves][LP];

  const int wave = threadIdx.x >> 6;
  const int t = threadIdx.x & 63;
  float2* buf = &lds[wave][0];
  const size_t tile = (size_t)xcd_tile(blockIdx.x, gridDim.x);
  const size_t sig = tile / (size_t)a.tiles;
  const size_t f0 = (tile - sig * a.tiles) * kFwdTile;
  const float* __restrict__ x = a.pcm + sig * a.len;

  float2 twr[E * NP];
  StockhamPlan<L, E, -1, Rs...>::load_twiddles(twr, t, a.tw, 2);
  float2 wsp[JF];
#pragma unroll
  for (int j = 0; j < JF; ++j) wsp[j] = load_tw<-1>(a.tw, t + j * T);
  const float2 wmid = load_tw<-1>(a.tw, L / 2);

  for (int i = 0; i < kFwdTile / kWaves; ++i) {
    const size_t f = f0 + (size_t)(i * kWaves + wave);
    if (f >= a.frames) break;  // uniform per wave
    float2 v[E];
    load_frame<L, E, FirstOf<Rs...>::value>(v, t, x, a.len, f * (size_t)a.hop, a.window, a.win);
    float2 X1[JF], X2[JF], Xm;
    frame_spectrum<L, E, Rs...>(v, t, buf, twr, wsp, wmid, X1, X2, Xm);
    const size_t row = (sig * a.frames + f) * (size_t)BINS;
    float* __restrict__ ore = a.o_re ? a.o_re + row : nullptr;
    float* __restrict__ oim = a.o_im ? a.o_im + row : nullptr;
    float* __restrict__ omg = a.o_mag ? a.o_mag + row : nullptr;
    auto put = [&](int k, float2 X) {
      float A, B;
      stage_vals<MODE>(X, A, B);
      if (ore) ore[k] = plane_value<MODE, 0>(A, B);
      if (oim) oim[k] = plane_value<MODE, 1>(A, B);
      if (omg) omg[k] = plane_value<MODE, 2>(A, B);
    };
#pragma unroll
    for (int j = 0; j < JF; ++j) {
      const int p = t + j * T;
      put(p, X1[j]);
      put(L - p, X2[j]);
    }
    if (t == 0) put(L / 2, Xm);
    wave_lds_fence();  // buf is rewritten by the next frame
  }
}

template <int L, int E, int MODE, int... Rs>
__global__ void __launch_bounds__(256) stft_adjoint_kernel(SignalArgs a) {
  constexpr int T = L / E;
  constexpr int N = 2 * L;
  constexpr int NP = sizeof...(Rs);
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int LP = padded_len<L>();
  constexpr int JF = (L / 2) / T;
  constexpr int BINS = L + 1;
  static_assert(T == 64, "one transform per wave");
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 ring[kRing][LP];

  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int t = tid & 63;
  const size_t tile = (size_t)xcd_tile(blockIdx.x, gridDim.x);
  const size_t sig = tile / (size_t)a.tiles;
  const long long f0 = (long long)(tile - sig * a.tiles) * (kWaves * kAdjSteps);
  const float* __restrict__ x = a.pcm + sig * a.len;
  const float* __restrict__ ga = a.g_mag ? a.g_mag + sig * a.frames * BINS : nullptr;
  const float* __restrict__ gb = a.g_phase ? a.g_phase + sig * a.frames * BINS : nullptr;
  float* __restrict__ gx = a.grad + sig * a.len;
  const float2* __restrict__ w2 = reinterpret_cast<const float2*>(a.win);
  const long long frames = (long long)a.frames;
  const long long hop = a.hop;
  const long long W = a.window;

  float2 twr[E * NP];
  StockhamPlan<L, E, -1, Rs...>::load_twiddles(twr, t, a.tw, 2);
  float2 wsp[JF];
#pragma unroll
  for (int j = 0; j < JF; ++j) wsp[j] = load_tw<-1>(a.tw, t + j * T);
  const float2 wmid = load_tw<-1>(a.tw, L / 2);

  for (int step = -1; step < kAdjSteps; ++step) {
    const long long fbase = f0 + (long long)step * kWaves;
    const long long f = fbase + wave;
    float2* buf = &ring[(int)(f & (kRing - 1))][0];
    if (f >= 0 && f < frames) {  // uniform per wave
      float2 v[E];
      load_frame<L, E, R0>(v, t, x, a.len, (size_t)(f * hop), a.window, a.win);
      float2 X1[JF], X2[JF], Xm;
      frame_spectrum<L, E, Rs...>(v, t, buf, twr, wsp, wmid, X1, X2, Xm);
      const size_t row = (size_t)f * BINS;
      auto gA = [&](int k) { return ga ? ga[row + k] : 0.0f; };
      auto gB = [&](int k) { return gb ? gb[row + k] : 0.0f; };
#pragma unroll
      for (int j = 0; j < JF; ++j) {
        const int p = t + j * T;
        X1[j] = grad_spectrum<MODE>(X1[j], gA(p), gB(p), p == 0);
        X2[j] = grad_spectrum<MODE>(X2[j], gA(L - p), gB(L - p), p == 0);
      }
      Xm = grad_spectrum<MODE>(Xm, gA(L / 2), gB(L / 2), false);
      wave_lds_fence();  // every lane has read Z
      // conj(Y) into buf, natural order
#pragma unroll
      for (int j = 0; j < JF; ++j) {
        const int p = t + j * T;
        const float2 Vp = X1[j], Vq = X2[j];  // V_p, V_{L-p}
        if (p == 0) {
          buf[pidx(0)] = make_float2(Vp.x + Vq.x, -(Vp.x - Vq.x));
        } else {
          const float2 ep = c_conj(wsp[j]);  // e^{+2 pi i p/N}
          const float2 s = c_scale(c_add(Vp, c_conj(Vq)), 0.5f);
          const float2 d = c_scale(c_sub(Vp, c_conj(Vq)), 0.5f);
          const float2 ed = c_mul(ep, d);
          const float2 Yp = make_float2(s.x - ed.y, s.y + ed.x);  // s + i e d
          // Y_{L-p} with e' = -conj(e): (V_q + conj V_p)/2 + i e' (V_q - conj V_p)/2
          const float2 s2 = c_conj(s);
          const float2 d2 = c_scale(c_conj(d), -1.0f);
          const float2 ed2 = c_mul(make_float2(-ep.x, ep.y), d2);
          const float2 Yq = make_float2(s2.x - ed2.y, s2.y + ed2.x);
          buf[pidx(p)] = c_conj(Yp);
          buf[pidx(L - p)] = c_conj(Yq);
        }
      }
      if (t == 0) {
        // p = L/2: e = i, so Y = Re V + i * i * (i Im V) = conj(V)
        buf[pidx(L / 2)] = Xm;  // conj(Y) = V
      }
      wave_lds_fence();
#pragma unroll
      for (int q = 0; q < E / R0; ++q)
#pragma unroll
        for (int r = 0; r < R0; ++r) v[q * R0 + r] = buf[pidx(t + q * T + r * (L / R0))];
      wave_lds_fence();
      WaveStockham<L, E, -1, 1, 0, Rs...>::run(v, t, buf, twr, [](int) {});
      wave_lds_fence();
      // y = conj(result): r[2m] = v.x, r[2m+1] = -v.y; windowed contributions into the slot
#pragma unroll
      for (int q = 0; q < E / Rl; ++q)
#pragma unroll
        for (int r = 0; r < Rl; ++r) {
          const int m = t + q * T + r * (L / Rl);
          const float2 y = v[q * Rl + r];
          const float2 w = w2 ? w2[m] : make_float2(1.0f, 1.0f);
          buf[pidx(m)] = make_float2(y.x * w.x, -y.y * w.y);
        }
    }
    __syncthreads();
    if (step >= 0) {
      // samples [fbase*H, (fbase+4)*H): covered by frames fbase-4 .. fbase+3, all in the ring
      const long long j0 = fbase * hop;
      for (int u = tid; u < kWaves * hop; u += 256) {
        const long long j = j0 + u;
        if (j >= (long long)a.len) break;
        long long fmax = j / hop;
        if (fmax > frames - 1) fmax = frames - 1;
        const long long fmin = j >= W ? (j - W + hop) / hop : 0;  // ceil((j-W+1)/H)
        float sum = 0.0f;
        for (long long g = fmin; g <= fmax; ++g) {
          const int n = (int)(j - g * hop);
          const float2 c = ring[(int)(g & (kRing - 1))][pidx(n >> 1)];
          sum += (n & 1) ? c.y : c.x;
        }
        gx[j] = sum;
      }
    }
    __syncthreads();
  }
}

template <int L, int E, int... Rs>
hipError_t signals_launch(const SignalArgs& a, int mode, hipStream_t s) {
  const size_t grid = a.signals * (size_t)a.tiles;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  switch (mode) {
    case kOutComplex:
      hipLaunchKernelGGL((stft_signals_kernel<L, E, kOutComplex, Rs...>), dim3((unsigned)grid), dim3(256), 0, s, a);
      break;
    case kOutMagnitude:
      hipLaunchKernelGGL((stft_signals_kernel<L, E, kOutMagnitude, Rs...>), dim3((unsigned)grid), dim3(256), 0, s, a);
      break;
    case kOutPolar:
      hipLaunchKernelGGL((stft_signals_kernel<L, E, kOutPolar, Rs...>), dim3((unsigned)grid), dim3(256), 0, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int L, int E, int... Rs>
hipError_t adjoint_launch(const SignalArgs& a, int mode, hipStream_t s) {
  const size_t grid = a.signals * (size_t)a.tiles;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  switch (mode) {
    case kOutComplex:
      hipLaunchKernelGGL((stft_adjoint_kernel<L, E, kOutComplex, Rs...>), dim3((unsigned)grid), dim3(256), 0, s, a);
      break;
    case kOutMagnitude:
      hipLaunchKernelGGL((stft_adjoint_kernel<L, E, kOutMagnitude, Rs...>), dim3((unsigned)grid), dim3(256), 0, s, a);
      break;
    case kOutPolar:
      hipLaunchKernelGGL((stft_adjoint_kernel<L, E, kOutPolar, Rs...>), dim3((unsigned)grid), dim3(256), 0, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

bool signals_supported(int n, int window, int hop, bool adjoint) {
  if (n != 512 && n != 1024 && n != 2048) return false;
  if (window <= 0 || window > n || hop <= 0) return false;
  // the adjoint's ring holds the 4 frames before a step: a sample may span at most 5
  return !adjoint || (window - 1) / hop <= kRing - kWaves;
}

hipError_t launch_stft_signals(SignalArgs a, int mode, hipStream_t s) {
  const int n = a.n;
  if (!signals_supported(n, a.window, a.hop, false) || a.len == 0) return hipErrorInvalidValue;
  a.tiles = (int)((a.frames + kFwdTile - 1) / kFwdTile);
  switch (n) {
    case 512: return signals_launch<256, 4, 4, 4, 4, 4>(a, mode, s);
    case 1024: return signals_launch<512, 8, 8, 8, 8>(a, mode, s);
    default: return signals_launch<1024, 16, 16, 16, 4>(a, mode, s);
  }
}

hipError_t launch_stft_adjoint(SignalArgs a, int mode, hipStream_t s) {
  const int n = a.n;
  if (!signals_supported(n, a.window, a.hop, true) || a.len == 0 || !a.grad) return hipErrorInvalidValue;
  // frame positions up to ceil(len / hop) so that every sample is written
  const size_t positions = (a.len + a.hop - 1) / a.hop;
  const size_t per = (size_t)kWaves * kAdjSteps;
  a.tiles = (int)((positions + per - 1) / per);
  switch (n) {
    case 512: return adjoint_launch<256, 4, 4, 4, 4, 4>(a, mode, s);
    case 1024: return adjoint_launch<512, 8, 8, 8, 8>(a, mode, s);
    default: return adjoint_launch<1024, 16, 16, 16, 4>(a, mode, s);
  }
}

}  // namespace fftfree

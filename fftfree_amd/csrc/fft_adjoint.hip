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
// From saved planes (SAVED): the forward's own output planes give V directly (polar:
// V = e^{i arg X} (Ga + i Gb/|X|), 0 where |X| = 0; complex: V = Ga + i Gb), so no frame
// is transformed forward and the samples are not read; the rest is the same.
//
// Overlap-add runs in a fixed order, so the result is deterministic. A workgroup owns a
// run of consecutive frame positions of one signal and accumulates its frames' windowed
// r_f, in ascending frame order, in an LDS ring; each sample is written once. The W-H
// samples at the start of a run also receive the previous run's last frames: each of
// the two workgroups leaves its partial sums for them in a halo slot, and a second,
// small kernel (stft_halo_kernel) writes previous + own, in that order.
#include "fft_device.hpp"
#include "fft_kernels.hpp"

#include <mutex>

namespace fftfree {
using namespace dev;

namespace {

constexpr int kWaves = 4;       // forward: frames in flight per workgroup (one per wave)
constexpr int kFwdTile = 32;    // forward: frames per workgroup

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

// One frame's samples, windowed, in first-pass order: z[m] = x[2m] w[2m] + i x[2m+1] w[2m+1],
// zero past the `cnt` valid samples. Window taps come from LDS (wl, N floats).
template <int L, int E, int R0>
__device__ __forceinline__ void load_frame(float2* v, int t, const float* __restrict__ x, size_t len, size_t start,
                                           int window, const float* wl) {
  constexpr int T = L / E;
  const size_t rem = len - start;
  const int cnt = (int)(rem < (size_t)window ? rem : (size_t)window);
  const float* __restrict__ src = x + start;
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int m = t + q * T + r * (L / R0);
      const float a = 2 * m < cnt ? src[2 * m] : 0.0f;
      const float b = 2 * m + 1 < cnt ? src[2 * m + 1] : 0.0f;
      v[q * R0 + r] = make_float2(a * wl[2 * m], b * wl[2 * m + 1]);
    }
}

// Raw samples of one frame in first-pass order (zero past `cnt`), for prefetching; the
// window is applied at use (window_frame).
template <int L, int E, int R0>
__device__ __forceinline__ void fetch_frame(float2* v, int t, const float* __restrict__ x, size_t len, size_t start,
                                            int window, bool valid) {
  constexpr int T = L / E;
  int cnt = 0;
  if (valid) {
    const size_t rem = len - start;
    cnt = (int)(rem < (size_t)window ? rem : (size_t)window);
  }
  const float* __restrict__ src = x + (valid ? start : 0);
  if (cnt == 2 * L && ((((uintptr_t)src) & 7) == 0)) {  // whole frame: paired loads
    const float2* __restrict__ s2 = reinterpret_cast<const float2*>(src);
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q * R0 + r] = s2[t + q * T + r * (L / R0)];
    return;
  }
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int m = t + q * T + r * (L / R0);
      v[q * R0 + r] = make_float2(2 * m < cnt ? src[2 * m] : 0.0f, 2 * m + 1 < cnt ? src[2 * m + 1] : 0.0f);
    }
}

template <int L, int E, int R0>
__device__ __forceinline__ void window_frame(float2* v, int t, const float* wl) {
  constexpr int T = L / E;
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int m = t + q * T + r * (L / R0);
      v[q * R0 + r].x *= wl[2 * m];
      v[q * R0 + r].y *= wl[2 * m + 1];
    }
}

// Window taps into LDS (1.0 when windows are off); the caller synchronises.
template <int N>
__device__ __forceinline__ void stage_window(float* wl, const float* __restrict__ win) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) wl[i] = win ? win[i] : 1.0f;
}

// Forward real DFT of one frame by one wave, then the split into bins: pair(j, p, X_p,
// X_{L-p}) runs for p = t + j*T < L/2 and mid(X_{L/2}) on lane 0 only. Each call reads
// only Z_p and Z_{L-p} (lane 0: also Z_{L/2}), so a callback may overwrite exactly
// those buf entries in place.
template <int L, int E, int... Rs, class Pair, class Mid>
__device__ __forceinline__ void frame_bins(float2* v, int t, float2* buf, const float2* twr, const float2* wsp,
                                           float2 wmid, Pair&& pair, Mid&& mid) {
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
    float2 X1 = c_add(e, wo);
    float2 X2 = c_conj(c_sub(e, wo));
    if (p == 0) {  // bins 0 and L are real
      X1.y = 0.0f;
      X2.y = 0.0f;
    }
    pair(j, p, X1, X2);
  }
  if (t == 0) {
    const float2 zm = buf[pidx(L / 2)];
    const float2 zc = c_conj(zm);
    const float2 e = c_scale(c_add(zm, zc), 0.5f);
    const float2 d = c_sub(zm, zc);
    mid(c_add(e, c_mul(wmid, make_float2(0.5f * d.y, -0.5f * d.x))));
  }
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
  __shared__ float wl[2 * L];

  const int wave = threadIdx.x >> 6;
  const int t = threadIdx.x & 63;
  float2* buf = &lds[wave][0];
  stage_window<2 * L>(wl, a.win);
  __syncthreads();
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
    load_frame<L, E, FirstOf<Rs...>::value>(v, t, x, a.len, f * (size_t)a.hop, a.window, wl);
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
    frame_bins<L, E, Rs...>(
        v, t, buf, twr, wsp, wmid,
        [&](int, int p, float2 X1, float2 X2) {
          put(p, X1);
          put(L - p, X2);
        },
        [&](float2 Xm) { put(L / 2, Xm); });
    wave_lds_fence();  // buf is rewritten by the next frame
  }
}

// Adjoint kernel. T = L/E threads transform one frame; a step transforms FPS = 256/T
// consecutive frame positions, one per thread group; the passes of a transform exchange
// through LDS with wave fences when a frame is one wave (T = 64), else with workgroup
// barriers. Per step:
//  1. each group recomputes X for its frame, forms V, overwrites Z with conj(Y) in place
//     (each thread owns the pair p, L-p), inverse-transforms and leaves r_f in its buffer;
//  2. every sample the step's frames touch gets their windowed r_f added to an LDS
//     accumulator ring, frames in ascending order (a fixed order: deterministic);
//  3. the FPS*H samples that no later frame touches are written out and their ring
//     entries cleared for reuse.
// The first W-H samples of the run go to the left halo slot, and after the last step
// the samples past the run that its frames touch go to the right halo slot;
// stft_halo_kernel completes them.
template <int L, int E, int MODE, int ACCM, bool SAVED, int... Rs>
__global__ void __launch_bounds__(256) stft_adjoint_kernel(SignalArgs a) {
  constexpr int T = L / E;
  constexpr int N = 2 * L;
  constexpr int FPS = 256 / T;
  constexpr int NP = sizeof...(Rs);
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int LP = padded_len<L>();
  constexpr int JF = (L / 2) / T;
  constexpr int BINS = L + 1;
  constexpr int ACC = ACCM * N;  // accumulator ring (samples); the host checks W + FPS*H <= ACC
  static_assert(T * FPS == 256 && JF * T == L / 2, "thread layout");
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 fbuf[FPS][LP];
  __shared__ float acc[ACC];
  __shared__ float wl[N];

  const int tid = threadIdx.x;
  const int slot = tid / T;
  const int t = tid - slot * T;
  float2* buf = &fbuf[slot][0];
  const size_t tile = (size_t)xcd_tile(blockIdx.x, gridDim.x);
  const size_t sig = tile / (size_t)a.tiles;
  const long long steps = a.steps;
  const int tix = (int)(tile - sig * a.tiles);  // run index within the signal
  const long long f0 = (long long)tix * FPS * steps;
  const long long fend = f0 + FPS * steps;
  const float* __restrict__ x = SAVED ? nullptr : a.pcm + sig * a.len;  // saved planes: samples unused
  const float* __restrict__ ga = a.g_mag ? a.g_mag + sig * a.frames * BINS : nullptr;
  const float* __restrict__ gb = a.g_phase ? a.g_phase + sig * a.frames * BINS : nullptr;
  // saved planes (SAVED, polar): the forward's |X| and arg X of this signal
  const float* __restrict__ pmg = SAVED && MODE == kOutPolar ? a.p_mag + sig * a.frames * BINS : nullptr;
  const float* __restrict__ pph = SAVED && MODE == kOutPolar ? a.p_phase + sig * a.frames * BINS : nullptr;
  float* __restrict__ gx = a.grad + sig * a.len;
  const long long frames = (long long)a.frames;
  const long long len = (long long)a.len;
  const long long hop = a.hop;
  const int W = a.window;

  float2 twr[E * NP];
  StockhamPlan<L, E, -1, Rs...>::load_twiddles(twr, t, a.tw, 2);
  float2 wsp[JF];
#pragma unroll
  for (int j = 0; j < JF; ++j) wsp[j] = load_tw<-1>(a.tw, t + j * T);
  const float2 wmid = load_tw<-1>(a.tw, L / 2);
  stage_window<N>(wl, a.win);
  if constexpr (SAVED) {
    // the forward's zero padding: taps past W reach no sample
    for (int i = a.window + tid; i < N; i += 256) wl[i] = 0.0f;
  }
  for (int i = tid; i < ACC; i += 256) acc[i] = 0.0f;
  __syncthreads();

  const int halo = W > (int)hop ? W - (int)hop : 0;  // samples shared with the previous run
  auto frame_valid = [&](long long f) { return f >= 0 && f < frames; };
  // samples of the first step's frame (recomputing) or its plane values (saved); later
  // steps' are prefetched a step ahead
  float2 nxt[SAVED ? 1 : E];
  constexpr bool kPolarSaved = SAVED && MODE == kOutPolar;
  float sa[SAVED ? 2 * JF + 1 : 1], sb[SAVED ? 2 * JF + 1 : 1];
  float sm[kPolarSaved ? 2 * JF + 1 : 1], sp[kPolarSaved ? 2 * JF + 1 : 1];
  const bool has_a = ga != nullptr, has_b = gb != nullptr;
  // plane values of frame f (index clamped so the loads are unconditional; a missing
  // gradient plane reads the other one, and its values are not used; the launcher never
  // runs this kernel with both missing)
  auto fetch_saved = [&](long long f) {
    const size_t row = (size_t)(f < 0 ? 0 : (f >= frames ? frames - 1 : f)) * BINS;
    const float* __restrict__ pa = (has_a ? ga : gb) + row;
    const float* __restrict__ pb = (has_b ? gb : ga) + row;
#pragma unroll
    for (int j = 0; j < JF; ++j) {
      const int p = t + j * T;
      sa[2 * j] = pa[p];
      sa[2 * j + 1] = pa[L - p];
      sb[2 * j] = pb[p];
      sb[2 * j + 1] = pb[L - p];
    }
    sa[2 * JF] = pa[L / 2];
    sb[2 * JF] = pb[L / 2];
    if constexpr (kPolarSaved) {
#pragma unroll
      for (int j = 0; j < JF; ++j) {
        const int p = t + j * T;
        sm[2 * j] = pmg[row + p];
        sm[2 * j + 1] = pmg[row + L - p];
        sp[2 * j] = pph[row + p];
        sp[2 * j + 1] = pph[row + L - p];
      }
      sm[2 * JF] = pmg[row + L / 2];
      sp[2 * JF] = pph[row + L / 2];
    }
  };
  // V of plane values (ga, gb, |X|, arg X); bins 0 and L take no Gb term
  auto v_saved = [&](int i, bool real_bin) -> float2 {
    const float g1 = has_a ? sa[i] : 0.0f;
    const float g2 = has_b && !real_bin ? sb[i] : 0.0f;
    if constexpr (kPolarSaved) {
      const float m = sm[i];
      if (m == 0.0f) return make_float2(0.0f, 0.0f);
      // hardware sine and cosine: arg X lies in [-pi, pi], where their error (a few 1e-7
      // of the unit direction) stays far inside the gradient's bound
      float sn, cs;
      __sincosf(sp[i], &sn, &cs);
      const float q = g2 / m;
      return make_float2(cs * g1 - sn * q, sn * g1 + cs * q);
    } else {
      return make_float2(g1, g2);
    }
  };
  {
    const long long f = f0 + slot;
    if constexpr (SAVED)
      fetch_saved(f);
    else
      fetch_frame<L, E, R0>(nxt, t, x, a.len, frame_valid(f) ? (size_t)(f * hop) : 0, W, frame_valid(f));
  }
  // Final samples of the previous step wait in the ring until the next step has used its
  // gradients; they are stored then, a whole transform ahead of the next wait for loads
  // (with stores pending, a wait for a load is a wait for every access in flight).
  long long pj0 = 0;
  int pcnt = 0, plh = 0;
  auto flush = [&]() {
    if (pcnt <= 0) return;
    const int prb = (int)(pj0 & (ACC - 1));
    float* __restrict__ out = gx + pj0;
    float* __restrict__ hl = a.halo + ((size_t)(sig * a.tiles + tix) * 2 + 1) * halo + (pj0 - f0 * hop);
    for (int u = tid; u < pcnt; u += 256) {
      const int ri = (prb + u) & (ACC - 1);
      if (u < plh)
        hl[u] = acc[ri];
      else
        out[u] = acc[ri];
      acc[ri] = 0.0f;
    }
  };
  for (long long step = 0; step < steps; ++step) {
    const long long fbase = f0 + step * FPS;
    const long long f = fbase + slot;
    const bool valid = frame_valid(f);  // invalid groups run on zeros (barriers stay uniform)

    // conj(Y_p) and conj(Y_{L-p}) into buf from V_p and V_{L-p} (thread t owns the pair)
    auto put_pair = [&](int j, int p, float2 Vp, float2 Vq) {
      if (p == 0) {
        buf[pidx(0)] = make_float2(Vp.x + Vq.x, -(Vp.x - Vq.x));
      } else {
        const float2 ep = c_conj(wsp[j]);  // e^{+2 pi i p/N}
        const float2 sv = c_scale(c_add(Vp, c_conj(Vq)), 0.5f);
        const float2 dv = c_scale(c_sub(Vp, c_conj(Vq)), 0.5f);
        const float2 ed = c_mul(ep, dv);
        const float2 Yp = make_float2(sv.x - ed.y, sv.y + ed.x);  // s + i e d
        // Y_{L-p} with e' = -conj(e): (V_q + conj V_p)/2 + i e' (V_q - conj V_p)/2
        const float2 s2 = c_conj(sv);
        const float2 d2 = c_scale(c_conj(dv), -1.0f);
        const float2 ed2 = c_mul(make_float2(-ep.x, ep.y), d2);
        const float2 Yq = make_float2(s2.x - ed2.y, s2.y + ed2.x);
        buf[pidx(p)] = c_conj(Yp);
        buf[pidx(L - p)] = c_conj(Yq);
      }
    };
    // 1. r_f into buf
    float2 v[E];
    float2* fb;
    if constexpr (SAVED) {
      // V from the prefetched plane values, then Y as below (invalid frames: V = 0)
      const float keep = valid ? 1.0f : 0.0f;
#pragma unroll
      for (int j = 0; j < JF; ++j) {
        const int p = t + j * T;
        const float2 Vp = c_scale(v_saved(2 * j, p == 0), keep);
        const float2 Vq = c_scale(v_saved(2 * j + 1, p == 0), keep);  // V_{L-p}
        put_pair(j, p, Vp, Vq);
      }
      if (t == 0) buf[pidx(L / 2)] = c_scale(v_saved(2 * JF, false), keep);
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = nxt[e];
      window_frame<L, E, R0>(v, t, wl);
      // Wait here for the samples (and the stores issued with them, a transform ago), so
      // that the gradient loads below are not part of that wait: s_waitcnt vmcnt(0)
      // expcnt(7) lgkmcnt(15).
      __builtin_amdgcn_s_waitcnt(0x0F70);
      __builtin_amdgcn_sched_barrier(0);
      // upstream gradients of this frame's bins, requested before the transform (a clamped
      // row: an invalid frame's spectrum is zero, or its r_f is never added)
      const size_t row = (size_t)(valid ? f : (f < 0 ? 0 : frames - 1)) * BINS;
      float gAp[JF], gAq[JF], gBp[JF], gBq[JF];
#pragma unroll
      for (int j = 0; j < JF; ++j) {
        const int p = t + j * T;
        gAp[j] = ga ? ga[row + p] : 0.0f;
        gAq[j] = ga ? ga[row + L - p] : 0.0f;
        gBp[j] = gb ? gb[row + p] : 0.0f;
        gBq[j] = gb ? gb[row + L - p] : 0.0f;
      }
      const float gAm = ga ? ga[row + L / 2] : 0.0f;
      const float gBm = gb ? gb[row + L / 2] : 0.0f;
      if constexpr (T == 64) {  // the frame's own wave: wave-local exchanges
        WaveStockham<L, E, -1, 1, 0, Rs...>::run(v, t, buf, twr, [](int) {});
        store_natural<L, E, Rl>(v, t, buf);
        wave_lds_fence();
      } else {
        StockhamPlan<L, E, -1, Rs...>::template run<true>(v, t, buf, buf, twr, fb);
        store_natural<L, E, Rl>(v, t, buf);
        __syncthreads();
      }
#pragma unroll
      for (int j = 0; j < JF; ++j) {
        const int p = t + j * T;
        const float2 zp = buf[pidx(p)];
        const float2 zc = c_conj(buf[pidx((L - p) & (L - 1))]);
        const float2 e = c_scale(c_add(zp, zc), 0.5f);
        const float2 d = c_sub(zp, zc);
        const float2 wo = c_mul(wsp[j], make_float2(0.5f * d.y, -0.5f * d.x));
        float2 X1 = c_add(e, wo);
        float2 X2 = c_conj(c_sub(e, wo));
        if (p == 0) {  // bins 0 and L are real
          X1.y = 0.0f;
          X2.y = 0.0f;
        }
        const float2 Vp = grad_spectrum<MODE>(X1, gAp[j], gBp[j], p == 0);
        const float2 Vq = grad_spectrum<MODE>(X2, gAq[j], gBq[j], p == 0);  // V_{L-p}
        put_pair(j, p, Vp, Vq);
      }
      if (t == 0) {  // p = L/2: e = i, so Y = Re V + i * i * (i Im V) = conj(V): conj(Y) = V
        const float2 zm = buf[pidx(L / 2)];
        const float2 zc = c_conj(zm);
        const float2 e = c_scale(c_add(zm, zc), 0.5f);
        const float2 d = c_sub(zm, zc);
        const float2 Xm = c_add(e, c_mul(wmid, make_float2(0.5f * d.y, -0.5f * d.x)));
        buf[pidx(L / 2)] = grad_spectrum<MODE>(Xm, gAm, gBm, false);
      }
    }
    // every gradient load has been used (lane 0 of a frame also used the middle bin's):
    // waiting here costs nothing and keeps the stores below out of later waits
    __builtin_amdgcn_s_waitcnt(0x0F70);
    flush();  // the previous step's final samples (ring entries no one reads before the next barrier)
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q * R0 + r] = buf[pidx(t + q * T + r * (L / R0))];
    if constexpr (T == 64) {  // the frame's own wave reads and scatters its buffer
      wave_lds_fence();
      WaveStockham<L, E, -1, 1, 0, Rs...>::run(v, t, buf, twr, [](int) {});
    } else {
      __syncthreads();  // every value read before the first in-place scatter
      StockhamPlan<L, E, -1, Rs...>::template run<true>(v, t, buf, buf, twr, fb);
    }
    // y = conj(result): r[2m] = v.x, r[2m+1] = -v.y
#pragma unroll
    for (int q = 0; q < E / Rl; ++q)
#pragma unroll
      for (int r = 0; r < Rl; ++r) {
        const float2 y = v[q * Rl + r];
        buf[pidx(t + q * T + r * (L / Rl))] = make_float2(y.x, -y.y);
      }
    if (step + 1 < steps) {  // the next step's samples or plane values, in flight during steps 2 and 3
      const long long fn = f + FPS;
      if constexpr (SAVED)
        fetch_saved(fn);
      else
        fetch_frame<L, E, R0>(nxt, t, x, a.len, frame_valid(fn) ? (size_t)(fn * hop) : 0, W, frame_valid(fn));
    }
    __syncthreads();

    // 2. accumulate the step's frames, ascending, over the samples they touch
    const long long j0 = fbase * hop;
    const int ulo = j0 < 0 ? (int)(-j0) : 0;  // the step's samples clipped to [0, len)
    const int rb = (int)(j0 & (ACC - 1));     // ring position of sample j0
    {
      const int span = (FPS - 1) * (int)hop + W;
      const long long room = len - j0;
      const int uhi = room < span ? (int)room : span;
      bool vg[FPS];
#pragma unroll
      for (int g = 0; g < FPS; ++g) vg[g] = fbase + g >= 0 && fbase + g < frames;
      if ((hop & 1) == 0) {
        // even hop: sample pairs (u, u+1), u even, sit in one float2 of every frame buffer
        // and of the ring (j0 and rb are even)
        float2* acc2 = reinterpret_cast<float2*>(acc);
        const float2* wl2 = reinterpret_cast<const float2*>(wl);
        for (int u = ulo + 2 * tid; u < uhi; u += 512) {
          const int ri = ((rb + u) & (ACC - 1)) >> 1;
          float2 sum = acc2[ri];
          const bool both = u + 1 < uhi;
#pragma unroll
          for (int g = 0; g < FPS; ++g) {
            const int n = u - g * (int)hop;
            if (vg[g] && n >= 0 && n < W) {
              const float2 c = fbuf[g][pidx(n >> 1)];
              const float2 w = wl2[n >> 1];
              sum.x = __builtin_fmaf(c.x, w.x, sum.x);
              if (n + 1 < W) sum.y = __builtin_fmaf(c.y, w.y, sum.y);
            }
          }
          if (both) {
            acc2[ri] = sum;
          } else {
            acc[ri << 1] = sum.x;
          }
        }
      } else {
        for (int u = ulo + tid; u < uhi; u += 256) {
          const int ri = (rb + u) & (ACC - 1);
          float sum = acc[ri];
#pragma unroll
          for (int g = 0; g < FPS; ++g) {
            const int n = u - g * (int)hop;
            if (vg[g] && n >= 0 && n < W) {
              const float2 c = fbuf[g][pidx(n >> 1)];
              sum = __builtin_fmaf((n & 1) ? c.y : c.x, wl[n], sum);
            }
          }
          acc[ri] = sum;
        }
      }
    }
    __syncthreads();

    // 3. samples [fbase*H, (fbase+FPS)*H) have all of this run's frames: they are stored
    //    by the next step's flush (the first `halo` samples of a run after the first go
    //    to the left halo slot), which also clears their ring entries
    {
      const long long room = len - j0;
      pj0 = j0;
      pcnt = room < FPS * (int)hop ? (int)(room > 0 ? room : 0) : FPS * (int)hop;
      plh = tix > 0 ? (int)(f0 * hop + halo - j0) : 0;  // u below plh: halo
    }
  }
  flush();

  // right halo: the samples past the run that its last frames touch
  const long long jr = fend * hop;
  const bool right = tix + 1 < a.tiles && jr < len && halo > 0;
  if (right) {
    float* __restrict__ hr = a.halo + (size_t)(sig * a.tiles + tix + 1) * 2 * halo;
    for (int i = tid; i < halo; i += 256) hr[i] = acc[(int)((jr + i) & (ACC - 1))];
  }
}

// dL/dx = 0 (an adjoint call with no upstream gradient).
__global__ void __launch_bounds__(256) zero_grad_kernel(float* __restrict__ out, size_t count) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) out[i] = 0.0f;
}

// Halo samples of every run boundary: previous run's partial sum + this run's. One
// workgroup per boundary, boundaries of all signals in turn.
__global__ void __launch_bounds__(256) stft_halo_kernel(SignalArgs a, int halo, size_t run_len) {
  const size_t b = blockIdx.x;  // boundary b of signal sig: between runs r-1 and r
  const size_t sig = b / (size_t)(a.tiles - 1);
  const size_t r = b - sig * (a.tiles - 1) + 1;
  const size_t jb = r * run_len;
  const float* pr = a.halo + ((sig * a.tiles + r) * 2 + 0) * (size_t)halo;
  float* __restrict__ g = a.grad + sig * a.len + jb;
  const int n = jb >= a.len ? 0 : (jb + halo <= a.len ? halo : (int)(a.len - jb));
  for (int i = threadIdx.x; i < n; i += 256) g[i] = pr[i] + pr[halo + i];
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

// Sizes the grid to the chip: each workgroup runs `steps` steps, chosen so that the
// workgroups of all signals fill the resident capacity (CUs x workgroups per CU) about
// once. Longer runs cost fewer warm-up frames; more workgroups hide more latency.
// Plans the adjoint launch of `a` (steps per run, runs per signal, halo slots) and
// launches it; with `need`, only reports the halo floats that launch would use.
template <int L, int E, int ACCM, int... Rs>
hipError_t adjoint_launch_acc(SignalArgs a, int mode, hipStream_t s, size_t* need, bool saved) {
  constexpr int FPS = 256 / (L / E);
  const void* kern = nullptr;
  if (saved) {
    switch (mode) {
      case kOutComplex:
        kern = reinterpret_cast<const void*>(&stft_adjoint_kernel<L, E, kOutComplex, ACCM, true, Rs...>);
        break;
      case kOutPolar:
        kern = reinterpret_cast<const void*>(&stft_adjoint_kernel<L, E, kOutPolar, ACCM, true, Rs...>);
        break;
      default: return hipErrorInvalidValue;  // magnitude planes hold no direction of X
    }
  } else {
    switch (mode) {
      case kOutComplex:
        kern = reinterpret_cast<const void*>(&stft_adjoint_kernel<L, E, kOutComplex, ACCM, false, Rs...>);
        break;
      case kOutMagnitude:
        kern = reinterpret_cast<const void*>(&stft_adjoint_kernel<L, E, kOutMagnitude, ACCM, false, Rs...>);
        break;
      case kOutPolar:
        kern = reinterpret_cast<const void*>(&stft_adjoint_kernel<L, E, kOutPolar, ACCM, false, Rs...>);
        break;
      default: return hipErrorInvalidValue;
    }
  }
  int dev = 0, cus = 0, occ = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 256, 0);
  if (e != hipSuccess) return e;
  const size_t capacity = (size_t)(cus > 0 ? cus : 1) * (size_t)(occ > 0 ? occ : 1);
  const size_t positions = (a.len + a.hop - 1) / a.hop;  // frame positions covering every sample
  const size_t per_signal = (capacity + a.signals - 1) / a.signals;
  size_t steps = (positions + FPS * per_signal - 1) / (FPS * per_signal);
  // a run must cover the W-H samples it shares with its neighbour
  const size_t halo = a.window > a.hop ? (size_t)(a.window - a.hop) : 0;
  const size_t min_steps = (halo + (size_t)FPS * a.hop - 1) / ((size_t)FPS * a.hop);
  steps = steps < 4 ? 4 : (steps > 64 ? 64 : steps);
  if (steps < min_steps) steps = min_steps;
  a.steps = (int)steps;
  a.tiles = (int)((positions + FPS * steps - 1) / (FPS * steps));
  const size_t grid = a.signals * (size_t)a.tiles;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  const bool halos = halo > 0 && a.tiles > 1;
  if (need) {
    *need = halos ? grid * 2 * halo : 0;
    return hipSuccess;
  }
  if (halos && (!a.halo || grid * 2 * halo > a.halo_capacity)) return hipErrorInvalidValue;
  void* args[] = {&a};
  e = hipLaunchKernel(kern, dim3((unsigned)grid), dim3(256), args, 0, s);
  if (e != hipSuccess || !halos) return e;
  const size_t hgrid = a.signals * (size_t)(a.tiles - 1);
  if (hgrid > 0x7fffffffull) return hipErrorInvalidValue;
  const size_t run_len = (size_t)FPS * steps * (size_t)a.hop;
  hipLaunchKernelGGL(stft_halo_kernel, dim3((unsigned)hgrid), dim3(256), 0, s, a, (int)halo, run_len);
  return hipGetLastError();
}

// The smallest accumulator ring (2N, 4N or 8N samples) that holds one step's span.
template <int L, int E, int... Rs>
hipError_t adjoint_launch(const SignalArgs& a, int span, int mode, hipStream_t s, size_t* need, bool saved) {
  if (span <= 4 * L) return adjoint_launch_acc<L, E, 2, Rs...>(a, mode, s, need, saved);
  if (span <= 8 * L) return adjoint_launch_acc<L, E, 4, Rs...>(a, mode, s, need, saved);
  return adjoint_launch_acc<L, E, 8, Rs...>(a, mode, s, need, saved);
}

}  // namespace

// Frames per adjoint step for each N (256 / threads per frame). N = 2048 runs a frame on
// two waves (8 values per lane, three workgroups per CU): one wave per frame (16 values
// per lane, 253 registers, two workgroups per CU) took the saved-plane adjoint of 32 x 2^18
// samples from 87-89 us to 94 us.
int adjoint_fps(int n) { return n == 2048 ? 2 : 4; }

namespace {
// Scratch of the allocation-free adjoint entry (fft_stft_adjoint_device_noalloc): its
// halo slots, kNoAllocHaloFloats floats (16 MiB), one copy on every GPU the library runs
// on, held until the process ends. Every call on a GPU shares that GPU's copy.
constexpr size_t kNoAllocHaloFloats = size_t(4) << 20;
__device__ __align__(256) float scratch[kNoAllocHaloFloats];
constexpr int kMaxDevices = 64;
}  // namespace

float* noalloc_halo_scratch(size_t* capacity) {
  // `scratch` names a host-side stand-in: the device address comes from
  // hipGetSymbolAddress, looked up once per device with that device current
  static std::mutex mu;
  static float* per_device[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!per_device[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    per_device[dev] = static_cast<float*>(p);
  }
  *capacity = kNoAllocHaloFloats;
  return per_device[dev];
}

size_t adjoint_scratch(int n, int window, int hop, size_t signals, size_t pcm_len) {
  if (hop <= 0 || window <= hop) return 0;
  // at least 4 steps per run: an upper bound on the runs the launcher makes
  const size_t positions = (pcm_len + hop - 1) / hop;
  const size_t per = (size_t)adjoint_fps(n) * 4;
  const size_t runs = signals * ((positions + per - 1) / per);
  return runs * 2 * (size_t)(window - hop);
}

bool signals_supported(int n, int window, int hop, bool adjoint) {
  if (n != 512 && n != 1024 && n != 2048) return false;
  if (window <= 0 || window > n || hop <= 0) return false;
  // the adjoint's accumulator ring (at most 8N samples) must hold one step's span
  return !adjoint || (long long)window + (long long)adjoint_fps(n) * hop <= 8LL * n;
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

hipError_t launch_stft_adjoint(SignalArgs a, int mode, hipStream_t s, size_t* halo_need, bool saved) {
  const int n = a.n;
  if (!signals_supported(n, a.window, a.hop, true) || a.len == 0 || !a.grad) return hipErrorInvalidValue;
  if (saved && mode == kOutPolar && (!a.p_mag || !a.p_phase)) return hipErrorInvalidValue;
  if (saved && !a.g_mag && !a.g_phase) {
    // no upstream gradient: the saved-plane kernel would have no plane to read its
    // gradient values from, and dL/dx is 0
    if (halo_need) {
      *halo_need = 0;
      return hipSuccess;
    }
    const size_t total = a.signals * a.len;
    size_t grid = (total + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(zero_grad_kernel, dim3((unsigned)grid), dim3(256), 0, s, a.grad, total);
    return hipGetLastError();
  }
  const int fps = adjoint_fps(n);
  const int span = a.window + fps * a.hop;
  // recomputing, with no phase gradient: polar V = Ga X/|X| is the magnitude map (fewer
  // registers)
  if (!saved && mode == kOutPolar && !a.g_phase) mode = kOutMagnitude;
  switch (n) {
    case 512: return adjoint_launch<256, 4, 4, 4, 4, 4>(a, span, mode, s, halo_need, saved);
    case 1024: return adjoint_launch<512, 8, 8, 8, 8>(a, span, mode, s, halo_need, saved);
    default: return adjoint_launch<1024, 8, 8, 8, 8, 2>(a, span, mode, s, halo_need, saved);
  }
}

}  // namespace fftfree

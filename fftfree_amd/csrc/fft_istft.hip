// fft_istft.hip -- fft_execute_complex_batched on the device: frame-major complex bins
// in, one transform per frame, real part times the synthesis window out, either frame by
// frame or overlap-added (with optional COLA normalisation).
//
// Reference semantics (fft_cffi.cpp:1042-1157):
//  * column rebuild (1062-1098): bins 0..bins-1 from the planes; for a C2R half-spectrum
//    context, imag(X_0) = imag(X_{N/2}) = 0 and X_{N-k} = conj X_k for 0 < k < N/2; else
//    zeros above the given bins;
//  * the context's plan with its hooks (pre/post eigen_fft.hpp:2484-2596; inverse 1/N);
//  * s = Re(x_f[i]) * ws[i] (ws only when windows are on), then either out[f*N + i] = s,
//    or out = 0 over frames*N samples and out[f*H + i] += s for f*H + i < frames*N, in
//    ascending frame order; with COLA normalisation, norm[t] += wa[i]*ws[i] (1 when
//    windows are off, cola_window_weight 210-226) alongside, and out[t] /= norm[t]
//    wherever norm[t] > 1e-12.
//
// Fused kernel (istft_kernel) for an inverse C2R half-spectrum plan with N = 2L in
// {512, 1024, 2048} and H <= N: x = IDFT_N(X) is real, so with e_p = e^{+2 pi i p/N}
//   2 Z_p = (X_p + conj X_{L-p}) + i e_p (X_p - conj X_{L-p})   (0 < p < L),
//   2 Z_0 = (Re X_0 + Re X_L) + i (Re X_0 - Re X_L),   2 Z_{L/2} = 2 conj X_{L/2},
// and the unscaled L-point inverse y = IDFT(2Z) gives x[2m] = Re y_m / N,
// x[2m+1] = Im y_m / N. The inverse runs as conj(DFT(conj 2Z)) on the forward engine.
// Only Re X_0 and Re X_L enter, which is the rebuild's imag(X_0) = imag(X_{N/2}) = 0.
//
// A workgroup owns a run of frame positions and overlap-adds its frames, in ascending
// order, in an LDS ring; a sample is written once, when no later frame touches it. The
// ceil(N/H) - 1 frames before a run that reach into it are transformed again by that
// run (warm-up steps whose samples are not written), so runs need no exchange and the
// whole result, COLA division included, comes from one launch. The samples past the
// last frame, up to frames*N, are zero-filled by the same launch.
#include "fft_device.hpp"
#include "fft_kernels.hpp"

namespace fftfree {
using namespace dev;

namespace {

constexpr int kFps = 4;  // frames per step: one 64-lane wave per frame
#ifndef FFTFREE_ISTFT_FILL
#define FFTFREE_ISTFT_FILL 100
#endif
// Workgroups per 100 resident slots of the chip: more means shorter runs (more warm-up
// frames transformed twice), fewer means less latency hiding.
constexpr int kIstftFill = FFTFREE_ISTFT_FILL;

typedef float f32x2 __attribute__((ext_vector_type(2)));

enum IstftMode : int { kFrameMajor = 0, kOverlap = 1, kOverlapCola = 2 };

// COLA weight of frame tap n (fft_cffi.cpp:210-226).
__device__ __forceinline__ float cola_weight(const IstftArgs& a, int n) {
  if (a.unit_cola) return 1.0f;
  const float wa = a.awin ? a.awin[n] : 1.0f;
  const float ws = a.swin ? a.swin[n] : 1.0f;
  return wa * ws;
}

// m * (cos, sin)(atan2(y, x)), the Griffin-Lim phase update (phase_infer.hpp:187-191),
// as m * (x, y) / |(x, y)|; for x = y = 0 the values atan2 gives for signed zeros
// (angle 0 or +-pi; sin(pi) in float is -8.742278e-8).
__device__ __forceinline__ float2 gl_dir(float m, float x, float y) {
  if (x == 0.0f && y == 0.0f) {
    if (!__builtin_signbit(x)) return make_float2(m, __builtin_copysignf(0.0f, y) * m);
    return make_float2(-m, (__builtin_signbit(y) ? 8.742278e-8f : -8.742278e-8f) * m);
  }
  // 1/|(x, y)|: the hardware reciprocal square root where x^2 + y^2 stays a normal float
  // (about 1 ulp, as 1/hypotf), else the scaled hypotf
  const float s2 = __builtin_fmaf(x, x, y * y);
  const float inv = (s2 > 1e-30f && s2 < 1e30f) ? __builtin_amdgcn_rsqf(s2) : 1.0f / hypotf(x, y);
  return make_float2(m * (x * inv), m * (y * inv));
}

// The float4-aligned part of out[from, to) that rank `rank` of `ranks` zero-fills, in
// float4 units from `base4`; the unaligned head and tail floats go to the first and last
// rank (zero_edges).
struct ZeroSpan {
  f32x4* base4 = nullptr;
  size_t b = 0, e = 0;
};
__device__ __forceinline__ ZeroSpan zero_span(float* __restrict__ out, size_t from, size_t to, int rank, int ranks) {
  ZeroSpan z;
  if (from >= to) return z;
  float* base = out + from;
  const size_t len = to - from;
  size_t head = ((16 - ((uintptr_t)base & 15)) & 15) >> 2;
  if (head > len) head = len;
  const size_t n4 = (len - head) >> 2;
  const size_t per = (n4 + ranks - 1) / ranks;
  z.base4 = reinterpret_cast<f32x4*>(base + head);
  z.b = (size_t)rank * per;
  z.e = z.b + per < n4 ? z.b + per : n4;
  if (z.b > z.e) z.b = z.e;
  return z;
}
__device__ __forceinline__ void zero_edges(float* __restrict__ out, size_t from, size_t to, int rank, int ranks) {
  if (from >= to) return;
  float* base = out + from;
  const size_t len = to - from;
  size_t head = ((16 - ((uintptr_t)base & 15)) & 15) >> 2;
  if (head > len) head = len;
  const size_t n4 = (len - head) >> 2;
  const size_t rest = (len - head) & 3;
  if (rank == 0 && threadIdx.x < head) base[threadIdx.x] = 0.0f;
  if (rank == ranks - 1 && threadIdx.x < rest) base[head + 4 * n4 + threadIdx.x] = 0.0f;
}

template <int L, int E, int MODE, int ACCM, bool GLM, int... Rs>
__global__ void __launch_bounds__(256) istft_kernel(IstftArgs a) {
  constexpr int T = L / E;
  constexpr int N = 2 * L;
  constexpr int NP = sizeof...(Rs);
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int LP = padded_len<L>();
  constexpr int JF = (L / 2) / T;
  constexpr int BINS = L + 1;
  constexpr bool OLA = MODE != kFrameMajor;
  constexpr bool COLA = MODE == kOverlapCola;
  constexpr int ACC = OLA ? ACCM * N : 4;  // ring; the host checks (FPS-1)*H + N <= ACC
  static_assert(T == 64 && JF * T == L / 2, "one frame per wave");
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 fbuf[kFps][LP];
  __shared__ __attribute__((aligned(16))) float acc[ACC];
  __shared__ float wl[N];                // synthesis taps / N
  __shared__ float nhr[COLA ? N : 1];    // 1 / COLA sum per sample residue mod H (1 where <= 1e-12)
  __shared__ float cw[COLA ? N : 1];     // COLA weights

  const int tid = threadIdx.x;
  const int slot = tid >> 6;
  const int t = tid & 63;
  float2* buf = &fbuf[slot][0];
  const int rank = xcd_tile(blockIdx.x, gridDim.x);
  const long long frames = (long long)a.frames;
  const int H = a.hop;
  const long long f0 = (long long)rank * kFps * a.steps;
  const bool vec4 = (((uintptr_t)a.out) & 15) == 0 && (H & 3) == 0;

  float2 twr[E * NP];
  StockhamPlan<L, E, -1, Rs...>::load_twiddles(twr, t, a.tw, 2);
  float2 ep[JF];  // e^{+2 pi i p/N}
#pragma unroll
  for (int j = 0; j < JF; ++j) ep[j] = load_tw<1>(a.tw, t + j * T);
  for (int i = tid; i < N; i += 256) wl[i] = (a.swin ? a.swin[i] : 1.0f) * (1.0f / (float)N);
  if constexpr (COLA)
    for (int i = tid; i < N; i += 256) cw[i] = cola_weight(a, i);
  if constexpr (OLA)
    for (int i = tid; i < ACC; i += 256) acc[i] = 0.0f;
  __syncthreads();
  if constexpr (COLA) {
    // sum of a sample whose frames are all present: taps r + kH, frames ascending (taps
    // descending), as the reference accumulates
    for (int r = tid; r < H; r += 256) {
      float sum = 0.0f;
      for (int n = r + ((N - 1 - r) / H) * H; n >= 0; n -= H) sum += cw[n];
      nhr[r] = sum > 1e-12f ? 1.0f / sum : 1.0f;
    }
    __syncthreads();
  }

  // bins of the frame a step ahead, in registers (the index is clamped so the loads are
  // unconditional; frames outside [0, frames) count as zero where the values are used)
  float xr[2 * JF], xi[2 * JF], mr, mi;
  float xm[GLM ? 2 * JF : 1], mm = 0.0f;  // Griffin-Lim: |X| of the same bins
  bool xvalid = false, xupd = false;
  auto fetch = [&](long long f) {
    xvalid = f >= 0 && f < frames;
    const long long fc = f < 0 ? 0 : (f >= frames ? frames - 1 : f);
    xupd = GLM && (size_t)fc < a.gl_frames;
    const float* __restrict__ re = (xupd ? a.gl_re : a.in_re) + (size_t)fc * BINS;
    const float* __restrict__ im = (xupd ? a.gl_im : a.in_im) + (size_t)fc * BINS;
#pragma unroll
    for (int j = 0; j < JF; ++j) {
      const int p = t + j * T;
      xr[2 * j] = re[p];
      xi[2 * j] = im[p];
      xr[2 * j + 1] = re[L - p];
      xi[2 * j + 1] = im[L - p];
    }
    mr = re[L / 2];
    mi = im[L / 2];
    if constexpr (GLM) {
      const float* __restrict__ mg = a.gl_mag + (size_t)fc * BINS;
#pragma unroll
      for (int j = 0; j < JF; ++j) {
        const int p = t + j * T;
        xm[2 * j] = mg[p];
        xm[2 * j + 1] = mg[L - p];
      }
      mm = mg[L / 2];
    }
  };

  const long long first = (OLA && f0 > 0) ? -(long long)a.warm : 0;
  // the zero tail past the overlap-add: this rank's share, spread over its steps
  ZeroSpan zs;
  size_t zper = 0;
  if constexpr (OLA) {
    zs = zero_span(a.out, a.ola_len, a.out_len, rank, gridDim.x);
    const size_t nsteps = (size_t)(a.steps - first);
    zper = (zs.e - zs.b + nsteps - 1) / nsteps;
    zero_edges(a.out, a.ola_len, a.out_len, rank, gridDim.x);
  }
  const f32x4 zero4 = {0.0f, 0.0f, 0.0f, 0.0f};

  // Final samples of the previous step wait in the ring, normalised; they are stored at
  // the start of the next step, ahead of its loads, so that no wait for those loads
  // also waits for fresh stores.
  long long pj0 = 0;
  int pfin = 0;
  bool pwrite = false;
  auto flush = [&]() {
    if (pfin <= 0) return;
    const int prb = (int)(pj0 & (ACC - 1));
    float* __restrict__ o = a.out + pj0;
    if (vec4) {
      const int n4 = pfin & ~3;
      for (int u = 4 * tid; u < n4; u += 1024) {
        f32x4* r = reinterpret_cast<f32x4*>(&acc[(prb + u) & (ACC - 1)]);
        const f32x4 v = *r;
        *r = zero4;
        // Griffin-Lim's forward STFT reads the signal straight back: ordinary stores leave
        // it in the caches; otherwise streaming stores
        if (pwrite) {
          if constexpr (GLM) *reinterpret_cast<f32x4*>(o + u) = v;
          else __builtin_nontemporal_store(v, reinterpret_cast<f32x4*>(o + u));
        }
      }
      if (tid < pfin - n4) {
        const int u = n4 + tid;
        float& r = acc[(prb + u) & (ACC - 1)];
        if (pwrite) o[u] = r;
        r = 0.0f;
      }
    } else {
      for (int u = tid; u < pfin; u += 256) {
        float& r = acc[(prb + u) & (ACC - 1)];
        if (pwrite) o[u] = r;
        r = 0.0f;
      }
    }
  };
  float2 pv[OLA ? 1 : E];  // frame-major: the previous step's frame, stored a step late
  fetch(f0 + first * kFps + slot);
  for (long long step = first; step < a.steps; ++step) {
    const long long fbase = f0 + step * kFps;
    const long long f = fbase + slot;
    // conj(2Z) into buf; thread t owns the pairs (p, L-p), lane 0 also L/2. Frame-major
    // output keeps a clamped frame's values: the wave then rewrites the last frame's
    // samples with the same values, so its stores need no condition.
    const float keep = (xvalid || !OLA) ? 1.0f : 0.0f;
    if constexpr (GLM) {
      if (xupd) {  // X = |X| dir(previous forward spectrum)
#pragma unroll
        for (int j = 0; j < 2 * JF; ++j) {
          const float2 z = gl_dir(xm[j], xr[j], xi[j]);
          xr[j] = z.x;
          xi[j] = z.y;
        }
        const float2 z = gl_dir(mm, mr, mi);
        mr = z.x;
        mi = z.y;
      }
    }
#pragma unroll
    for (int j = 0; j < JF; ++j) {
      const int p = t + j * T;
      const float2 Xp = make_float2(xr[2 * j] * keep, xi[2 * j] * keep);
      const float2 Xq = make_float2(xr[2 * j + 1] * keep, xi[2 * j + 1] * keep);
      if (p == 0) {
        buf[pidx(0)] = make_float2(Xp.x + Xq.x, Xq.x - Xp.x);
      } else {
        const float2 sv = make_float2(Xp.x + Xq.x, Xp.y - Xq.y);  // X_p + conj X_q
        const float2 dv = make_float2(Xp.x - Xq.x, Xp.y + Xq.y);  // X_p - conj X_q
        const float2 ed = c_mul(ep[j], dv);
        buf[pidx(p)] = make_float2(sv.x - ed.y, -(sv.y + ed.x));
        buf[pidx(L - p)] = make_float2(sv.x + ed.y, sv.y - ed.x);
      }
    }
    if (t == 0) buf[pidx(L / 2)] = make_float2(2.0f * keep * mr, 2.0f * keep * mi);
    // The next step's loads, then this step's stores: the wait for those loads (at the
    // next step's conj(2Z)) also covers every store pending at that point, so the stores
    // are issued a whole step ahead of it.
    fetch(f + kFps);  // unconditional, so the waits stay counted
    if constexpr (OLA) {
      // the previous step's final samples and a share of the zero tail
      flush();
      const size_t k = (size_t)(step - first);
      const size_t zb = zs.b + k * zper;
      const size_t ze = zb + zper < zs.e ? zb + zper : zs.e;
      for (size_t i = zb + tid; i < ze; i += 256) __builtin_nontemporal_store(zero4, &zs.base4[i]);
    } else {
      // the previous step's frame (a clamped frame index: see above)
      if (step > first) {
        const long long pf = f - kFps < frames ? f - kFps : frames - 1;
        f32x2* __restrict__ o = reinterpret_cast<f32x2*>(a.out + (size_t)pf * N);
        const float2* wl2 = reinterpret_cast<const float2*>(wl);
#pragma unroll
        for (int q = 0; q < E / Rl; ++q)
#pragma unroll
          for (int r = 0; r < Rl; ++r) {
            const int m = t + q * T + r * (L / Rl);
            const float2 y = pv[q * Rl + r];
            const float2 w = wl2[m];
            const f32x2 val = {y.x * w.x, -y.y * w.y};
            __builtin_nontemporal_store(val, &o[m]);
          }
      }
    }
    wave_lds_fence();
    float2 v[E];
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q * R0 + r] = buf[pidx(t + q * T + r * (L / R0))];
    wave_lds_fence();
    WaveStockham<L, E, -1, 1, 0, Rs...>::run(v, t, buf, twr, [](int) {});
    // v[q*Rl + r] = DFT(conj 2Z)[m], m = t + q*T + r*L/Rl; y = conj of it:
    // x[2m] = Re y_m / N, x[2m+1] = Im y_m / N (the 1/N is in wl)
    if constexpr (!OLA) {
#pragma unroll
      for (int e = 0; e < E; ++e) pv[e] = v[e];
      wave_lds_fence();  // buf is rewritten by the next step
    } else {
#pragma unroll
      for (int q = 0; q < E / Rl; ++q)
#pragma unroll
        for (int r = 0; r < Rl; ++r) {
          const float2 y = v[q * Rl + r];
          buf[pidx(t + q * T + r * (L / Rl))] = make_float2(y.x, -y.y);
        }
      __syncthreads();

      // Overlap-add of the step's frames over the samples they touch, ascending. The
      // first FPS*H samples get no later frame: they are final and stay in the ring,
      // normalised, until the next step stores them (unless this is a warm-up step).
      const long long j0 = fbase * H;
      const long long room = (long long)a.ola_len - j0;
      const int span = (kFps - 1) * H + N;
      const int uhi = room < span ? (int)(room > 0 ? room : 0) : span;
      const int fin = kFps * H;
      const int rb = (int)(j0 & (ACC - 1));
      bool vg[kFps];
#pragma unroll
      for (int g = 0; g < kFps; ++g) vg[g] = fbase + g >= 0 && fbase + g < frames;
      // every frame touching the step's final samples exists: the periodic sum applies
      const bool inner = fbase >= (long long)a.kd && fbase + kFps <= frames;
      auto edge_norm = [&](int u) -> float {
        float sum = 0.0f;
        for (int d = -a.kd; d < kFps; ++d) {  // frames ascending
          const int n = u - d * H;
          const long long fr = fbase + d;
          if (n >= 0 && n < N && fr >= 0 && fr < frames) sum += cw[n];
        }
        return sum;
      };
      auto finish = [&](float sum, int u, int r) -> float {
        if constexpr (COLA) {
          if (inner) return sum * nhr[r];
          const float nrm = edge_norm(u);
          return nrm > 1e-12f ? sum / nrm : sum;
        } else {
          return sum;
        }
      };
      if ((H & 1) == 0) {
        // even hop: sample pairs (u, u+1), u even, sit in one float2 of every frame
        // buffer and of the ring (j0 and rb are even)
        float2* acc2 = reinterpret_cast<float2*>(acc);
        const float2* wl2 = reinterpret_cast<const float2*>(wl);
        int r = COLA ? (2 * tid) % H : 0;  // residue of u mod H
        const int dr = COLA ? 512 % H : 0;
        for (int u = 2 * tid; u < uhi; u += 512) {
          const int ri = ((rb + u) & (ACC - 1)) >> 1;
          float2 sum = acc2[ri];
#pragma unroll
          for (int g = 0; g < kFps; ++g) {
            const int n = u - g * H;
            if (vg[g] && n >= 0 && n < N) {
              const float2 c = fbuf[g][pidx(n >> 1)];
              const float2 w = wl2[n >> 1];
              sum.x = __builtin_fmaf(c.x, w.x, sum.x);
              sum.y = __builtin_fmaf(c.y, w.y, sum.y);
            }
          }
          if (u < fin) {
            const int r1 = r + 1 == H ? 0 : r + 1;
            sum = make_float2(finish(sum.x, u, r), finish(sum.y, u + 1, r1));
          }
          acc2[ri] = sum;
          if (COLA) {
            r += dr;
            r = r >= H ? r - H : r;
          }
        }
      } else {
        int r = COLA ? tid % H : 0;
        const int dr = COLA ? 256 % H : 0;
        for (int u = tid; u < uhi; u += 256) {
          const int ri = (rb + u) & (ACC - 1);
          float sum = acc[ri];
#pragma unroll
          for (int g = 0; g < kFps; ++g) {
            const int n = u - g * H;
            if (vg[g] && n >= 0 && n < N) {
              const float2 c = fbuf[g][pidx(n >> 1)];
              sum = __builtin_fmaf((n & 1) ? c.y : c.x, wl[n], sum);
            }
          }
          acc[ri] = u < fin ? finish(sum, u, r) : sum;
          if (COLA) {
            r += dr;
            r = r >= H ? r - H : r;
          }
        }
      }
      pj0 = j0;
      pfin = uhi < fin ? uhi : fin;
      pwrite = step >= 0;
      __syncthreads();
    }
  }
  if constexpr (OLA) {
    flush();
  } else if (a.steps > first) {
    const long long fl = f0 + (a.steps - 1) * kFps + slot;
    const long long pf = fl < frames ? fl : frames - 1;
    f32x2* __restrict__ o = reinterpret_cast<f32x2*>(a.out + (size_t)pf * N);
    const float2* wl2 = reinterpret_cast<const float2*>(wl);
#pragma unroll
    for (int q = 0; q < E / Rl; ++q)
#pragma unroll
      for (int r = 0; r < Rl; ++r) {
        const int m = t + q * T + r * (L / Rl);
        const float2 y = pv[q * Rl + r];
        const float2 w = wl2[m];
        const f32x2 val = {y.x * w.x, -y.y * w.y};
        __builtin_nontemporal_store(val, &o[m]);
      }
  }
}

// General path, step 1: the column rebuild of fft_cffi.cpp:1062-1098, one thread per
// column entry (frames x N complex values).
__global__ void __launch_bounds__(256) complex_columns_kernel(IstftArgs a, int hermitian, float2* __restrict__ cols) {
  const size_t total = a.frames * (size_t)a.n;
  const size_t n = (size_t)a.n;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
    const size_t f = idx / n;
    const size_t i = idx - f * n;
    const size_t row = f * (size_t)a.bins;
    float2 z = make_float2(0.0f, 0.0f);
    if (hermitian) {
      const size_t ny = n / 2;
      if (i == 0 || i == ny) {
        z = make_float2(a.in_re[row + i], 0.0f);
      } else if (i < ny) {
        z = make_float2(a.in_re[row + i], a.in_im[row + i]);
      } else {
        z = make_float2(a.in_re[row + n - i], -a.in_im[row + n - i]);
      }
    } else if (i < (size_t)a.bins) {
      z = make_float2(a.in_re[row + i], a.in_im[row + i]);
    }
    cols[idx] = z;
  }
}

// General path, step 3: real parts of the transformed columns to the output, one thread
// per output sample (a gather, so the frame order of every sum is fixed).
__global__ void __launch_bounds__(256) overlap_kernel(IstftArgs a, const float2* __restrict__ cols) {
  const size_t n = (size_t)a.n;
  const size_t total = a.frames * n;
  const size_t H = (size_t)a.hop;
  for (size_t s = (size_t)blockIdx.x * 256 + threadIdx.x; s < total; s += (size_t)gridDim.x * 256) {
    if (!a.ola) {
      const size_t i = s & (n - 1);
      const float x = cols[s].x;
      a.out[s] = a.swin ? x * a.swin[i] : x;
      continue;
    }
    const size_t flo = s >= n ? (s - n) / H + 1 : 0;
    size_t fhi = s / H;
    if (fhi > a.frames - 1) fhi = a.frames - 1;
    float sum = 0.0f, nrm = 0.0f;
    for (size_t f = flo; f <= fhi; ++f) {
      const size_t i = s - f * H;
      const float x = cols[f * n + i].x;
      sum += a.swin ? x * a.swin[i] : x;
      if (a.cola) nrm += cola_weight(a, (int)i);
    }
    a.out[s] = (a.cola && nrm > 1e-12f) ? sum / nrm : sum;
  }
}

// Griffin-Lim start (phase_infer.hpp:124-134): the float uniform on [-pi, pi) that
// std::uniform_real_distribution<float> makes of one mt19937 word w: u = float(w) / 2^32
// (generate_canonical<float, 24>, with u = 1 taken to the float below 1), angle =
// u * (b - a) + a with one rounding (the reference is built with -ffast-math
// -march=native, CMakeLists.txt:16, which fuses it), then mag * (cos, sin).
__global__ void __launch_bounds__(256) gl_start_kernel(const float* __restrict__ mag,
                                                       const uint32_t* __restrict__ words, size_t count,
                                                       float* __restrict__ o_re, float* __restrict__ o_im) {
  constexpr float kPi = 3.14159265358979323846f;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) {
    float u = (float)words[i] * 2.3283064365386963e-10f;  // 2^-32
    if (u >= 1.0f) u = 0.99999994f;                      // nextafter(1, 0)
    const float ang = __builtin_fmaf(u, kPi - (-kPi), -kPi);
    float sn, cs;
    sincosf(ang, &sn, &cs);
    const float m = mag[i];
    o_re[i] = m * cs;
    o_im[i] = m * sn;
  }
}

// Griffin-Lim phase update (phase_infer.hpp:185-200).
__global__ void __launch_bounds__(256) gl_update_kernel(const float* __restrict__ mag, const float* __restrict__ t_re,
                                                        const float* __restrict__ t_im, size_t count, int bins,
                                                        int half, int n_even, float* __restrict__ o_re,
                                                        float* __restrict__ o_im) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) {
    const float2 z = gl_dir(mag[i], t_re[i], t_im[i]);
    const int k = (int)(i % (size_t)bins);
    const bool real_bin = half && (k == 0 || (n_even && bins > 1 && k == bins - 1));
    o_re[i] = z.x;
    o_im[i] = real_bin ? 0.0f : z.y;
  }
}

// Griffin-Lim overlap-add (phase_infer.hpp:153-173): each sample averages the frames that
// cover it, summed in ascending frame order.
__global__ void __launch_bounds__(256) gl_overlap_kernel(const float* __restrict__ fr, size_t frames, int n, int hop,
                                                         float* __restrict__ out, size_t ola_len) {
  const size_t N = (size_t)n, H = (size_t)hop;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < ola_len; t += (size_t)gridDim.x * 256) {
    const size_t flo = t >= N ? (t - N) / H + 1 : 0;
    size_t fhi = t / H;
    if (fhi > frames - 1) fhi = frames - 1;
    float sum = 0.0f, cnt = 0.0f;
    for (size_t f = flo; f <= fhi; ++f) {
      if (f * H + N > ola_len) break;
      sum += fr[f * N + (t - f * H)];
      cnt += 1.0f;
    }
    out[t] = cnt > 0.0f ? sum / cnt : sum;
  }
}

// The streaming inverse's overlap-add (fft_cffi.cpp:1602-1742). The reference adds each
// frame into an N-sample accumulator and then shifts it by e samples (e = min(hop, N));
// laid out flat, frame f lands at offset f*e of a buffer that starts with the accumulator
// carried in, so sample t is carry[t] (t < n) plus the frames covering it in ascending
// order, the same additions in the same order. With weights (COLA normalisation) the
// norm buffer gets the carried norm plus w[t - f*e] the same way.
__global__ void __launch_bounds__(256) stream_ola_kernel(const float* __restrict__ fr, size_t frames, int n, int e,
                                                         const float* __restrict__ carry,
                                                         const float* __restrict__ carry_norm,
                                                         const float* __restrict__ w, float* __restrict__ out,
                                                         float* __restrict__ out_norm, size_t len) {
  const size_t N = (size_t)n, E = (size_t)e;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < len; t += (size_t)gridDim.x * 256) {
    const size_t flo = t >= N ? (t - N) / E + 1 : 0;  // first frame with f*e + n > t
    size_t fhi = t / E;                               // last frame with f*e <= t
    if (fhi > frames - 1) fhi = frames - 1;
    float v = t < N ? carry[t] : 0.0f;
    float m = (w && t < N) ? carry_norm[t] : 0.0f;
    for (size_t f = flo; f <= fhi; ++f) {
      v += fr[f * N + (t - f * E)];
      if (w) m += w[t - f * E];
    }
    out[t] = v;
    if (w) out_norm[t] = m;
  }
}

template <int L, int E, int MODE, int ACCM, bool GLM, int... Rs>
hipError_t istft_launch_mode(IstftArgs a, hipStream_t s) {
  const void* kern = reinterpret_cast<const void*>(&istft_kernel<L, E, MODE, ACCM, GLM, Rs...>);
  int dev = 0, cus = 0, occ = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 256, 0);
  if (e != hipSuccess) return e;
  const size_t capacity = (size_t)(cus > 0 ? cus : 1) * (size_t)(occ > 0 ? occ : 1);
  // frame positions the runs cover: every frame, or every hop of the overlap-added output
  const size_t positions = MODE == kFrameMajor ? a.frames : (a.ola_len + a.hop - 1) / a.hop;
  const size_t slots = (capacity * kIstftFill + 99) / 100;
  size_t steps = (positions + kFps * slots - 1) / (kFps * slots);
  const size_t lo = MODE == kFrameMajor ? 1 : (size_t)(a.warm > 1 ? a.warm : 1);
  steps = steps < lo ? lo : (steps > 256 ? 256 : steps);
  a.steps = (int)steps;
  const size_t grid = (positions + kFps * steps - 1) / (kFps * steps);
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  a.tiles = (int)grid;
  void* args[] = {&a};
  return hipLaunchKernel(kern, dim3((unsigned)grid), dim3(256), args, 0, s);
}

template <int L, int E, int... Rs>
hipError_t istft_launch(const IstftArgs& a, hipStream_t s) {
  if (!a.ola) return istft_launch_mode<L, E, kFrameMajor, 2, false, Rs...>(a, s);
  const int span = (kFps - 1) * a.hop + 2 * L;
  if (a.gl_mag) {  // Griffin-Lim: overlap-add divided by the frame count
    if (!a.cola || !a.unit_cola) return hipErrorInvalidValue;
    if (span <= 4 * L) return istft_launch_mode<L, E, kOverlapCola, 2, true, Rs...>(a, s);
    return istft_launch_mode<L, E, kOverlapCola, 4, true, Rs...>(a, s);
  }
  if (a.cola) {
    if (span <= 4 * L) return istft_launch_mode<L, E, kOverlapCola, 2, false, Rs...>(a, s);
    return istft_launch_mode<L, E, kOverlapCola, 4, false, Rs...>(a, s);
  }
  if (span <= 4 * L) return istft_launch_mode<L, E, kOverlap, 2, false, Rs...>(a, s);
  return istft_launch_mode<L, E, kOverlap, 4, false, Rs...>(a, s);
}

}  // namespace

bool istft_fused_supported(int n, int hop, int ola) {
  if (n != 512 && n != 1024 && n != 2048) return false;
  return !ola || (hop >= 1 && hop <= n);
}

size_t istft_ola_len(size_t frames, int n, int hop) {
  if (frames == 0) return 0;
  const size_t total = frames * (size_t)n;
  const size_t span = (frames - 1) * (size_t)hop + (size_t)n;
  return span < total ? span : total;
}

hipError_t launch_istft(IstftArgs a, hipStream_t s) {
  if (!istft_fused_supported(a.n, a.hop, a.ola) || a.frames == 0 || !a.out) return hipErrorInvalidValue;
  if (a.ola) {
    a.ola_len = istft_ola_len(a.frames, a.n, a.hop);
    if (a.out_len == 0) a.out_len = a.frames * (size_t)a.n;
    if (a.out_len < a.ola_len) return hipErrorInvalidValue;
    a.kd = (a.n + a.hop - 1) / a.hop;              // frames that touch one sample, at most
    a.warm = (a.kd - 1 + kFps - 1) / kFps;          // warm-up steps before a run
  }
  switch (a.n) {
    case 512: return istft_launch<256, 4, 4, 4, 4, 4>(a, s);
    case 1024: return istft_launch<512, 8, 8, 8, 8>(a, s);
    default: return istft_launch<1024, 16, 16, 16, 4>(a, s);
  }
}

hipError_t launch_complex_columns(const IstftArgs& a, int hermitian, float2* cols, hipStream_t s) {
  const size_t total = a.frames * (size_t)a.n;
  if (total == 0) return hipSuccess;
  size_t grid = (total + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(complex_columns_kernel, dim3((unsigned)grid), dim3(256), 0, s, a, hermitian, cols);
  return hipGetLastError();
}

hipError_t launch_overlap(IstftArgs a, const float2* cols, hipStream_t s) {
  const size_t total = a.frames * (size_t)a.n;
  if (total == 0) return hipSuccess;
  if (a.ola && a.hop <= 0) return hipErrorInvalidValue;
  size_t grid = (total + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(overlap_kernel, dim3((unsigned)grid), dim3(256), 0, s, a, cols);
  return hipGetLastError();
}

hipError_t launch_gl_start(const float* mag, const uint32_t* words, size_t count, float* o_re, float* o_im,
                           hipStream_t s) {
  if (count == 0) return hipSuccess;
  size_t grid = (count + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(gl_start_kernel, dim3((unsigned)grid), dim3(256), 0, s, mag, words, count, o_re, o_im);
  return hipGetLastError();
}

hipError_t launch_gl_update(const float* mag, const float* t_re, const float* t_im, size_t frames, int bins,
                            int half, int n_even, float* o_re, float* o_im, hipStream_t s) {
  const size_t count = frames * (size_t)bins;
  if (count == 0) return hipSuccess;
  size_t grid = (count + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(gl_update_kernel, dim3((unsigned)grid), dim3(256), 0, s, mag, t_re, t_im, count, bins, half,
                     n_even, o_re, o_im);
  return hipGetLastError();
}

hipError_t launch_stream_ola(const float* frames_pcm, size_t frames, int n, int e, const float* carry,
                             const float* carry_norm, const float* w, float* out, float* out_norm, size_t len,
                             hipStream_t s) {
  if (len == 0 || frames == 0) return hipSuccess;
  if (n <= 0 || e <= 0 || e > n || len != (frames - 1) * (size_t)e + (size_t)n) return hipErrorInvalidValue;
  size_t grid = (len + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(stream_ola_kernel, dim3((unsigned)grid), dim3(256), 0, s, frames_pcm, frames, n, e, carry,
                     carry_norm, w, out, out_norm, len);
  return hipGetLastError();
}

hipError_t launch_gl_overlap(const float* frames_pcm, size_t frames, int n, int hop, float* out, size_t ola_len,
                             hipStream_t s) {
  if (ola_len == 0 || frames == 0) return hipSuccess;
  if (hop <= 0) return hipErrorInvalidValue;
  size_t grid = (ola_len + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(gl_overlap_kernel, dim3((unsigned)grid), dim3(256), 0, s, frames_pcm, frames, n, hop, out,
                     ola_len);
  return hipGetLastError();
}

}  // namespace fftfree

// fft_phase.hip -- phase inference from magnitude spectrograms (phase_infer_execute,
// fft_cffi.cpp:2072-2162 and phase_infer.hpp:23-66), frame-major planes of `bins` floats.
//
// Mode 0 (infer_linear, phase_infer.hpp:36-66): phi[f][k] is the double-precision running
// sum of dphi[k] = 2 pi kk hop / N over the frames before f (kk = k mod N), and
// out = (float)(m cos phi), (float)(m sin phi) in double. The running sum is reproduced
// exactly: one thread per (64-frame block, bin) finds phi at its first frame in closed
// form (running_sum, phase_sum.hpp: a constant step per binade) and walks on through
// segments of constant exact steps, rotating (cos phi, sin phi) by each step; the
// rotation's rounding (~64 * 2^-53) stays far below the float outputs' ulp.
//
// Mode 1 (minimum phase, fft_cffi.cpp:2095-2158): per frame the real-even log spectrum
// log(max(1e-20, m)) (mirrored for a half spectrum), its inverse DFT (the real cepstrum),
// the causal fold (samples 1 .. (N+1)/2-1 doubled, N/2+1 .. N-1 cleared), the forward DFT,
// and out = exp(Re L) (cos Im L, sin Im L). The DFTs run on the batched complex kernels;
// the three elementwise steps are here.
#include "fft_device.hpp"
#include "fft_kernels.hpp"
#include "phase_sum.hpp"

namespace fftfree {
using namespace dev;

namespace {

constexpr int kPhaseBlock = 64;  // frames per thread of the linear phase
constexpr int kPhaseChunk = 32;  // magnitudes loaded together

__device__ __forceinline__ double linear_dphi(int k, int n, int hop) {
  const int kk = k < n ? k : k % n;
  return 6.28318530717958647692 * (double)kk * (double)hop / (double)n;  // phase_infer.hpp:47-51
}

__global__ void __launch_bounds__(256) phase_linear_kernel(const float* __restrict__ mag, size_t frames, int bins,
                                                           int n, int hop, float* __restrict__ o_re,
                                                           float* __restrict__ o_im) {
  const size_t blocks = (frames + kPhaseBlock - 1) / kPhaseBlock;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= blocks * (size_t)bins) return;
  const size_t b = idx / (size_t)bins;
  const int k = (int)(idx - b * (size_t)bins);
  const double d = linear_dphi(k, n, hop);
  const size_t f0 = b * kPhaseBlock;
  const int count = frames - f0 < (size_t)kPhaseBlock ? (int)(frames - f0) : kPhaseBlock;
  double phi = running_sum(d, (uint64_t)f0);
  // Within a segment of constant exact steps (running_segment) cos and sin of the next
  // phase come from rotating by the step; at a segment's end the literal addition is
  // made and cos, sin are evaluated afresh. All lanes advance frame by frame together.
  double c = 1.0, s = 0.0, cd = 1.0, sd = 0.0, step = 0.0;
  uint64_t run = 0;
  bool seed = true;
  for (int i0 = 0; i0 < count; i0 += kPhaseChunk) {
    float mv[kPhaseChunk];  // this chunk's magnitudes, all loads issued before any is used
#pragma unroll
    for (int j = 0; j < kPhaseChunk; ++j) {
      const int i = i0 + j < count ? i0 + j : count - 1;
      mv[j] = mag[(f0 + i) * (size_t)bins + k];
    }
#pragma unroll
    for (int j = 0; j < kPhaseChunk; ++j) {
      const int i = i0 + j;
      if (i >= count) break;
      if (seed) {
        running_segment(phi, d, &step, &run);
        sincos(phi, &s, &c);
        sincos(step, &sd, &cd);
        seed = false;
      }
      const size_t at = (f0 + i) * (size_t)bins + k;
      const double m = (double)mv[j];
      o_re[at] = (float)(m * c);
      o_im[at] = (float)(m * s);
      if (run == 0) {
        phi = phi + d;
        seed = true;
      } else {
        const double c2 = c * cd - s * sd;
        s = s * cd + c * sd;
        c = c2;
        phi += step;
        --run;
      }
    }
  }
}

// Real-even log spectrum of every frame into N complex values per frame.
__global__ void __launch_bounds__(256) phase_log_kernel(const float* __restrict__ mag, size_t frames, int bins, int n,
                                                        int half, float2* __restrict__ cols) {
  const size_t total = frames * (size_t)n;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
    const size_t f = idx / (size_t)n;
    const int i = (int)(idx - f * (size_t)n);
    int k = i;
    if (half && i > n / 2) k = n - i;  // spec[N-k] = spec[k] (fft_cffi.cpp:2119-2126)
    // half spectrum: DC always, bins 1 .. N/2-1 and their mirrors, N/2 when N is even;
    // the rest stays 0 (fft_cffi.cpp:2113-2126)
    const bool set = !half || i == 0 || i < n / 2 || i > n - n / 2 || (i == n / 2 && (n % 2) == 0);
    const float m = mag[f * (size_t)bins + k];
    cols[idx] = make_float2(set ? logf(fmaxf(1e-20f, m)) : 0.0f, 0.0f);
  }
}

// Causal fold of the real cepstrum (fft_cffi.cpp:2138-2143).
__global__ void __launch_bounds__(256) phase_fold_kernel(size_t frames, int n, float2* __restrict__ cols) {
  const size_t total = frames * (size_t)n;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
    const int i = (int)(idx % (size_t)n);
    float2 v = cols[idx];
    if (i >= 1 && i < (n + 1) / 2) {
      v.x *= 2.0f;
      v.y *= 2.0f;
    } else if (i > n / 2) {
      v = make_float2(0.0f, 0.0f);
    }
    cols[idx] = v;
  }
}

// exp of the complex log spectrum into the output planes (fft_cffi.cpp:2150-2156).
__global__ void __launch_bounds__(256) phase_exp_kernel(const float2* __restrict__ cols, size_t frames, int bins,
                                                        int n, float* __restrict__ o_re, float* __restrict__ o_im) {
  const size_t total = frames * (size_t)bins;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
    const size_t f = idx / (size_t)bins;
    const int k = (int)(idx - f * (size_t)bins);
    const float2 L = cols[f * (size_t)n + k];
    const float a = expf(L.x);
    float s, c;
    sincosf(L.y, &s, &c);
    o_re[idx] = a * c;
    o_im[idx] = a * s;
  }
}

// Minimum phase, every step of a frame in one wave (N = 512 and 1024: one N-point
// transform per wave, T = 64 lanes of E = N/64 values). The log spectrum x is real, so
// its inverse DFT is conj(DFT(x)) / N exactly (conjugation commutes with every rounded
// operation): both transforms are forward ones over one twiddle set. The first
// transform leaves lane t holding indices t + 64j, j < E, in its last-pass order; the
// second reads the same indices in its first-pass order, so the fold between them is a
// register permutation. Samples per step as in phase_log_kernel, phase_fold_kernel and
// phase_exp_kernel (fft_cffi.cpp:2095-2158), the transforms those of the C2C kernels.
template <int L, int E, int... Rs>
__global__ void __launch_bounds__(256) phase_minimum_kernel(const float* __restrict__ mag, size_t frames, int bins,
                                                            int half, const float2* __restrict__ tw,
                                                            float* __restrict__ o_re, float* __restrict__ o_im) {
  constexpr int T = L / E;
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int NP = sizeof...(Rs);
  static_assert(T == 64 && product<Rs...>() == L, "one transform per wave");
  __shared__ float2 lds[4][padded_len<L>()];
  const int wave = threadIdx.x >> 6;
  const int t = threadIdx.x & 63;
  const size_t f = (size_t)blockIdx.x * 4 + (size_t)wave;
  if (f >= frames) return;  // a whole wave: the transforms below use wave-local exchanges only
  float2* buf = &lds[wave][0];
  const float* __restrict__ m = mag + f * (size_t)bins;
  float2 twr[E * NP];
  StockhamPlan<L, E, -1, Rs...>::load_twiddles(twr, t, tw, 1);
  float2 v[E];
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int i = t + q * T + r * (L / R0);
      const int k = half && i > L / 2 ? L - i : i;  // spec[N-k] = spec[k]
      v[q * R0 + r] = make_float2(logf(fmaxf(1e-20f, m[k])), 0.0f);  // every index is set for even N
    }
  WaveStockham<L, E, -1, 1, 0, Rs...>::run(v, t, buf, twr, [](int) {});
  // cepstrum c = conj(y) / N at i = t + 64 j, j = q + r * (E / Rl); causal fold; into the
  // first-pass order of the next transform (index j = q0 + r0 * (E / R0))
  float2 c[E];
  constexpr float inv_n = 1.0f / (float)L;
#pragma unroll
  for (int q = 0; q < E / Rl; ++q)
#pragma unroll
    for (int r = 0; r < Rl; ++r) {
      const int j = q + r * (E / Rl);
      const int i = t + 64 * j;
      float2 z = make_float2(v[q * Rl + r].x * inv_n, -v[q * Rl + r].y * inv_n);
      if (i >= 1 && i < (L + 1) / 2) {
        z.x *= 2.0f;
        z.y *= 2.0f;
      } else if (i > L / 2) {
        z = make_float2(0.0f, 0.0f);
      }
      c[(j % (E / R0)) * R0 + j / (E / R0)] = z;
    }
  WaveStockham<L, E, -1, 1, 0, Rs...>::run(c, t, buf, twr, [](int) {});
  float* __restrict__ ore = o_re + f * (size_t)bins;
  float* __restrict__ oim = o_im + f * (size_t)bins;
#pragma unroll
  for (int q = 0; q < E / Rl; ++q)
#pragma unroll
    for (int r = 0; r < Rl; ++r) {
      const int k = t + q * T + r * (L / Rl);
      if (k < bins) {
        const float2 Lk = c[q * Rl + r];
        const float a = expf(Lk.x);
        float sn, cs;
        sincosf(Lk.y, &sn, &cs);
        ore[k] = a * cs;
        oim[k] = a * sn;
      }
    }
}

unsigned grid_for(size_t work) {
  size_t g = (work + 255) / 256;
  return (unsigned)(g > 65536 ? 65536 : (g == 0 ? 1 : g));
}

}  // namespace

hipError_t launch_phase_linear(const float* mag, size_t frames, int bins, int n, int hop, float* o_re, float* o_im,
                               hipStream_t s) {
  if (frames == 0 || bins <= 0) return hipSuccess;
  const size_t blocks = (frames + kPhaseBlock - 1) / kPhaseBlock;
  const size_t work = blocks * (size_t)bins;
  if ((work + 255) / 256 > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(phase_linear_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, mag, frames, bins, n,
                     hop, o_re, o_im);
  return hipGetLastError();
}

bool phase_minimum_fused(int n) { return n == 512 || n == 1024; }

hipError_t launch_phase_minimum(const float* mag, size_t frames, int bins, int n, int half, const float2* tw,
                                float* o_re, float* o_im, hipStream_t s) {
  if (frames == 0) return hipSuccess;
  const size_t grid = (frames + 3) / 4;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  if (n == 512)
    hipLaunchKernelGGL((phase_minimum_kernel<512, 8, 8, 8, 8>), dim3((unsigned)grid), dim3(256), 0, s, mag, frames, bins,
                       half, tw, o_re, o_im);
  else if (n == 1024)
    hipLaunchKernelGGL((phase_minimum_kernel<1024, 16, 16, 16, 4>), dim3((unsigned)grid), dim3(256), 0, s, mag, frames,
                       bins, half, tw, o_re, o_im);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_phase_log(const float* mag, size_t frames, int bins, int n, int half, float2* cols, hipStream_t s) {
  hipLaunchKernelGGL(phase_log_kernel, dim3(grid_for(frames * (size_t)n)), dim3(256), 0, s, mag, frames, bins, n, half,
                     cols);
  return hipGetLastError();
}

hipError_t launch_phase_fold(size_t frames, int n, float2* cols, hipStream_t s) {
  hipLaunchKernelGGL(phase_fold_kernel, dim3(grid_for(frames * (size_t)n)), dim3(256), 0, s, frames, n, cols);
  return hipGetLastError();
}

hipError_t launch_phase_exp(const float2* cols, size_t frames, int bins, int n, float* o_re, float* o_im,
                            hipStream_t s) {
  hipLaunchKernelGGL(phase_exp_kernel, dim3(grid_for(frames * (size_t)bins)), dim3(256), 0, s, cols, frames, bins, n,
                     o_re, o_im);
  return hipGetLastError();
}

}  // namespace fftfree

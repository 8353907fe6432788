// cpu_ref.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// A CPU restatement of the reference library's batched FFT/STFT path
// (sangderenard/fftfree), used for two things only:
//   * as the parity checker of the HIP product path (tests, smoke());
//   * as the CPU baseline that bench.py times beside the GPU.
// Nothing in the product (fftfree_amd/, include/) links, loads or calls this file.
//
// The reference itself cannot be compiled here: it needs Eigen 3.4.0 through a CMake
// FetchContent download (CMakeLists.txt:105-111) and no Eigen exists offline. This file
// restates, in plain C++ over std::complex<float>, what these reference regions compute:
//   twiddles + bit reversal           eigen_fft.hpp:2234-2253
//   pre/post transform hooks          eigen_fft.hpp:2484-2596
//   Cooley-Tukey permute + stages     eigen_fft.hpp:2777-2880, 2902-3085
//   Stockham autosort stages          eigen_fft.hpp:3439-3593
//   radix-2 DIT scatter/in-place      butterfly_kernel.hpp:77-98, 128-233
//   window builder + normalisation    fft_cffi.cpp:163-208
//   frame count + framing + output    fft_cffi.cpp:701-825
//   complex-input batched inverse     fft_cffi.cpp:1042-1157 (output stage 1119-1157, 210-226)
//   Griffin-Lim                       phase_infer.hpp:69-204, fft_cffi.cpp:2176-2248
//   phase inference (linear, cepstral) phase_infer.hpp:36-66, fft_cffi.cpp:2095-2158
//   time-frequency grid pipeline      ft_tools.hpp:144-477, ft_tools_cffi.cpp:18-170
//   metadata writers (index maps)     eigen_fft.hpp:1529-1888 as called at 2672-2682,
//                                     2981-3011, 3239-3244, 3478-3505
// The reference's WorkerPool (eigen_fft.hpp:818-1419) is restated as a plain
// std::thread fan-out over contiguous frame ranges; every frame is transformed
// independently, exactly as the reference's lanes=1 column chunks are.
//
// Float arithmetic follows the reference's order of operations; bit-level equality with
// the reference binary is not claimed (its butterflies go through Eigen packet math
// compiled with -ffast-math, which no reference test pins bit-for-bit).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

namespace {

using cf = std::complex<float>;

struct RefPlan {
  int N = 0;
  int lgN = 0;
  bool inverse = false;
  std::vector<cf> W;      // N/2 twiddles, float angle (eigen_fft.hpp:2234-2242)
  std::vector<int> bitrev;  // N entries (eigen_fft.hpp:2244-2253)
};

bool make_plan(int N, bool inverse, RefPlan& P) {
  if (N <= 0) return false;
  int lg = 0, t = N;
  while ((t & 1) == 0) { ++lg; t >>= 1; }
  if ((1 << lg) != N) return false;  // eigen_fft.hpp:2228-2232: power of two only
  P.N = N;
  P.lgN = lg;
  P.inverse = inverse;
  // eigen_fft.hpp:2234-2242 -- the angle is formed in float, then cos/sin in float.
  const float sgn = inverse ? 1.0f : -1.0f;
  const float tau = sgn * float(2 * std::acos(float(-1))) / float(N);
  P.W.assign(std::max(1, N / 2), cf(1.0f, 0.0f));
  for (int k = 0; k < N / 2; ++k) {
    const float ang = tau * float(k);
    P.W[k] = cf(float(std::cos(ang)), float(std::sin(ang)));
  }
  P.bitrev.assign(N, 0);
  for (int i = 0; i < N; ++i) {  // eigen_fft.hpp:2244-2253
    unsigned x = unsigned(i), r = 0;
    for (int b = 0; b < lg; ++b) { r = (r << 1) | (x & 1u); x >>= 1; }
    P.bitrev[i] = int(r);
  }
  return true;
}

struct Flags {
  int transform = 0;  // 0 C2C, 1 R2C, 2 C2R, 3 R2R
  bool reduce_magnitude = false;
  bool store_polar = false;
  bool half_spectrum = false;
};

// eigen_fft.hpp:2484-2533 (Cooley-Tukey only; the Stockham path never calls it).
void pre_transform(const RefPlan& P, const Flags& f, cf* col) {
  const int N = P.N;
  if (f.transform == 1) {
    for (int i = 0; i < N; ++i) if (col[i].imag() != 0.0f) col[i].imag(0.0f);
    return;
  }
  if (f.transform == 2 && f.half_spectrum) {
    if (f.store_polar) {
      for (int k = 0; k <= N / 2; ++k) {
        const float mag = col[k].real(), ph = col[k].imag();
        col[k] = cf(mag * std::cos(ph), mag * std::sin(ph));
      }
    }
    col[0].imag(0.0f);
    col[N / 2].imag(0.0f);
    for (int k = 1; k < N / 2; ++k) col[N - k] = std::conj(col[k]);
  }
}

// eigen_fft.hpp:2535-2596.
void post_transform(const RefPlan& P, const Flags& f, cf* col) {
  const int N = P.N;
  const bool r2c_half = (f.transform == 1 && f.half_spectrum);
  if (r2c_half) {
    col[0].imag(0.0f);
    col[N / 2].imag(0.0f);
    for (int k = N / 2 + 1; k < N; ++k) col[k] = cf(0.0f, 0.0f);
  }
  const int upper = r2c_half ? (N / 2 + 1) : N;
  if (f.store_polar) {
    for (int i = 0; i < upper; ++i) {
      const float mag = std::abs(col[i]);
      const float ph = std::arg(col[i]);
      col[i] = cf(mag, ph);
    }
    return;
  }
  if (f.reduce_magnitude) {
    for (int i = 0; i < upper; ++i) col[i] = cf(std::abs(col[i]), 0.0f);
    return;
  }
  if (f.transform == 2) {
    for (int i = 0; i < N; ++i) if (col[i].imag() != 0.0f) col[i].imag(0.0f);
  }
}

// Index-map capture for one column (frame 0), matching the reference metadata kinds.
// Buffers hold complex<float> entries like the reference's MetadataRequest buffers.
struct Capture {
  cf* layout_map = nullptr;     // N
  cf* twiddle_index = nullptr;  // N * lgN
  cf* pairs = nullptr;          // N * lgN
  cf* stage_params = nullptr;   // lgN
  cf* twiddle = nullptr;        // 2^lgN - 1
  cf* omega = nullptr;          // N (Cooley-Tukey only)
};

// Cooley-Tukey: bit-reverse swap then lgN in-place radix-2 DIT stages
// (eigen_fft.hpp:2805-2865 permute, 2902-3016 stages, 3076-3085 scaling).
void ct_column(const RefPlan& P, const Flags& f, cf* col, const Capture* cap) {
  const int N = P.N;
  pre_transform(P, f, col);
  for (int i = 0; i < N; ++i) {
    const int src = P.bitrev[i];
    if (src <= i) continue;
    std::swap(col[i], col[src]);
  }
  for (int len = 2, s = 0; len <= N; len <<= 1, ++s) {
    const int half = len >> 1;
    const int step = N / len;
    const int blocks = N / len;
    if (cap && cap->stage_params) cap->stage_params[s] = cf(float(len), float(step));
    for (int k = 0; k < half; ++k) {
      const cf w = P.W[k * step];
      for (int block = 0; block < blocks; ++block) {
        const int a = block * len + k;
        const int b = a + half;
        const cf ai = col[a];
        const cf bt = col[b] * w;  // butterfly_kernel.hpp:93-97 (DIT)
        col[a] = ai + bt;
        col[b] = ai - bt;
        if (cap && cap->twiddle_index) {
          cap->twiddle_index[size_t(s) * N + a] = cf(float(k * step), 0.0f);
          cap->twiddle_index[size_t(s) * N + b] = cf(float(k * step), 0.0f);
        }
        if (cap && cap->pairs) {
          cap->pairs[size_t(s) * N + a] = cf(float(a), float(b));
          cap->pairs[size_t(s) * N + b] = cf(float(a), float(b));
        }
      }
    }
  }
  if (P.inverse) {
    const float scale = 1.0f / float(N);
    for (int i = 0; i < N; ++i) col[i] *= scale;
  }
  post_transform(P, f, col);
}

// Stockham autosort: ping-pong radix-2 DIT scatter stages with natural-order output
// (eigen_fft.hpp:3439-3593); no pre-transform hook, post hook after the batch (3617).
void stockham_column(const RefPlan& P, const Flags& f, cf* col, std::vector<cf>& ping,
                     std::vector<cf>& pong, const Capture* cap) {
  const int N = P.N;
  ping.assign(col, col + N);
  pong.resize(N);
  cf* in = ping.data();
  cf* out = pong.data();
  const int halfN = N >> 1;
  for (int s = 0; s < P.lgN; ++s) {
    const int m = 1 << s;
    const int seg = N >> (s + 1);
    if (cap && cap->stage_params) cap->stage_params[s] = cf(float(m << 1), float(seg));
    for (int j = 0; j < m; ++j) {
      const int tw = j * seg;
      const cf w = P.W[tw];
      for (int g = 0; g < seg; ++g) {
        const int i0 = 2 * j * seg + g;
        const int i1 = i0 + seg;
        const int o0 = j * seg + g;
        const int o1 = o0 + halfN;
        const cf a = in[i0];
        const cf bt = in[i1] * w;  // butterfly_kernel.hpp:171-176 (DIT scatter)
        out[o0] = a + bt;
        out[o1] = a - bt;
        if (cap && cap->pairs) {
          cap->pairs[size_t(s) * N + o0] = cf(float(i0), float(i1));
          cap->pairs[size_t(s) * N + o1] = cf(float(i0), float(i1));
        }
        if (cap && cap->twiddle_index) {
          cap->twiddle_index[size_t(s) * N + o0] = cf(float(tw), 0.0f);
          cap->twiddle_index[size_t(s) * N + o1] = cf(float(tw), 0.0f);
        }
      }
    }
    std::swap(in, out);
  }
  const float scale = P.inverse ? 1.0f / float(N) : 1.0f;
  for (int i = 0; i < N; ++i) col[i] = P.inverse ? in[i] * scale : in[i];
  post_transform(P, f, col);
}

// Kernel integer semantics at the C level (fft_cffi.cpp:799-800, eigen_fft.hpp:2296):
// 2 selects Stockham, everything else runs the Cooley-Tukey default.
bool is_stockham(int kernel) { return kernel == 2; }

template <class Fn>
void fan_out(size_t total, int threads, Fn fn) {
  if (threads <= 1 || total < 2) { fn(size_t(0), total); return; }
  const size_t nt = std::min<size_t>(size_t(threads), total);
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (size_t t = 0; t < nt; ++t) {
    const size_t b = total * t / nt, e = total * (t + 1) / nt;
    pool.emplace_back([=]() { fn(b, e); });
  }
  for (auto& th : pool) th.join();
}

void transform_batch(const RefPlan& P, const Flags& f, int kernel, cf* data, size_t B,
                     int threads) {
  fan_out(B, threads, [&](size_t b, size_t e) {
    std::vector<cf> ping, pong;
    for (size_t c = b; c < e; ++c) {
      cf* col = data + c * size_t(P.N);
      if (is_stockham(kernel)) stockham_column(P, f, col, ping, pong, nullptr);
      else ct_column(P, f, col, nullptr);
    }
  });
}

}  // namespace

extern "C" {

// Plan twiddles W[k], k < N/2, as interleaved floats (eigen_fft.hpp:2234-2242).
int ref_twiddles(int N, int inverse, float* out) {
  RefPlan P;
  if (!make_plan(N, inverse != 0, P)) return 0;
  std::memcpy(out, P.W.data(), sizeof(cf) * size_t(N / 2));
  return 1;
}

// lgN-bit reversal table (eigen_fft.hpp:2244-2253).
int ref_bitrev(int N, int* out) {
  RefPlan P;
  if (!make_plan(N, false, P)) return 0;
  std::memcpy(out, P.bitrev.data(), sizeof(int) * size_t(N));
  return 1;
}

// build_window + normalize_window (fft_cffi.cpp:163-208); writes W floats. Evaluated
// in the source's order of operations (no fast-math reassociation) so the taps are a
// deterministic function of the formulas: with reassociation, e.g. the Blackman end taps
// come out as 0 or -1.4e-17 depending on the compiler.
#pragma GCC push_options
#pragma GCC optimize("no-fast-math")
void ref_build_window(int kind, int W, float p1, float p2, int norm, float* out) {
  (void)p2;
  const double PI = 3.141592653589793238462643383279502884;
  std::vector<float> w(size_t(std::max(W, 0)), 1.0f);
  switch (kind) {
    case 1:  // HANN
    case 5:  // KAISER placeholder = Hann (fft_cffi.cpp:200-203)
      for (int n = 0; n < W; ++n) w[n] = float(0.5 * (1.0 - std::cos(2.0 * PI * n / (W - 1))));
      break;
    case 2:
      for (int n = 0; n < W; ++n) w[n] = float(0.54 - 0.46 * std::cos(2.0 * PI * n / (W - 1)));
      break;
    case 3:
      for (int n = 0; n < W; ++n)
        w[n] = float(0.42 - 0.5 * std::cos(2.0 * PI * n / (W - 1)) +
                     0.08 * std::cos(4.0 * PI * n / (W - 1)));
      break;
    case 4: {
      const double alpha = (p1 == 0.0f ? 0.5 : p1);
      const int L = std::max(1, W - 1);
      for (int n = 0; n < W; ++n) {
        const double x = double(n) / double(L);
        if (x < alpha / 2) w[n] = float(0.5 * (1 + std::cos(PI * (2 * x / alpha - 1))));
        else if (x <= 1 - alpha / 2) w[n] = 1.0f;
        else w[n] = float(0.5 * (1 + std::cos(PI * (2 * x / alpha - 2 / alpha + 1))));
      }
      break;
    }
    default:
      break;
  }
  if (!w.empty() && norm == 1) {
    double s2 = 0.0;
    for (float v : w) s2 += double(v) * double(v);
    s2 = std::sqrt(std::max(1e-30, s2));
    for (auto& v : w) v = float(v / s2);
  } else if (!w.empty() && norm == 2) {
    double s = 0.0;
    for (float v : w) s += double(v);
    if (std::abs(s) > 1e-30) for (auto& v : w) v = float(v / s);
  }
  if (W > 0) std::memcpy(out, w.data(), sizeof(float) * size_t(W));
}
#pragma GCC pop_options

// Frame count of fft_execute_batched (fft_cffi.cpp:715-750); 0 means "refused".
size_t ref_frame_count(size_t pcm_len, int N, int window, int hop, int ctx_pad, int pad_mode,
                       size_t max_frames) {
  const int W = window ? window : N;
  const int H = hop ? hop : W;
  if (W <= 0 || H <= 0 || W > N || pcm_len == 0) return 0;
  const int pad = (pad_mode == 0) ? ctx_pad : pad_mode;
  size_t frames = 0;
  if (pcm_len < size_t(W)) {
    if (pad == 1) frames = 1; else return 0;
  } else {
    frames = 1 + (pcm_len - size_t(W)) / size_t(H);
    const size_t last_start = (frames - 1) * size_t(H);
    const size_t remaining = (pcm_len > last_start) ? (pcm_len - last_start) : 0;
    if (remaining < size_t(W)) {
      if (pad == 1) {
      } else if (pad == 2) {
        if (frames > 0) frames -= 1;
      } else {
        return 0;
      }
    }
  }
  if (frames == 0) return 0;
  if (max_frames != 0 && frames > max_frames) frames = max_frames;
  return frames;
}

// In-place batched transform over a column-major N x B complex buffer, i.e. the
// reference's fft_inplace_batched (eigen_fft.hpp:3152-3175) with the plan's
// transform/reduce/polar/half flags and the C-level kernel code.
int ref_fft_batched(int N, int inverse, int kernel, int transform, int reduce_magnitude,
                    int store_polar, int half_spectrum, float* data, size_t B, int threads) {
  RefPlan P;
  if (!make_plan(N, inverse != 0, P)) return 0;
  Flags f{transform, reduce_magnitude != 0, store_polar != 0, half_spectrum != 0};
  transform_batch(P, f, kernel, reinterpret_cast<cf*>(data), B, threads);
  return 1;
}

// fft_execute_batched (fft_cffi.cpp:701-825, shared-plan fast path): frames the PCM,
// applies the analysis window (first W taps, only when windows are enabled), zero-pads to
// N, transforms, then writes three frame-major planes of `bins` floats each.
// `window_taps` holds N floats (taps >= W are 1.0) or is NULL when windows are off.
size_t ref_execute_batched(int N, int inverse, int kernel, int window, int hop, int ctx_pad,
                           int transform, int reduce_magnitude, int store_polar,
                           int half_spectrum, const float* window_taps, const float* pcm,
                           size_t pcm_len, float* out_real, float* out_imag, float* out_mag,
                           int pad_mode, size_t max_frames, int threads) {
  if (!pcm || pcm_len == 0 || !out_real || !out_imag || !out_mag) return 0;
  RefPlan P;
  if (!make_plan(N, inverse != 0, P)) return 0;
  const int W = window ? window : N;
  const int H = hop ? hop : W;
  const size_t frames =
      ref_frame_count(pcm_len, N, window, hop, ctx_pad, pad_mode, max_frames);
  if (frames == 0) return 0;
  Flags f{transform, reduce_magnitude != 0, store_polar != 0, half_spectrum != 0};
  const size_t bins = half_spectrum ? size_t(N / 2 + 1) : size_t(N);
  fan_out(frames, threads, [&](size_t b, size_t e) {
    std::vector<cf> col{}, ping, pong;
    col.resize(size_t(N));
    for (size_t fr = b; fr < e; ++fr) {
      const size_t start = fr * size_t(H);
      size_t copy = 0;
      if (start < pcm_len) copy = std::min(size_t(W), pcm_len - start);
      for (size_t i = 0; i < copy; ++i)
        col[i] = cf(window_taps ? pcm[start + i] * window_taps[i] : pcm[start + i], 0.0f);
      for (size_t i = copy; i < size_t(N); ++i) col[i] = cf(0.0f, 0.0f);
      if (is_stockham(kernel)) stockham_column(P, f, col.data(), ping, pong, nullptr);
      else ct_column(P, f, col.data(), nullptr);
      const size_t ob = fr * bins;
      for (size_t i = 0; i < bins; ++i) {
        out_real[ob + i] = col[i].real();
        out_imag[ob + i] = col[i].imag();
        out_mag[ob + i] = store_polar ? col[i].real() : std::abs(col[i]);
      }
    }
  });
  return frames;
}

// fft_execute_complex_batched without OLA (fft_cffi.cpp:1042-1157): frame-major complex
// bins in, Hermitian rebuild for half-spectrum C2R, transform, real part times the
// synthesis window (NULL = off) out, N floats per frame.
size_t ref_execute_complex_batched(int N, int inverse, int kernel, int transform,
                                   int reduce_magnitude, int store_polar, int half_spectrum,
                                   const float* synth_taps, const float* in_real,
                                   const float* in_imag, size_t frames, float* out_pcm,
                                   size_t max_frames, int threads) {
  if (!in_real || !in_imag || !out_pcm || frames == 0) return 0;
  RefPlan P;
  if (!make_plan(N, inverse != 0, P)) return 0;
  if (max_frames != 0 && frames > max_frames) frames = max_frames;
  Flags f{transform, reduce_magnitude != 0, store_polar != 0, half_spectrum != 0};
  const size_t bins = half_spectrum ? size_t(N / 2 + 1) : size_t(N);
  fan_out(frames, threads, [&](size_t b, size_t e) {
    std::vector<cf> col{}, ping, pong;
    col.resize(size_t(N));
    for (size_t fr = b; fr < e; ++fr) {
      for (size_t i = 0; i < bins; ++i) col[i] = cf(in_real[fr * bins + i], in_imag[fr * bins + i]);
      if (half_spectrum && transform == 2) {
        col[0].imag(0.0f);
        const size_t ny = size_t(N) / 2;
        if (ny < bins) col[ny].imag(0.0f);
        const size_t maxk = (bins > ny + 1) ? ny : (bins - 1);
        for (size_t k = 1; k < maxk; ++k) col[size_t(N) - k] = std::conj(col[k]);
      } else {
        for (size_t i = bins; i < size_t(N); ++i) col[i] = cf(0.0f, 0.0f);
      }
      if (is_stockham(kernel)) stockham_column(P, f, col.data(), ping, pong, nullptr);
      else ct_column(P, f, col.data(), nullptr);
      for (size_t i = 0; i < size_t(N); ++i) {
        float s = col[i].real();
        if (synth_taps) s *= synth_taps[i];
        out_pcm[fr * size_t(N) + i] = s;
      }
    }
  });
  return frames;
}

// Output stage of fft_execute_complex_batched (fft_cffi.cpp:1119-1157) from the real
// parts of the transformed columns (frames x N, frame-major). With windows enabled, s =
// re * synthesis[i]; apply_ola = 0: out[f*N + i] = s. apply_ola = 1: out is zeroed over
// frames*N samples and out[f*H + i] += s for f*H + i < frames*N, frames ascending; with
// cola_mode 1 (FFT_COLA_NORMALIZE) norm[t] accumulates cola_window_weight (analysis[i] *
// synthesis[i] with windows enabled, else 1; fft_cffi.cpp:210-226) and out[t] /= norm[t]
// where norm[t] > 1e-12. NULL taps stand for the reference's empty window vectors.
void ref_complex_output(int N, size_t frames, const float* cols_real, int windows_enabled, const float* analysis,
                        const float* synthesis, int apply_ola, int hop, int cola_mode, float* out_pcm) {
  const size_t n = size_t(N);
  const bool win = windows_enabled && synthesis;
  auto weight = [&](size_t i) -> float {
    if (!windows_enabled) return 1.0f;
    const float a = analysis ? analysis[i] : 1.0f;
    const float b = synthesis ? synthesis[i] : 1.0f;
    return a * b;
  };
  if (!apply_ola) {
    for (size_t f = 0; f < frames; ++f)
      for (size_t i = 0; i < n; ++i) {
        float v = cols_real[f * n + i];
        if (win) v *= synthesis[i];
        out_pcm[f * n + i] = v;
      }
    return;
  }
  const size_t total = frames * n;
  for (size_t t = 0; t < total; ++t) out_pcm[t] = 0.0f;
  std::vector<float> norm;
  if (cola_mode == 1) norm.assign(total, 0.0f);
  for (size_t f = 0; f < frames; ++f) {
    const size_t start = f * size_t(hop);
    for (size_t i = 0; i < n; ++i) {
      float v = cols_real[f * n + i];
      if (win) v *= synthesis[i];
      const size_t t = start + i;
      if (t < total) {
        out_pcm[t] += v;
        if (!norm.empty()) norm[t] += weight(i);
      }
    }
  }
  if (!norm.empty())
    for (size_t t = 0; t < total; ++t)
      if (norm[t] > 1e-12f) out_pcm[t] /= norm[t];
}

// The start angles of Griffin-Lim (phase_infer.hpp:124-129): std::mt19937 seeded with
// `seed` (0x47d7bfad when 0) through std::uniform_real_distribution<float>(-pi, pi), in
// the order the bins are visited. Built with the reference's flags.
void ref_gl_angles(unsigned seed, size_t count, float* out) {
  constexpr float kPi = 3.14159265358979323846f;
  std::mt19937 rng(seed ? seed : 0x47d7bfadu);
  std::uniform_real_distribution<float> dist(-kPi, kPi);
  for (size_t i = 0; i < count; ++i) out[i] = dist(rng);
}

// Griffin-Lim (phase_infer.hpp:69-204) for a forward R2C context (kernel, window, hop,
// context pad mode, analysis taps or NULL) and an inverse context (kernel, transform,
// synthesis taps or NULL; no polar/magnitude hooks, no internal overlap-add), both of
// size N with the same half_spectrum flag; hop is the overlap-add hop the C entry
// resolves (fft_cffi.cpp:2214-2222). Returns 1, or 0 where the reference returns false.
int ref_griffin_lim(int N, int half, const float* mag, size_t frames, int hop, int iterations, int pad_mode,
                    unsigned seed, int inv_kernel, int inv_transform, const float* inv_synth, int fwd_kernel,
                    int fwd_window, int fwd_hop, int fwd_ctx_pad, const float* fwd_taps, float* out_real,
                    float* out_imag, int threads) {
  if (!mag || !out_real || !out_imag || frames == 0 || N <= 0 || hop <= 0) return 0;
  const size_t n = size_t(N);
  const size_t bins = half ? n / 2 + 1 : n;
  const size_t total = bins * frames;
  const size_t ola_len = (frames - 1) * size_t(hop) + n;
  const int iters = iterations > 0 ? iterations : 32;
  std::vector<float> frame_pcm(frames * n), ola(ola_len), weight(ola_len), tr(total), ti(total), tm(total);
  std::vector<float> ang(total);
  ref_gl_angles(seed, total, ang.data());
  for (size_t i = 0; i < total; ++i) {
    out_real[i] = mag[i] * std::cos(ang[i]);
    out_imag[i] = mag[i] * std::sin(ang[i]);
  }
  for (int it = 0; it < iters; ++it) {
    const size_t produced = ref_execute_complex_batched(N, 1, inv_kernel, inv_transform, 0, 0, half, inv_synth,
                                                        out_real, out_imag, frames, frame_pcm.data(), frames, threads);
    if (produced == 0) return 0;
    std::fill(ola.begin(), ola.end(), 0.0f);
    std::fill(weight.begin(), weight.end(), 0.0f);
    const size_t usable = std::min(produced, frames);
    for (size_t f = 0; f < usable; ++f) {
      const size_t start = f * size_t(hop);
      if (start + n > ola_len) break;
      for (size_t i = 0; i < n; ++i) {
        ola[start + i] += frame_pcm[f * n + i];
        weight[start + i] += 1.0f;
      }
    }
    for (size_t t = 0; t < ola_len; ++t)
      if (weight[t] > 0.0f) ola[t] /= weight[t];
    const size_t fwd = ref_execute_batched(N, 0, fwd_kernel, fwd_window, fwd_hop, fwd_ctx_pad, 1, 0, 0, half,
                                           fwd_taps, ola.data(), ola_len, tr.data(), ti.data(), tm.data(),
                                           pad_mode, frames, threads);
    if (fwd == 0) return 0;
    const size_t uf = std::min(fwd, frames);
    for (size_t i = 0; i < uf * bins; ++i) {
      const float ph = std::atan2(ti[i], tr[i]);
      out_real[i] = mag[i] * std::cos(ph);
      out_imag[i] = mag[i] * std::sin(ph);
    }
    if (half)
      for (size_t f = 0; f < uf; ++f) {
        out_imag[f * bins] = 0.0f;
        if ((n & 1) == 0 && bins > 1) out_imag[f * bins + bins - 1] = 0.0f;
      }
  }
  return 1;
}

// infer_linear (phase_infer.hpp:36-66): per-bin phase advance accumulated in double.
void ref_phase_linear(int N, int hop, int half, const float* mag, size_t frames, float* out_real, float* out_imag) {
  const int H = hop > 0 ? hop : N;
  const int bins = half ? N / 2 + 1 : N;
  const double two_pi = 6.28318530717958647692;
  std::vector<double> dphi(static_cast<size_t>(bins), 0.0), phi(static_cast<size_t>(bins), 0.0);
  for (int k = 0; k < bins; ++k) {
    const int kk = k < N ? k : k % N;
    dphi[size_t(k)] = two_pi * double(kk) * double(H) / double(N);
  }
  for (size_t f = 0; f < frames; ++f)
    for (int k = 0; k < bins; ++k) {
      const size_t i = f * size_t(bins) + size_t(k);
      const double m = double(mag[i]);
      out_real[i] = float(m * std::cos(phi[size_t(k)]));
      out_imag[i] = float(m * std::sin(phi[size_t(k)]));
      phi[size_t(k)] += dphi[size_t(k)];
    }
}

// Minimum-phase spectra through the real cepstrum (fft_cffi.cpp:2095-2158), on the
// restated float transforms (Cooley-Tukey, the plan cache's default kernel).
int ref_phase_minimum(int N, int half, const float* mag, size_t frames, float* out_real, float* out_imag) {
  RefPlan Pi, Pf;
  if (!make_plan(N, true, Pi) || !make_plan(N, false, Pf)) return 0;
  const int bins = half ? N / 2 + 1 : N;
  const Flags c2c{0, false, false, false};
  std::vector<cf> spec(size_t(N), cf(0, 0)), tmp;
  const float eps = 1e-20f;
  for (size_t f = 0; f < frames; ++f) {
    const size_t base = f * size_t(bins);
    auto set_log = [&](int k, float m) { spec[size_t(k)] = cf(std::log(std::max(eps, m)), 0.0f); };
    set_log(0, mag[base]);
    if (half) {
      for (int k = 1; k < N / 2; ++k) {
        set_log(k, mag[base + size_t(k)]);
        spec[size_t(N - k)] = spec[size_t(k)];
      }
      if (N % 2 == 0) set_log(N / 2, mag[base + size_t(N / 2)]);
    } else {
      for (int k = 1; k < N; ++k) set_log(k, mag[base + size_t(k)]);
    }
    tmp = spec;
    ct_column(Pi, c2c, tmp.data(), nullptr);
    for (int n = 1; n < (N + 1) / 2; ++n) tmp[size_t(n)] *= 2.0f;
    for (int n = N / 2 + 1; n < N; ++n) tmp[size_t(n)] = cf(0.0f, 0.0f);
    ct_column(Pf, c2c, tmp.data(), nullptr);
    for (int k = 0; k < bins; ++k) {
      const cf L = tmp[size_t(k)];
      const float a = std::exp(L.real());
      out_real[base + size_t(k)] = a * std::cos(L.imag());
      out_imag[base + size_t(k)] = a * std::sin(L.imag());
    }
  }
  return 1;
}

// The grid pipeline of ft_apply_complex_pipeline / ft_apply_polar_pipeline: bilinear
// upsample (ft_tools.hpp:144-183), harmonic quantizer (262-356), local contrast booster
// (358-461), box downsample (185-227), polar grids resampled through complex values
// (229-245). Settings arrive unpacked; returns 1, or 0 when the result's shape differs.
namespace {
struct FtField {
  int tb = 0, fb = 0;
  std::vector<cf> z;  // complex grid
  std::vector<float> mag, ph;  // polar grid
};

std::vector<cf> ft_up(const std::vector<cf>& src, int ts, int fs, int tf, int ff) {
  const int td = ts * tf, fd = fs * ff;
  std::vector<cf> dst(size_t(td) * size_t(fd));
  auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
  for (int t = 0; t < td; ++t) {
    const double st = double(t) / double(tf);
    const int t0 = int(std::floor(st)), t1 = std::min(t0 + 1, ts - 1);
    const float tw = float(st - double(t0));
    for (int f = 0; f < fd; ++f) {
      const double sf = double(f) / double(ff);
      const int f0 = int(std::floor(sf)), f1 = std::min(f0 + 1, fs - 1);
      const float fw = float(sf - double(f0));
      const cf c00 = src[size_t(t0) * fs + f0], c10 = src[size_t(t1) * fs + f0];
      const cf c01 = src[size_t(t0) * fs + f1], c11 = src[size_t(t1) * fs + f1];
      const cf a(lerp(c00.real(), c10.real(), tw), lerp(c00.imag(), c10.imag(), tw));
      const cf b(lerp(c01.real(), c11.real(), tw), lerp(c01.imag(), c11.imag(), tw));
      dst[size_t(t) * fd + f] = cf(lerp(a.real(), b.real(), fw), lerp(a.imag(), b.imag(), fw));
    }
  }
  return dst;
}

std::vector<cf> ft_down(const std::vector<cf>& src, int ts, int fs, int tf, int ff, bool preserve) {
  const int td = (ts + tf - 1) / tf, fd = (fs + ff - 1) / ff;
  std::vector<cf> dst(size_t(td) * size_t(fd));
  for (int t = 0; t < td; ++t)
    for (int f = 0; f < fd; ++f) {
      cf acc(0.0f, 0.0f);
      int count = 0;
      for (int a = t * tf; a < std::min(t * tf + tf, ts); ++a)
        for (int b = f * ff; b < std::min(f * ff + ff, fs); ++b) {
          acc += src[size_t(a) * fs + b];
          ++count;
        }
      if (preserve && count > 0) acc *= 1.0f / float(count);
      dst[size_t(t) * fd + f] = acc;
    }
  return dst;
}

float ft_quant_mag(float m, float freq, float base, float step, float mix, float decay) {
  float w;
  if (mix <= 0.0f) {
    w = 0.0f;
  } else if (base <= 0.0f) {
    w = std::min(std::max(mix, 0.0f), 1.0f);
  } else {
    const float ratio = freq / base;
    if (!std::isfinite(double(ratio))) {
      w = 0.0f;
    } else {
      const float nearest = float(std::round(double(ratio)));
      const float attraction = std::exp(-decay * std::fabs(ratio - nearest));
      w = std::min(std::max(mix * attraction, 0.0f), 1.0f);
    }
  }
  if (w <= 0.0f) return m;
  float target = m;
  if (step > 0.0f) target = float(std::round(double(m / step))) * step;
  return m + (target - m) * w;
}

void ft_contrast_mags(std::vector<float>& mags, int tb, int fb, int kt, int kf, float bright, float contrast,
                      float gamma, bool clamp_zero) {
  std::vector<float> out(mags.size());
  const bool gt = kt <= 0, gf = kf <= 0;
  const int rt = gt ? tb : std::max(kt / 2, 0), rf = gf ? fb : std::max(kf / 2, 0);
  float gmean = 0.0f;
  if (gt && gf) {
    for (float v : mags) gmean += v;
    gmean /= float(mags.size());
  }
  const float g = gamma > 0.0f ? gamma : 1.0f;
  const bool use_gamma = std::fabs(g - 1.0f) > 1e-5f;
  for (int t = 0; t < tb; ++t)
    for (int f = 0; f < fb; ++f) {
      float mean = 0.0f;
      if (gt && gf) {
        mean = gmean;
      } else {
        int count = 0;
        for (int a = gt ? 0 : std::max(0, t - rt); a <= (gt ? tb - 1 : std::min(tb - 1, t + rt)); ++a)
          for (int b = gf ? 0 : std::max(0, f - rf); b <= (gf ? fb - 1 : std::min(fb - 1, f + rf)); ++b) {
            mean += mags[size_t(a) * fb + b];
            ++count;
          }
        if (count > 0) mean /= float(count);
      }
      const size_t i = size_t(t) * fb + f;
      float adj = mean + (mags[i] - mean) * (1.0f + contrast) + bright;
      if (clamp_zero) adj = std::max(adj, 0.0f);
      if (use_gamma) adj = float(std::pow(double(std::max(adj, 1e-12f)), double(1.0f / g)));
      out[i] = adj;
    }
  mags.swap(out);
}
}  // namespace

int ref_ft_pipeline(int polar, double f_origin, double f_step, int ut, int uf, int dt, int df, int preserve,
                    int q_on, float q_base, float q_step, float q_mix, float q_decay, int c_on, int c_kt, int c_kf,
                    float c_bright, float c_contrast, float c_gamma, int c_clamp, float* a, float* b, int tb,
                    int fb) {
  if (!a || !b || tb <= 0 || fb <= 0) return 0;
  ut = ut > 0 ? ut : 1; uf = uf > 0 ? uf : 1; dt = dt > 0 ? dt : 1; df = df > 0 ? df : 1;
  if (f_step == 0.0) f_step = 1.0;
  int t = tb, f = fb;
  const size_t n0 = size_t(tb) * fb;
  std::vector<cf> z(n0);
  std::vector<float> mag(a, a + n0), ph(b, b + n0);
  auto to_z = [&]() {
    z.resize(mag.size());
    for (size_t i = 0; i < mag.size(); ++i) z[i] = std::polar(mag[i], ph[i]);
  };
  auto from_z = [&]() {
    mag.resize(z.size());
    ph.resize(z.size());
    for (size_t i = 0; i < z.size(); ++i) {
      mag[i] = std::abs(z[i]);
      ph[i] = std::arg(z[i]);
    }
  };
  if (!polar)
    for (size_t i = 0; i < n0; ++i) z[i] = cf(a[i], b[i]);
  if (ut > 1 || uf > 1) {
    if (polar) to_z();
    z = ft_up(z, t, f, ut, uf);
    t *= ut;
    f *= uf;
    f_step /= double(uf);
    if (polar) from_z();
  }
  if (q_on && q_mix > 0.0f) {
    for (int i = 0; i < t; ++i)
      for (int k = 0; k < f; ++k) {
        const float freq = float(f_origin + f_step * double(k));
        const size_t at = size_t(i) * f + k;
        if (polar) {
          mag[at] = ft_quant_mag(mag[at], freq, q_base, q_step, q_mix, q_decay);
        } else {
          const float m = std::abs(z[at]);
          if (m <= std::numeric_limits<float>::epsilon()) continue;
          const float adj = std::max(ft_quant_mag(m, freq, q_base, q_step, q_mix, q_decay), 0.0f);
          z[at] = std::polar(adj, std::arg(z[at]));
        }
      }
  }
  if (c_on) {
    if (polar) {
      ft_contrast_mags(mag, t, f, c_kt, c_kf, c_bright, c_contrast, c_gamma, c_clamp != 0);
    } else {
      std::vector<float> m(z.size());
      for (size_t i = 0; i < z.size(); ++i) m[i] = std::abs(z[i]);
      ft_contrast_mags(m, t, f, c_kt, c_kf, c_bright, c_contrast, c_gamma, c_clamp != 0);
      for (size_t i = 0; i < z.size(); ++i) z[i] = std::polar(m[i], std::arg(z[i]));
    }
  }
  if (dt > 1 || df > 1) {
    if (polar) to_z();
    z = ft_down(z, t, f, dt, df, preserve != 0);
    t = (t + dt - 1) / dt;
    f = (f + df - 1) / df;
    if (polar) from_z();
  }
  if (t != tb || f != fb) return 0;
  for (size_t i = 0; i < n0; ++i) {
    a[i] = polar ? mag[i] : z[i].real();
    b[i] = polar ? ph[i] : z[i].imag();
  }
  return 1;
}

// Metadata of one column as the reference writers record it. kind follows
// eigfft::MetadataKind order (eigen_fft.hpp:1477-1491): 0 Twiddle, 1 LayoutMap,
// 3 StageParams, 4 ButterflyPairs, 6 TwiddleIndexMap, 7 Omega. Returns the entry count
// written (complex<float> entries), 0 if the kind is not recorded by that kernel.
size_t ref_metadata(int N, int inverse, int kernel, int transform, int half_spectrum, int kind,
                    float* out, size_t capacity) {
  RefPlan P;
  if (!make_plan(N, inverse != 0, P)) return 0;
  const int S = P.lgN;
  std::vector<cf> buf;
  Capture cap;
  size_t count = 0;
  switch (kind) {
    case 0: {  // Twiddle: record_plan_twiddles (eigen_fft.hpp:1566-1577)
      count = (size_t(1) << S) - 1;
      buf.assign(count, cf(0, 0));
      for (int s = 0; s < S; ++s) {
        const int m = 1 << s, step = N / (m << 1);
        for (int j = 0; j < m; ++j) buf[size_t(m - 1 + j)] = P.W[j * step];
      }
      break;
    }
    case 1:  // LayoutMap: identity for both kernels (eigen_fft.hpp:2672-2682, 3239-3244)
      count = size_t(N);
      buf.resize(count);
      for (int p = 0; p < N; ++p) buf[p] = cf(float(p), 0.0f);
      break;
    case 7: {  // Omega: Cooley-Tukey only (eigen_fft.hpp:1752-1762, 2665-2667)
      if (is_stockham(kernel)) return 0;
      const size_t upper = (transform == 1 && half_spectrum) ? size_t(N / 2 + 1) : size_t(N);
      count = upper;
      buf.assign(count, cf(0, 0));
      const float two_pi = float(2 * std::acos(float(-1)));
      for (size_t k = 0; k < upper; ++k) buf[k] = cf(two_pi * float(k) / float(N), 0.0f);
      break;
    }
    case 3:
    case 4:
    case 6: {
      count = (kind == 3) ? size_t(S) : size_t(N) * size_t(S);
      buf.assign(count, cf(0, 0));
      if (kind == 3) cap.stage_params = buf.data();
      if (kind == 4) cap.pairs = buf.data();
      if (kind == 6) cap.twiddle_index = buf.data();
      std::vector<cf> col(size_t(N), cf(0, 0)), ping, pong;
      Flags f{};
      if (is_stockham(kernel)) stockham_column(P, f, col.data(), ping, pong, &cap);
      else ct_column(P, f, col.data(), &cap);
      break;
    }
    default:
      return 0;
  }
  if (count > capacity) return 0;
  std::memcpy(out, buf.data(), sizeof(cf) * count);
  return count;
}

}  // extern "C"

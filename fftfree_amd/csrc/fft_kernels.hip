// fft_kernels.hip -- gfx950 kernels of the batched FFT/STFT path and their launchers.
//
// stft_real_kernel  fused framing + analysis window + zero-pad (fft_cffi.cpp:766-791),
//                   real-input DFT of size N as a complex DFT of size N/2 plus a
//                   split/twiddle step, post-transform hook (eigen_fft.hpp:2535-2596) and
//                   the three-plane output copy (fft_cffi.cpp:811-825), one launch.
// c2c_kernel        batched complex DFT of size N with the pre/post hooks
//                   (eigen_fft.hpp:2484-2596), the C++ fft_inplace_batched path
//                   (eigen_fft.hpp:3152-3175).
// *_tiny_kernel     direct DFT for N too small for a multi-thread Stockham pass.
// metadata_kernel   the reference's integer metadata maps from closed forms
//                   (eigen_fft.hpp:1529-1888 writers as called by both kernels).
//
// Data layout in HBM: PCM is one float array; spectra planes are frame-major
// (frame f, bin k at f*bins + k); complex batches are transform-major float2.
// One workgroup owns FPW whole frames: nothing crosses workgroups, so no atomics.
#include "fft_device.hpp"
#include "fft_kernels.hpp"

namespace fftfree {
using namespace dev;

namespace {

// Reads framed, windowed real samples x[m] = pcm[start+m] * win[m] for m < cnt, else 0
// (fft_cffi.cpp:766-791; the product is a single float multiply as in the reference).
__device__ __forceinline__ float frame_sample(const float* __restrict__ src, const float* __restrict__ win,
                                             size_t cnt, int m) {
  if ((size_t)m >= cnt) return 0.0f;
  const float x = src[m];
  return win ? x * win[m] : x;
}

template <int L, int E, int FPW, int DIR, int... Rs>
__global__ void __launch_bounds__(FPW*(L / E)) stft_real_kernel(StftArgs a) {
  constexpr int T = L / E;
  constexpr int N = 2 * L;
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 lds[2][FPW][L];

  const int tr = threadIdx.x / T;
  const int t = threadIdx.x - tr * T;
  const size_t tile = (size_t)xcd_tile(blockIdx.x, gridDim.x);
  const size_t f = tile * FPW + tr;
  const bool active = f < a.frames;
  const size_t start = (f + a.frame_offset) * (size_t)a.hop;
  size_t cnt = 0;
  if (active && start < a.pcm_len) {
    const size_t rem = a.pcm_len - start;
    cnt = rem < (size_t)a.window ? rem : (size_t)a.window;
  }
  const float* __restrict__ src = a.pcm + (active ? start : 0);
  const float* __restrict__ win = a.win;

  // z[n] = x[2n] + i x[2n+1], loaded in first-pass butterfly order.
  float2 v[E];
  if (cnt == (size_t)N && ((((uintptr_t)src) & 7) == 0)) {
    const float2* __restrict__ s2 = reinterpret_cast<const float2*>(src);
    const float2* __restrict__ w2 = reinterpret_cast<const float2*>(win);
#pragma unroll
    for (int q = 0; q < E / R0; ++q) {
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
    }
  } else {
#pragma unroll
    for (int q = 0; q < E / R0; ++q) {
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int n = t + q * T + r * (L / R0);
        v[q * R0 + r] = make_float2(frame_sample(src, win, cnt, 2 * n), frame_sample(src, win, cnt, 2 * n + 1));
      }
    }
  }

  float2* free_buf;
  run_passes<L, E, DIR, 1, Rs...>(v, t, &lds[0][tr][0], &lds[1][tr][0], a.tw, 2, free_buf);
  store_natural<L, E, Rl>(v, t, free_buf);
  __syncthreads();

  // Split: E_k = (Z_k + conj Z_{L-k})/2, O_k = -i (Z_k - conj Z_{L-k})/2,
  // X_k = E_k + w_N^k O_k for k = 0..L; bins above L (full-spectrum output) are
  // conj(X_{N-k}) because the input is real.
  const float2* Z = free_buf;
  const size_t obase = f * (size_t)a.bins;
  for (int k = t; k < a.bins; k += T) {
    const int kk = k <= L ? k : N - k;
    const float2 zk = Z[kk & (L - 1)];
    const float2 zc = c_conj(Z[(L - kk) & (L - 1)]);
    const float2 e = c_scale(c_add(zk, zc), 0.5f);
    const float2 d = c_sub(zk, zc);
    const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
    const float2 w = kk < L ? load_tw<DIR>(a.tw, kk) : make_float2(-1.0f, 0.0f);
    float2 X = c_add(e, c_mul(w, o));
    X = c_scale(X, a.scale);
    if (kk == 0 || kk == L) X.y = 0.0f;  // exact for real input; R2C half rule
    if (k > L) X = c_conj(X);
    if (active) store_planes(X, a.out_mode, obase + k, a.o_re, a.o_im, a.o_mag);
  }
}

// Pre-transform hook on load (eigen_fft.hpp:2484-2533): mode 1 clears imag (R2C);
// mode 2 rebuilds the Hermitian spectrum of a C2R half-spectrum column from bins
// 0..N/2 (optionally stored as (mag, phase)).
__device__ __forceinline__ float2 load_pre(const float2* __restrict__ col, int n, int N, int mode, int polar) {
  if (mode == 0) return col[n];
  if (mode == 1) return make_float2(col[n].x, 0.0f);
  const int h = N >> 1;
  const int src = n <= h ? n : N - n;
  float2 z = col[src];
  if (polar) {
    float sn, cs;
    sincosf(z.y, &sn, &cs);
    z = make_float2(z.x * cs, z.x * sn);
  }
  if (src == 0 || src == h) z.y = 0.0f;
  return n <= h ? z : c_conj(z);
}

// Post-transform hook on store (eigen_fft.hpp:2535-2596) for output index idx.
__device__ __forceinline__ float2 post_op(float2 X, int idx, int N, const C2CArgs& a) {
  X = c_scale(X, a.scale);
  if (a.post_half) {
    if (idx == 0 || idx == (N >> 1)) X.y = 0.0f;
    if (idx > (N >> 1)) X = make_float2(0.0f, 0.0f);
  }
  if (a.post_mode >= 2 && idx < a.post_upper) {
    const float m = __builtin_sqrtf(__builtin_fmaf(X.x, X.x, X.y * X.y));
    X = make_float2(m, a.post_mode == 3 ? atan2f(X.y, X.x) : 0.0f);
  } else if (a.post_mode == 1) {
    X.y = 0.0f;
  }
  return X;
}

template <int L, int E, int FPW, int DIR, int... Rs>
__global__ void __launch_bounds__(FPW*(L / E)) c2c_kernel(C2CArgs a) {
  constexpr int T = L / E;
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 lds[2][FPW][L];

  const int tr = threadIdx.x / T;
  const int t = threadIdx.x - tr * T;
  const size_t tile = (size_t)xcd_tile(blockIdx.x, gridDim.x);
  const size_t f = tile * FPW + tr;
  const bool active = f < a.batch;
  const float2* __restrict__ col = a.in + (active ? f : 0) * (size_t)L;

  float2 v[E];
#pragma unroll
  for (int q = 0; q < E / R0; ++q) {
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int n = t + q * T + r * (L / R0);
      v[q * R0 + r] = active ? load_pre(col, n, L, a.pre_mode, a.pre_polar) : make_float2(0.0f, 0.0f);
    }
  }
  float2* free_buf;
  run_passes<L, E, DIR, 1, Rs...>(v, t, &lds[0][tr][0], &lds[1][tr][0], a.tw, 1, free_buf);
  if (active) {
    float2* __restrict__ dst = a.out + f * (size_t)L;
#pragma unroll
    for (int q = 0; q < E / Rl; ++q) {
#pragma unroll
      for (int r = 0; r < Rl; ++r) {
        const int idx = t + q * T + r * (L / Rl);
        dst[idx] = post_op(v[q * Rl + r], idx, L, a);
      }
    }
  }
}

// Direct DFT for tiny N, one thread per (frame, bin).
template <int DIR>
__global__ void stft_tiny_kernel(StftArgs a) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t f = gid / (size_t)a.bins;
  const int k = (int)(gid - f * (size_t)a.bins);
  if (f >= a.frames) return;
  const int N = a.n;
  const size_t start = (f + a.frame_offset) * (size_t)a.hop;
  size_t cnt = 0;
  if (start < a.pcm_len) {
    const size_t rem = a.pcm_len - start;
    cnt = rem < (size_t)a.window ? rem : (size_t)a.window;
  }
  const float* src = a.pcm + start;
  float2 X = make_float2(0.0f, 0.0f);
  for (int m = 0; m < N; ++m) {
    const float x = frame_sample(src, a.win, cnt, m);
    const float2 w = load_tw<DIR>(a.tw, (int)(((long)k * m) & (N - 1)));
    X.x = __builtin_fmaf(x, w.x, X.x);
    X.y = __builtin_fmaf(x, w.y, X.y);
  }
  X = c_scale(X, a.scale);
  if (k == 0 || k == N / 2) X.y = 0.0f;
  store_planes(X, a.out_mode, f * (size_t)a.bins + k, a.o_re, a.o_im, a.o_mag);
}

template <int DIR>
__global__ void c2c_tiny_kernel(C2CArgs a) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.n;
  const size_t f = gid / (size_t)N;
  const int k = (int)(gid - f * (size_t)N);
  if (f >= a.batch) return;
  const float2* col = a.in + f * (size_t)N;
  float2 X = make_float2(0.0f, 0.0f);
  for (int m = 0; m < N; ++m) {
    const float2 x = load_pre(col, m, N, a.pre_mode, a.pre_polar);
    X = c_add(X, c_mul(x, load_tw<DIR>(a.tw, (int)(((long)k * m) & (N - 1)))));
  }
  // in may alias out: N < 8 divides the wave size, so a transform's N threads share one
  // wave, and the store depends on every load of that wave having returned.
  a.out[f * (size_t)N + k] = post_op(X, k, N, a);
}

// Integer metadata from closed forms; one thread per entry.
__global__ void metadata_kernel(int N, int S, int kernel, int kind, int upper, const float2* tw, int inverse,
                                float2* out, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const bool stockham = (kernel == 2);
  float2 v = make_float2(0.0f, 0.0f);
  switch (kind) {
    case kMetaTwiddle: {  // slot (2^s - 1) + j holds W[j * N / 2^{s+1}]
      const int s = 31 - __builtin_clz(i + 1);
      const int j = i - ((1 << s) - 1);
      const float2 w = tw[j * (N >> (s + 1))];
      v = inverse ? c_conj(w) : w;
      break;
    }
    case kMetaLayoutMap:
      v = make_float2((float)i, 0.0f);
      break;
    case kMetaStageParams:
      v = make_float2((float)(2 << i), (float)(N >> (i + 1)));
      break;
    case kMetaButterflyPairs:
    case kMetaTwiddleIndexMap: {
      const int s = i / N;
      const int p = i - s * N;
      int lhs, rhs, twi;
      if (stockham) {
        const int seg = N >> (s + 1);
        const int pp = p & ((N >> 1) - 1);
        const int j = pp / seg, g = pp - j * seg;
        lhs = 2 * j * seg + g;
        rhs = lhs + seg;
        twi = j * seg;
      } else {
        const int half = 1 << s;
        lhs = (p & half) ? p - half : p;
        rhs = lhs + half;
        twi = (p & (half - 1)) * (N >> (s + 1));
      }
      v = kind == kMetaButterflyPairs ? make_float2((float)lhs, (float)rhs) : make_float2((float)twi, 0.0f);
      break;
    }
    case kMetaOmega:  // float(2 pi) * k / N in float, as the reference computes it
      v = make_float2((6.28318548202514648f * (float)i) / (float)N, 0.0f);
      break;
    default:
      break;
  }
  (void)S;
  (void)upper;
  out[i] = v;
}

// Generic real-input path, step 1: windowed, zero-padded frames as complex columns
// (imag 0), N per frame (fft_cffi.cpp:766-791).
__global__ void frame_complex_kernel(StftArgs a, float2* out) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t f = gid / (size_t)a.n;
  const int m = (int)(gid - f * (size_t)a.n);
  if (f >= a.frames) return;
  const size_t start = (f + a.frame_offset) * (size_t)a.hop;
  size_t cnt = 0;
  if (start < a.pcm_len) {
    const size_t rem = a.pcm_len - start;
    cnt = rem < (size_t)a.window ? rem : (size_t)a.window;
  }
  out[gid] = make_float2(frame_sample(a.pcm + start, a.win, cnt, m), 0.0f);
}

// Generic path, step 3: output copy of transformed columns into the three planes
// (fft_cffi.cpp:811-825): re, im, and |v| (or re when store_polar).
__global__ void planes_kernel(const float2* in, size_t frames, int n, int bins, int polar, float* o_re, float* o_im,
                              float* o_mag) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t f = gid / (size_t)bins;
  const int k = (int)(gid - f * (size_t)bins);
  if (f >= frames) return;
  const float2 v = in[f * (size_t)n + k];
  if (o_re) o_re[gid] = v.x;
  if (o_im) o_im[gid] = v.y;
  if (o_mag) o_mag[gid] = polar ? v.x : __builtin_sqrtf(__builtin_fmaf(v.x, v.x, v.y * v.y));
}

template <int L, int E, int FPW, int DIR, int... Rs>
hipError_t stft_launch_impl(const StftArgs& a, hipStream_t s) {
  const size_t grid = (a.frames + FPW - 1) / FPW;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((stft_real_kernel<L, E, FPW, DIR, Rs...>), dim3((unsigned)grid), dim3(FPW * (L / E)), 0, s,
                     a);
  return hipGetLastError();
}

template <int L, int E, int FPW, int DIR, int... Rs>
hipError_t c2c_launch_impl(const C2CArgs& a, hipStream_t s) {
  const size_t grid = (a.batch + FPW - 1) / FPW;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((c2c_kernel<L, E, FPW, DIR, Rs...>), dim3((unsigned)grid), dim3(FPW * (L / E)), 0, s, a);
  return hipGetLastError();
}

using StftFn = hipError_t (*)(const StftArgs&, hipStream_t);
using C2CFn = hipError_t (*)(const C2CArgs&, hipStream_t);

struct StftEntry {
  int n;
  VariantInfo info;
  StftFn fwd, inv;
};
struct C2CEntry {
  int n;
  VariantInfo info;
  C2CFn fwd, inv;
};

#define FFTFREE_STFT(L, E, FPW, RS, ...)                                                  \
  {2 * (L), {L, E, FPW, RS}, &stft_launch_impl<L, E, FPW, -1, __VA_ARGS__>,               \
   &stft_launch_impl<L, E, FPW, 1, __VA_ARGS__>}
#define FFTFREE_C2C(L, E, FPW, RS, ...)                                                   \
  {L, {L, E, FPW, RS}, &c2c_launch_impl<L, E, FPW, -1, __VA_ARGS__>,                     \
   &c2c_launch_impl<L, E, FPW, 1, __VA_ARGS__>}

// Real-input variants per N (first one per N is the default).
const StftEntry kStft[] = {
    FFTFREE_STFT(8, 8, 64, "8", 8),
    FFTFREE_STFT(16, 16, 64, "16", 16),
    FFTFREE_STFT(32, 8, 64, "8x4", 8, 4),
    FFTFREE_STFT(64, 8, 32, "8x8", 8, 8),
    FFTFREE_STFT(64, 16, 64, "16x4", 16, 4),
    FFTFREE_STFT(128, 16, 32, "16x8", 16, 8),
    FFTFREE_STFT(128, 8, 16, "8x8x2", 8, 8, 2),
    FFTFREE_STFT(256, 16, 16, "16x16", 16, 16),
    FFTFREE_STFT(256, 8, 8, "8x8x4", 8, 8, 4),
    FFTFREE_STFT(512, 8, 4, "8x8x8", 8, 8, 8),
    FFTFREE_STFT(512, 16, 8, "16x16x2", 16, 16, 2),
    FFTFREE_STFT(512, 8, 8, "8x8x8/8", 8, 8, 8),
    FFTFREE_STFT(1024, 16, 4, "16x16x4", 16, 16, 4),
    FFTFREE_STFT(1024, 8, 2, "8x8x8x2", 8, 8, 8, 2),
    FFTFREE_STFT(2048, 16, 2, "16x16x8", 16, 16, 8),
    FFTFREE_STFT(2048, 8, 1, "8x8x8x4", 8, 8, 8, 4),
    FFTFREE_STFT(4096, 16, 1, "16x16x16", 16, 16, 16),
};

const C2CEntry kC2C[] = {
    FFTFREE_C2C(8, 8, 64, "8", 8),
    FFTFREE_C2C(16, 16, 64, "16", 16),
    FFTFREE_C2C(32, 8, 64, "8x4", 8, 4),
    FFTFREE_C2C(64, 8, 32, "8x8", 8, 8),
    FFTFREE_C2C(128, 16, 32, "16x8", 16, 8),
    FFTFREE_C2C(256, 16, 16, "16x16", 16, 16),
    FFTFREE_C2C(512, 8, 4, "8x8x8", 8, 8, 8),
    FFTFREE_C2C(1024, 16, 4, "16x16x4", 16, 16, 4),
    FFTFREE_C2C(1024, 8, 2, "8x8x8x2", 8, 8, 8, 2),
    FFTFREE_C2C(2048, 16, 2, "16x16x8", 16, 16, 8),
    FFTFREE_C2C(4096, 16, 1, "16x16x16", 16, 16, 16),
};

#undef FFTFREE_STFT
#undef FFTFREE_C2C

template <class Entry, size_t K>
const Entry* find_variant(const Entry (&tab)[K], int n, int idx) {
  int seen = 0;
  for (size_t i = 0; i < K; ++i) {
    if (tab[i].n != n) continue;
    if (seen == idx) return &tab[i];
    ++seen;
  }
  return nullptr;
}

template <class Entry, size_t K>
int count_variants(const Entry (&tab)[K], int n) {
  int c = 0;
  for (size_t i = 0; i < K; ++i) c += (tab[i].n == n);
  return c;
}

constexpr int kTinyBlock = 256;

}  // namespace

int stft_max_lds_n() { return 8192; }
int c2c_max_lds_n() { return 4096; }
int stft_variant_count(int n) { return count_variants(kStft, n); }
const VariantInfo* stft_variant(int n, int idx) {
  const StftEntry* e = find_variant(kStft, n, idx);
  return e ? &e->info : nullptr;
}
int c2c_variant_count(int n) { return count_variants(kC2C, n); }
const VariantInfo* c2c_variant(int n, int idx) {
  const C2CEntry* e = find_variant(kC2C, n, idx);
  return e ? &e->info : nullptr;
}

hipError_t launch_stft(const StftArgs& a, int inverse, int variant, hipStream_t s) {
  if (a.frames == 0) return hipSuccess;
  if (a.n < 16) {
    const size_t total = a.frames * (size_t)a.bins;
    const size_t grid = (total + kTinyBlock - 1) / kTinyBlock;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    if (inverse)
      hipLaunchKernelGGL(stft_tiny_kernel<1>, dim3((unsigned)grid), dim3(kTinyBlock), 0, s, a);
    else
      hipLaunchKernelGGL(stft_tiny_kernel<-1>, dim3((unsigned)grid), dim3(kTinyBlock), 0, s, a);
    return hipGetLastError();
  }
  const StftEntry* e = find_variant(kStft, a.n, variant < 0 ? 0 : variant);
  if (!e) return hipErrorInvalidValue;
  return inverse ? e->inv(a, s) : e->fwd(a, s);
}

hipError_t launch_c2c(const C2CArgs& a, int inverse, int variant, hipStream_t s) {
  if (a.batch == 0) return hipSuccess;
  if (a.n < 8) {
    const size_t total = a.batch * (size_t)a.n;
    // whole transforms per block so a transform never straddles waves (in == out safe)
    const size_t grid = (total + kTinyBlock - 1) / kTinyBlock;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    if (inverse)
      hipLaunchKernelGGL(c2c_tiny_kernel<1>, dim3((unsigned)grid), dim3(kTinyBlock), 0, s, a);
    else
      hipLaunchKernelGGL(c2c_tiny_kernel<-1>, dim3((unsigned)grid), dim3(kTinyBlock), 0, s, a);
    return hipGetLastError();
  }
  const C2CEntry* e = find_variant(kC2C, a.n, variant < 0 ? 0 : variant);
  if (!e) return hipErrorInvalidValue;
  return inverse ? e->inv(a, s) : e->fwd(a, s);
}

hipError_t launch_frame_complex(const StftArgs& a, float2* out, hipStream_t s) {
  const size_t total = a.frames * (size_t)a.n;
  if (total == 0) return hipSuccess;
  const size_t grid = (total + 255) / 256;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(frame_complex_kernel, dim3((unsigned)grid), dim3(256), 0, s, a, out);
  return hipGetLastError();
}

hipError_t launch_planes(const float2* in, size_t frames, int n, int bins, int polar, float* o_re, float* o_im,
                         float* o_mag, hipStream_t s) {
  const size_t total = frames * (size_t)bins;
  if (total == 0) return hipSuccess;
  const size_t grid = (total + 255) / 256;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(planes_kernel, dim3((unsigned)grid), dim3(256), 0, s, in, frames, n, bins, polar, o_re, o_im,
                     o_mag);
  return hipGetLastError();
}

hipError_t launch_metadata(int n, int kernel, int kind, int r2c_half, const float2* tw, int inverse, float2* out,
                           size_t capacity, size_t* count, hipStream_t s) {
  int S = 0;
  while ((1 << S) < n) ++S;
  size_t c = 0;
  switch (kind) {
    case kMetaTwiddle: c = (size_t(1) << S) - 1; break;
    case kMetaLayoutMap: c = (size_t)n; break;
    case kMetaStageParams: c = (size_t)S; break;
    case kMetaButterflyPairs:
    case kMetaTwiddleIndexMap: c = (size_t)n * (size_t)S; break;
    case kMetaOmega: c = (kernel == 2) ? 0 : (r2c_half ? (size_t)(n / 2 + 1) : (size_t)n); break;
    default: c = 0; break;
  }
  *count = c;
  if (c == 0 || out == nullptr) return hipSuccess;  // size query
  if (c > capacity || c > 0x7fffffffull) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((c + 255) / 256);
  hipLaunchKernelGGL(metadata_kernel, dim3(grid), dim3(256), 0, s, n, S, kernel, kind, (int)c, tw, inverse, out,
                     (int)c);
  return hipGetLastError();
}

}  // namespace fftfree

// fft_kernels.hip -- gfx950 batched complex transform and metadata kernels.
//
// c2c_kernel        batched complex DFT of size N with the pre/post hooks
//                   (eigen_fft.hpp:2484-2596), the C++ fft_inplace_batched path
//                   (eigen_fft.hpp:3152-3175).
// c2c_tiny_kernel   direct DFT for N too small for a multi-thread Stockham pass.
// metadata_kernel   the reference's integer metadata maps from closed forms
//                   (eigen_fft.hpp:1529-1888 writers as called by both kernels).
//
// Complex batches are transform-major float2. One workgroup owns FPW whole transforms:
// nothing crosses workgroups, so no atomics.
#include "fft_device.hpp"
#include "fft_kernels.hpp"

namespace fftfree {
using namespace dev;

namespace {

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
    X = make_float2(m, a.post_mode == 3 ? fast_atan2f(X.y, X.x) : 0.0f);
  } else if (a.post_mode == 1) {
    X.y = 0.0f;
  }
  return X;
}

template <int L, int E, int FPW, int DIR, int... Rs>
__global__ void __launch_bounds__(FPW*(L / E)) c2c_kernel(C2CArgs a) {
  constexpr int T = L / E;
  constexpr int NP = sizeof...(Rs);
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int LP = padded_len<L>();
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 lds[2][FPW][LP];

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
  float2 twr[E * NP];
  StockhamPlan<L, E, DIR, Rs...>::load_twiddles(twr, t, a.tw, 1);
  float2* free_buf;
  StockhamPlan<L, E, DIR, Rs...>::run(v, t, &lds[0][tr][0], &lds[1][tr][0], twr, free_buf);
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

// The 512-point transform with the exchanges between its three radix-8 passes in
// registers (WaveRegs888: DPP and permlane swaps, no LDS): one transform per wave, FPW per
// workgroup. Same butterflies and twiddles as c2c_kernel<512, 8, ..., 8, 8, 8>, so the same
// bits; lane l ends holding X[rot_lane(l) + 64 r], which it stores.
template <int FPW, int DIR>
__global__ void __launch_bounds__(FPW * 64) c2c_regx512_kernel(C2CArgs a) {
  constexpr int L = 512;
  const int tr = threadIdx.x >> 6;  // wave-uniform: a wave is one transform
  const int t = threadIdx.x & 63;
  const size_t tile = (size_t)xcd_tile(blockIdx.x, gridDim.x);
  const size_t f = tile * FPW + tr;
  const bool active = f < a.batch;
  const float2* __restrict__ col = a.in + (active ? f : 0) * (size_t)L;
  float2 v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = active ? load_pre(col, t + 64 * r, L, a.pre_mode, a.pre_polar) : make_float2(0.0f, 0.0f);
  float2 twr[24];
  StockhamPlan<L, 8, DIR, 8, 8, 8>::load_twiddles(twr, rot_lane(t), a.tw, 1);
  WaveRegs888<DIR>::run(v, t, nullptr, twr, [](int) {});
  if (active) {
    float2* __restrict__ dst = a.out + f * (size_t)L;
    const int c = rot_lane(t);
#pragma unroll
    for (int r = 0; r < 8; ++r) dst[c + 64 * r] = post_op(v[r], c + 64 * r, L, a);
  }
}

// The 1024-point transform on one wave (16 values per lane, radices 16, 16, 4) with the
// exchanges in registers (WaveRegs16x16x4): no LDS, no barriers, FPW transforms per
// workgroup. Same butterflies and twiddles as c2c_kernel<1024, 16, ..., 16, 16, 4>.
template <int FPW, int DIR>
__global__ void __launch_bounds__(FPW * 64) c2c_regx1024_kernel(C2CArgs a) {
  constexpr int L = 1024;
  const int tr = threadIdx.x >> 6;  // wave-uniform: a wave is one transform
  const int t = threadIdx.x & 63;
  const size_t tile = (size_t)xcd_tile(blockIdx.x, gridDim.x);
  const size_t f = tile * FPW + tr;
  const bool active = f < a.batch;
  const float2* __restrict__ col = a.in + (active ? f : 0) * (size_t)L;
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = active ? load_pre(col, t + 64 * r, L, a.pre_mode, a.pre_polar) : make_float2(0.0f, 0.0f);
  float2 twr[48];
  WaveRegs16x16x4<DIR>::load_twiddles(twr, t, a.tw, 1);
  WaveRegs16x16x4<DIR>::run(v, t, twr, [](int) {});
  if (active) {
    float2* __restrict__ dst = a.out + f * (size_t)L;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int idx = regx1024_pass2(t, q) + 256 * r;
        dst[idx] = post_op(v[4 * q + r], idx, L, a);
      }
  }
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

template <int L, int E, int FPW, int DIR, int... Rs>
hipError_t c2c_launch_impl(const C2CArgs& a, hipStream_t s) {
  const size_t grid = (a.batch + FPW - 1) / FPW;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((c2c_kernel<L, E, FPW, DIR, Rs...>), dim3((unsigned)grid), dim3(FPW * (L / E)), 0, s, a);
  return hipGetLastError();
}

template <int FPW, int DIR>
hipError_t c2c_regx512_launch(const C2CArgs& a, hipStream_t s) {
  const size_t grid = (a.batch + FPW - 1) / FPW;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((c2c_regx512_kernel<FPW, DIR>), dim3((unsigned)grid), dim3(FPW * 64), 0, s, a);
  return hipGetLastError();
}
template <int FPW, int DIR>
hipError_t c2c_regx1024_launch(const C2CArgs& a, hipStream_t s) {
  const size_t grid = (a.batch + FPW - 1) / FPW;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((c2c_regx1024_kernel<FPW, DIR>), dim3((unsigned)grid), dim3(FPW * 64), 0, s, a);
  return hipGetLastError();
}

using C2CFn = hipError_t (*)(const C2CArgs&, hipStream_t);

struct C2CEntry {
  int n;
  VariantInfo info;
  C2CFn fwd, inv;
};

#define FFTFREE_C2C(L, E, FPW, RS, ...)                                                   \
  {L, {L, E, FPW, RS}, &c2c_launch_impl<L, E, FPW, -1, __VA_ARGS__>,                     \
   &c2c_launch_impl<L, E, FPW, 1, __VA_ARGS__>}

const C2CEntry kC2C[] = {
    FFTFREE_C2C(8, 8, 64, "8", 8),
    FFTFREE_C2C(16, 16, 64, "16", 16),
    FFTFREE_C2C(32, 8, 64, "8x4", 8, 4),
    FFTFREE_C2C(64, 8, 32, "8x8", 8, 8),
    FFTFREE_C2C(128, 16, 32, "16x8", 16, 8),
    FFTFREE_C2C(256, 16, 16, "16x16", 16, 16),
    // exchanges in registers: 136 us against 160 us for the LDS kernel at 65536 transforms
    {512, {512, 8, 4, "8x8x8 dpp"}, &c2c_regx512_launch<4, -1>, &c2c_regx512_launch<4, 1>},
    FFTFREE_C2C(512, 8, 4, "8x8x8", 8, 8, 8),
    FFTFREE_C2C(1024, 8, 2, "8x8x8x2", 8, 8, 8, 2),  // 13% faster at 65533 transforms, equal at 64
    FFTFREE_C2C(1024, 16, 4, "16x16x4", 16, 16, 4),
    {1024, {1024, 16, 1, "16x16x4 dpp"}, &c2c_regx1024_launch<1, -1>, &c2c_regx1024_launch<1, 1>},
    {1024, {1024, 16, 4, "16x16x4 dpp"}, &c2c_regx1024_launch<4, -1>, &c2c_regx1024_launch<4, 1>},
    FFTFREE_C2C(2048, 16, 2, "16x16x8", 16, 16, 8),
    FFTFREE_C2C(4096, 16, 1, "16x16x16", 16, 16, 16),
};

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

int c2c_max_lds_n() { return 4096; }
int c2c_variant_count(int n) { return count_variants(kC2C, n); }
const VariantInfo* c2c_variant(int n, int idx) {
  const C2CEntry* e = find_variant(kC2C, n, idx);
  return e ? &e->info : nullptr;
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

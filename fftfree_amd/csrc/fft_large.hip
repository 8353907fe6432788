// fft_large.hip -- batched complex DFT for N beyond one workgroup's LDS (N = N1 * N2).
//
// With n = n1*N2 + n2 and k = k1 + N1*k2:
//   X[k1 + N1 k2] = sum_{n2} W_N2^{n2 k2} [ W_N^{n2 k1} sum_{n1} x[n1 N2 + n2] W_N1^{n1 k1} ]
// Pass A transforms the N2 columns (length N1, stride N2) and multiplies by W_N^{n2 k1};
// pass B transforms the N1 rows (length N2, contiguous) and writes X[k1 + N1 k2].
// Both passes hold whole lines in LDS. Column data and pass B's output are strided in
// HBM, so a workgroup takes C adjacent lines and moves them through LDS: global reads
// and writes are runs of C consecutive complex values. The inter-pass twiddle is the
// product of two entries of the plan's FP64-built table (a few KB of it stay hot):
// W_N^m = W_N^{m with its low S bits cleared} * W_N^{m mod 2^S}.
//
// The default kernels (col_line_kernel / row_line_kernel) give each line one wave of 64
// lanes (16 values each), so the passes inside a line need no workgroup barrier, and read
// the per-pass twiddles from an LDS table rather than registers: about 110 VGPRs, so two
// 8-line workgroups of N1 = 1024 share a CU. The host runs the batch in chunks whose
// pass-A output (about 32 MiB) is still in the caches when pass B reads it, alternating
// two streams so that one chunk's pass A overlaps the other's pass B.
//
// The reference runs every N through lgN radix-2 stages in place (eigen_fft.hpp:2902-3057
// Cooley-Tukey, 3472-3532 Stockham); the sum computed is the same DFT.
#include "fft_device.hpp"
#include "fft_kernels.hpp"

#include <algorithm>
#include <cstdio>

namespace fftfree {
using namespace dev;

int device_cus();

namespace {

// Pre-transform hook on load: mode 1 clears imag (R2C).
__device__ __forceinline__ float2 pre1(float2 z, int mode) { return mode == 1 ? make_float2(z.x, 0.0f) : z; }

// Post-transform hook (eigen_fft.hpp:2535-2596) at global frequency idx of total length N.
__device__ __forceinline__ float2 post_large(float2 X, size_t idx, size_t N, const C2CArgs& a) {
  X = c_scale(X, a.scale);
  if (a.post_half) {
    if (idx == 0 || idx == (N >> 1)) X.y = 0.0f;
    if (idx > (N >> 1)) X = make_float2(0.0f, 0.0f);
  }
  if (a.post_mode >= 2 && idx < (size_t)a.post_upper) {
    const float m = __builtin_sqrtf(__builtin_fmaf(X.x, X.x, X.y * X.y));
    X = make_float2(m, a.post_mode == 3 ? fast_atan2f(X.y, X.x) : 0.0f);
  } else if (a.post_mode == 1) {
    X.y = 0.0f;
  }
  return X;
}

// Pre-transform hook of C2R half-spectrum plans (eigen_fft.hpp:2505-2531): bins 0..N/2
// (optionally (mag, phase)) become the full Hermitian spectrum. Thread n <= N/2 reads
// src[n] only and writes dst[n] and dst[N-n], so src == dst is safe.
__global__ void __launch_bounds__(256) hermitian_rebuild_kernel(const float2* in, float2* out, size_t n, int polar) {
  const size_t h = n >> 1;
  const size_t b = blockIdx.y;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i > h) return;
  const float2* __restrict__ src = in + b * n;
  float2* __restrict__ dst = out + b * n;
  float2 z = src[i];
  if (polar) {
    float sn, cs;
    sincosf(z.y, &sn, &cs);
    z = make_float2(z.x * cs, z.x * sn);
  }
  if (i == 0 || i == h) z.y = 0.0f;
  dst[i] = z;
  if (i != 0 && i != h) dst[n - i] = c_conj(z);
}

struct LargeArgs {
  const float2* in;
  float2* out;
  size_t batch;
  size_t n;            // N
  int n1, n2;          // column length, row length
  const float2* tw;    // forward table exp(-2 pi i m / N), m < N
  int s;               // S
  C2CArgs hooks;       // pre_mode (pass A), scale/post (pass B)
};

template <int DIR>
__device__ __forceinline__ float2 big_twiddle(const LargeArgs& a, size_t m) {
  const size_t mask = (size_t(1) << a.s) - 1;
  const float2 hi = load_tw<DIR>(a.tw, (int)(m & ~mask));
  const float2 lo = load_tw<DIR>(a.tw, (int)(m & mask));
  return c_mul(hi, lo);
}

// Pass A: C adjacent columns (length L = N1, stride N2) per workgroup.
template <int L, int E, int C, int DIR, int... Rs>
__global__ void __launch_bounds__(C*(L / E)) col_pass_kernel(LargeArgs a) {
  constexpr int T = L / E;
  constexpr int NT = C * T;
  constexpr int NP = sizeof...(Rs);
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int LP = padded_len<L>();
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 lds[C][LP];

  const int tid = threadIdx.x;
  // 1-D grid of batch * (N2 / C) tiles; each XCD walks a contiguous run of tiles, so the
  // neighbouring column groups that share HBM lines are read through the same L2
  const int groups = a.n2 / C;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const size_t b = (size_t)(tile / groups);
  const size_t c0 = (size_t)(tile - (int)b * groups) * C;
  const size_t N2 = (size_t)a.n2;
  const float2* __restrict__ src = a.in + b * a.n + c0;
  float2* __restrict__ dst = a.out + b * a.n + c0;

  // coalesced column load: thread = (col = tid % C, rows tid / C + k*T)
  {
    const int col = tid % C;
    const int row0 = tid / C;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int j = row0 + k * (NT / C);
      lds[col][pidx(j)] = pre1(src[(size_t)j * N2 + col], a.hooks.pre_mode);
    }
  }
  __syncthreads();

  const int slot = tid / T;
  const int t = tid - slot * T;
  float2* buf = &lds[slot][0];
  float2 v[E];
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) v[q * R0 + r] = buf[pidx(t + q * T + r * (L / R0))];
  float2 twr[E * NP];
  StockhamPlan<L, E, DIR, Rs...>::load_twiddles(twr, t, a.tw, (int)(a.n / L));
  __syncthreads();  // every line read before the first in-place scatter
  float2* Zb;
  StockhamPlan<L, E, DIR, Rs...>::template run<true>(v, t, buf, buf, twr, Zb);
  // twiddle W_N^{n2 k1} and natural-order store into LDS
  const size_t n2 = c0 + slot;
#pragma unroll
  for (int q = 0; q < E / Rl; ++q)
#pragma unroll
    for (int r = 0; r < Rl; ++r) {
      const int k1 = t + q * T + r * (L / Rl);
      buf[pidx(k1)] = c_mul(v[q * Rl + r], big_twiddle<DIR>(a, n2 * (size_t)k1));
    }
  __syncthreads();
  {
    const int col = tid % C;
    const int row0 = tid / C;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int j = row0 + k * (NT / C);
      dst[(size_t)j * N2 + col] = lds[col][pidx(j)];
    }
  }
}

// Pass B: C adjacent rows (length L = N2, contiguous) per workgroup; output element
// (row k1, k2) goes to k1 + N1*k2, written as runs of C consecutive values.
template <int L, int E, int C, int DIR, int... Rs>
__global__ void __launch_bounds__(C*(L / E)) row_pass_kernel(LargeArgs a) {
  constexpr int T = L / E;
  constexpr int NT = C * T;
  constexpr int NP = sizeof...(Rs);
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int LP = padded_len<L>();
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 lds[C][LP];

  const int tid = threadIdx.x;
  const int groups = a.n1 / C;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const size_t b = (size_t)(tile / groups);
  const size_t r0 = (size_t)(tile - (int)b * groups) * C;
  const size_t N1 = (size_t)a.n1;
  const int slot = tid / T;
  const int t = tid - slot * T;
  float2* buf = &lds[slot][0];
  const float2* __restrict__ row = a.in + b * a.n + (r0 + slot) * (size_t)L;

  float2 v[E];
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) v[q * R0 + r] = row[t + q * T + r * (L / R0)];
  float2 twr[E * NP];
  StockhamPlan<L, E, DIR, Rs...>::load_twiddles(twr, t, a.tw, (int)(a.n / L));
  float2* Zb;
  StockhamPlan<L, E, DIR, Rs...>::template run<true>(v, t, buf, buf, twr, Zb);
  const size_t k1 = r0 + slot;
#pragma unroll
  for (int q = 0; q < E / Rl; ++q)
#pragma unroll
    for (int r = 0; r < Rl; ++r) {
      const int k2 = t + q * T + r * (L / Rl);
      buf[pidx(k2)] = post_large(v[q * Rl + r], k1 + N1 * (size_t)k2, a.n, a.hooks);
    }
  __syncthreads();
  float2* __restrict__ dst = a.out + b * a.n + r0;
  {
    const int col = tid % C;
    const int row0 = tid / C;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int k2 = row0 + k * (NT / C);
      dst[(size_t)k2 * N1 + col] = lds[col][pidx(k2)];
    }
  }
}

// ---- line kernels with LDS twiddle tables ------------------------------------------
//
// One line (a column of pass A, a row of pass B) is held by T threads, E = L/T values
// each; C lines per workgroup. Exchanges between passes go through the line's LDS
// buffer; with T = 64 a line is one wave and needs no workgroup barrier. The pass
// twiddles come from per-workgroup LDS tables, not registers: pass P (span Ns > 1,
// radix R) has table[m*(R-1) + r-1] = exp(DIR 2 pi i m r / (Ns R)), m < Ns. So the
// kernels fit 128 VGPRs and two workgroups share a CU.

// Line stride of the C-line staging buffers, in float2: the padded line plus 8, so that
// stride = 8 (mod 32). The transposing accesses (lane = column c + C * row) then put the C
// lines 16 dwords apart in the banks; with a stride that is a multiple of 32 every column
// of a row hits the same bank pair.
template <int L>
constexpr int staged_stride() {
  constexpr int p = padded_len<L>();
  return p + ((8 - p % 32) + 32) % 32;
}
static_assert(staged_stride<1024>() % 32 == 8 && staged_stride<2048>() % 32 == 8, "stride 8 mod 32");

// Twiddle-table floats of the passes of Rs (pass 0, span 1, has none).
template <int Ns, int R, int... Rest>
constexpr int line_tw_size() {
  constexpr int own = Ns > 1 ? Ns * (R - 1) : 0;
  if constexpr (sizeof...(Rest) == 0) return own;
  else return own + line_tw_size<Ns * R, Rest...>();
}

// HT: one table tab[j] = exp(DIR 2 pi i j / L), j < L/2, for every pass, with
// W_L^{j + L/2} = -W_L^j: half the LDS of the per-pass tables, a few more VALU per twiddle.
template <int L, bool HT, int... Rs>
constexpr int line_tab_size() {
  if constexpr (HT) return L / 2;
  else return line_tw_size<1, Rs...>();
}

template <int L, int DIR>
__device__ __forceinline__ void fill_half_table(float2* tab, const float2* __restrict__ tw, int stride, int tid,
                                                int nt) {
  for (int i = tid; i < L / 2; i += nt) tab[i] = load_tw<DIR>(tw, i * stride);
}

template <int L>
__device__ __forceinline__ float2 half_table_tw(const float2* tab, int idx) {
  static_assert((L & (L - 1)) == 0, "power of two");
  const float2 w = tab[idx & (L / 2 - 1)];
  const unsigned sg = (unsigned)(idx & (L / 2)) << (31 - __builtin_ctz(L / 2));
  return make_float2(__uint_as_float(__float_as_uint(w.x) ^ sg), __uint_as_float(__float_as_uint(w.y) ^ sg));
}

template <int L, int DIR, int Ns, int OFF, int R, int... Rest>
struct LineTw {
  // fill every table; tw is exp(-2 pi i m / (L*stride))
  __device__ static __forceinline__ void fill(float2* tab, const float2* __restrict__ tw, int stride, int tid,
                                              int nt) {
    if constexpr (Ns > 1) {
      for (int i = tid; i < Ns * (R - 1); i += nt) {
        const int m = i / (R - 1), r = i - m * (R - 1) + 1;
        tab[OFF + i] = load_tw<DIR>(tw, m * (L / R / Ns) * stride * r);
      }
    }
    if constexpr (sizeof...(Rest) > 0)
      LineTw<L, DIR, Ns * R, OFF + (Ns > 1 ? Ns * (R - 1) : 0), Rest...>::fill(tab, tw, stride, tid, nt);
  }
};

// The passes of one line; result in last-pass order v[q*Rl + r] = X[t + q*T + r*L/Rl].
template <int L, int E, int DIR, bool HT, int Ns, int OFF, int R, int... Rest>
struct LineStockham {
  static constexpr int T = L / E;
  template <class Sync>
  __device__ static __forceinline__ void run(float2* v, int t, float2* buf, const float2* tab, Sync&& sync) {
#pragma unroll
    for (int q = 0; q < E / R; ++q) {
      float2* u = v + q * R;
      if constexpr (Ns > 1) {
        if constexpr (HT) {
          const int base = ((t + q * T) & (Ns - 1)) * (L / R / Ns);
#pragma unroll
          for (int r = 1; r < R; ++r) u[r] = c_mul(u[r], half_table_tw<L>(tab, base * r));
        } else {
          const float2* w = tab + OFF + ((t + q * T) & (Ns - 1)) * (R - 1);
#pragma unroll
          for (int r = 1; r < R; ++r) u[r] = c_mul(u[r], w[r - 1]);
        }
      }
      reg_dft<R, DIR>(u);
    }
    if constexpr (sizeof...(Rest) > 0) {
#pragma unroll
      for (int q = 0; q < E / R; ++q) {
        const int bb = t + q * T;
        const int k = bb & (Ns - 1);
        const int o = (bb - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[pidx(o + r * Ns)] = v[q * R + r];
      }
      sync();
      constexpr int Rn = FirstOf<Rest...>::value;
#pragma unroll
      for (int q = 0; q < E / Rn; ++q) {
        const int bb = t + q * T;
#pragma unroll
        for (int r = 0; r < Rn; ++r) v[q * Rn + r] = buf[pidx(bb + r * (L / Rn))];
      }
      sync();
      LineStockham<L, E, DIR, HT, Ns * R, OFF + (Ns > 1 ? Ns * (R - 1) : 0), Rest...>::run(v, t, buf, tab, sync);
    }
  }
};

// Pass A with LDS twiddles: C adjacent columns (length L = N1, stride N2) per workgroup.
// The inter-pass twiddle W_N^{n2 k1} with k1 = t + T*j is W_N^{n2 t} * W_N^{n2 T j}, the
// second factor the same for the whole line; each factor is read as the product of the
// table's hot subsets (big_twiddle), so the whole N-entry table is never swept. Pass A's
// output is read back by pass B: ordinary stores.
template <int L, int E, int C, bool HT, int DIR, int... Rs>
__global__ void __launch_bounds__(C*(L / E)) col_line_kernel(LargeArgs a) {
  constexpr int T = L / E;
  constexpr int NT = C * T;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int LP = staged_stride<L>();
  constexpr int TWN = line_tab_size<L, HT, Rs...>();
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  static_assert(T % 64 == 0, "whole waves per line");
  __shared__ float2 lds[C][LP];
  __shared__ float2 tab[TWN > 0 ? TWN : 1];

  const int tid = threadIdx.x;
  const int groups = a.n2 / C;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const size_t b = (size_t)(tile / groups);
  const size_t c0 = (size_t)(tile - (int)b * groups) * C;
  const size_t N2 = (size_t)a.n2;
  const float2* __restrict__ src = a.in + b * a.n + c0;
  float2* __restrict__ dst = a.out + b * a.n + c0;

  if constexpr (HT) fill_half_table<L, DIR>(tab, a.tw, (int)(a.n / L), tid, NT);
  else LineTw<L, DIR, 1, 0, Rs...>::fill(tab, a.tw, (int)(a.n / L), tid, NT);
  {  // coalesced column load: thread = (column tid % C, rows tid / C + k * NT / C)
    const int col = tid % C;
    const int row0 = tid / C;
    float2 x[E];
#pragma unroll
    for (int k = 0; k < E; ++k) x[k] = src[(size_t)(row0 + k * (NT / C)) * N2 + col];
#pragma unroll
    for (int k = 0; k < E; ++k) lds[col][pidx(row0 + k * (NT / C))] = pre1(x[k], a.hooks.pre_mode);
  }
  __syncthreads();

  const int line = __builtin_amdgcn_readfirstlane(tid / T);
  const int t = tid - line * T;
  float2* buf = &lds[line][0];
  auto sync = [&]() {
    if constexpr (T == 64) wave_lds_fence();
    else __syncthreads();
  };
  float2 v[E];
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) v[q * R0 + r] = buf[pidx(t + q * T + r * (L / R0))];
  sync();  // every value of the line read before the first in-place scatter
  LineStockham<L, E, DIR, HT, 1, 0, Rs...>::run(v, t, buf, tab, sync);
  const size_t n2 = c0 + (size_t)line;
  const float2 wt = big_twiddle<DIR>(a, n2 * (size_t)t);
  sync();  // the last pass's reads of buf are done before the natural-order writes
#pragma unroll
  for (int q = 0; q < E / Rl; ++q)
#pragma unroll
    for (int r = 0; r < Rl; ++r) {
      const int j = q + r * (E / Rl);  // k1 = t + T*j
      const float2 wj = big_twiddle<DIR>(a, n2 * (size_t)(T * j));
      buf[pidx(t + T * j)] = c_mul(v[q * Rl + r], c_mul(wt, wj));
    }
  __syncthreads();
  {
    const int col = tid % C;
    const int row0 = tid / C;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int j = row0 + k * (NT / C);
      dst[(size_t)j * N2 + col] = lds[col][pidx(j)];
    }
  }
}

// Pass B with LDS twiddles: C adjacent rows (length L = N2, contiguous) per workgroup;
// output (row k1, k2) goes to k1 + N1*k2 through LDS as runs of C values. NTS: streaming
// stores; otherwise ordinary ones, which let L2 merge the partial lines of neighbouring
// workgroups.
template <int L, int E, int C, bool HT, bool NTS, int DIR, int... Rs>
__global__ void __launch_bounds__(C*(L / E)) row_line_kernel(LargeArgs a) {
  constexpr int T = L / E;
  constexpr int NT = C * T;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int LP = staged_stride<L>();
  constexpr int TWN = line_tab_size<L, HT, Rs...>();
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  static_assert(T % 64 == 0, "whole waves per line");
  __shared__ float2 lds[C][LP];
  __shared__ float2 tab[TWN > 0 ? TWN : 1];

  const int tid = threadIdx.x;
  const int groups = a.n1 / C;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const size_t b = (size_t)(tile / groups);
  const size_t r0 = (size_t)(tile - (int)b * groups) * C;
  const size_t N1 = (size_t)a.n1;
  const int line = __builtin_amdgcn_readfirstlane(tid / T);
  const int t = tid - line * T;
  float2* buf = &lds[line][0];
  const float2* __restrict__ row = a.in + b * a.n + (r0 + line) * (size_t)L;

  float2 v[E];
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) v[q * R0 + r] = row[t + q * T + r * (L / R0)];
  if constexpr (HT) fill_half_table<L, DIR>(tab, a.tw, (int)(a.n / L), tid, NT);
  else LineTw<L, DIR, 1, 0, Rs...>::fill(tab, a.tw, (int)(a.n / L), tid, NT);
  __syncthreads();
  auto sync = [&]() {
    if constexpr (T == 64) wave_lds_fence();
    else __syncthreads();
  };
  LineStockham<L, E, DIR, HT, 1, 0, Rs...>::run(v, t, buf, tab, sync);
  const size_t k1 = r0 + (size_t)line;
  sync();
#pragma unroll
  for (int q = 0; q < E / Rl; ++q)
#pragma unroll
    for (int r = 0; r < Rl; ++r) {
      const int k2 = t + q * T + r * (L / Rl);
      buf[pidx(k2)] = post_large(v[q * Rl + r], k1 + N1 * (size_t)k2, a.n, a.hooks);
    }
  __syncthreads();
  float2* __restrict__ dst = a.out + b * a.n + r0;
  {
    const int col = tid % C;
    const int row0 = tid / C;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int k2 = row0 + k * (NT / C);
      const float2 z = lds[col][pidx(k2)];
      if constexpr (NTS) {
        float* o = reinterpret_cast<float*>(dst + (size_t)k2 * N1 + col);
        __builtin_nontemporal_store(z.x, o);
        __builtin_nontemporal_store(z.y, o + 1);
      } else {
        dst[(size_t)k2 * N1 + col] = z;
      }
    }
  }
}

// Persistent, software-pipelined passes (one workgroup per CU, one wave per line): the
// workgroups of XCD x walk the contiguous run of tiles [lo_x, hi_x) with stride = the
// XCD's workgroup count, so the tiles in flight on one XCD are neighbours (sharing HBM
// lines through one L2). While the waves transform tile i, the global loads of the tile
// after it are in flight into registers; the twiddle tables are filled once per
// workgroup. Every global load the transform needs is part of that prefetch (the
// inter-pass twiddle factors too), so no wait inside the transform waits for the next
// tile's data. Measured on the 2^20 x 16 round trip (variant "pipelined"): 309-312 us
// with one pass pipelined, 330 us with both, against 297 us for the one-tile
// workgroups at two per CU, whose second workgroup (and the other stream's launch)
// overlaps memory with arithmetic as well; the prefetch's registers leave room for
// one 8-wave workgroup per CU only.
__device__ __forceinline__ void xcd_range(int ntiles, int* lo, int* hi, int* step) {
  const int x = blockIdx.x & 7;
  const int q8 = ntiles >> 3, r8 = ntiles & 7;
  *lo = x * q8 + (x < r8 ? x : r8);
  *hi = *lo + q8 + (x < r8 ? 1 : 0);
  *step = ((int)gridDim.x - x + 7) >> 3;  // workgroups dealt to XCD x
  *lo += (int)(blockIdx.x >> 3);
}

// Workgroup barrier that orders LDS only: the pipelined passes exchange nothing through
// global memory inside a launch, so the barrier need not wait for the stores of the last
// tile or the loads of the next one (a plain __syncthreads() waits for both).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// s_waitcnt vmcnt(0) alone (gfx9 encoding: expcnt and lgkmcnt at their maxima). The
// pipelined passes place it after the transform, before the tile's stores are issued:
// the next tile's loads (issued before the transform) and the last tile's stores are
// then complete, and nothing later waits on a counter that mixes loads with the stores
// just issued (such a wait would be for all of them).
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ float2 readlane2(float2 v, int j) {
  return make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), j)),
                     __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), j)));
}

// Pass A, pipelined: tile = C adjacent columns (length L, stride N2) of one transform.
template <int L, int E, int C, int DIR, int... Rs>
__global__ void __launch_bounds__(C*(L / E), 1) col_pipe_kernel(LargeArgs a) {
  constexpr int T = L / E;
  constexpr int NT = C * T;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int LP = staged_stride<L>();
  constexpr int TWN = line_tab_size<L, false, Rs...>();
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  static_assert(T == 64 && E <= 64, "one wave per line; the line factors of E rows fit a wave");
  __shared__ float2 lds[C][LP];
  __shared__ float2 tab[TWN > 0 ? TWN : 1];

  const int tid = threadIdx.x;
  const int groups = a.n2 / C;
  int tile, hi, step;
  xcd_range(groups * (int)a.batch, &tile, &hi, &step);
  const size_t N2 = (size_t)a.n2;
  const int col = tid % C, row0 = tid / C;
  const int line = __builtin_amdgcn_readfirstlane(tid / T);
  const int t = tid - line * T;
  float2* buf = &lds[line][0];
  const size_t mask = (size_t(1) << a.s) - 1;

  float2 nx[E];
  float2 th, tl, uh, ul;  // W_N^{n2 t} = th * tl; W_N^{n2 T j} = uh * ul for j = t mod E
  auto fetch = [&](int tl_) {
    const size_t b = (size_t)(tl_ / groups);
    const size_t c0 = (size_t)(tl_ - (int)b * groups) * C;
    const float2* __restrict__ src = a.in + b * a.n + c0;
#pragma unroll
    for (int k = 0; k < E; ++k) nx[k] = src[(size_t)(row0 + k * (NT / C)) * N2 + col];
    const size_t n2 = c0 + (size_t)line;
    const size_t m1 = n2 * (size_t)t, m2 = n2 * (size_t)(T * (t & (E - 1)));
    th = load_tw<DIR>(a.tw, (int)(m1 & ~mask));
    tl = load_tw<DIR>(a.tw, (int)(m1 & mask));
    uh = load_tw<DIR>(a.tw, (int)(m2 & ~mask));
    ul = load_tw<DIR>(a.tw, (int)(m2 & mask));
  };
  LineTw<L, DIR, 1, 0, Rs...>::fill(tab, a.tw, (int)(a.n / L), tid, NT);
  if (tile < hi) fetch(tile);
  wait_vm();  // the loop's first pass then starts, as every later one, with no loads pending
  for (; tile < hi; tile += step) {
    const int b = tile / groups;
    const size_t c0 = (size_t)(tile - b * groups) * C;
#pragma unroll
    for (int k = 0; k < E; ++k) lds[col][pidx(row0 + k * (NT / C))] = pre1(nx[k], a.hooks.pre_mode);
    const float2 wt = c_mul(th, tl);
    const float2 wu = c_mul(uh, ul);
    lds_barrier();  // the tile is in LDS (and, on the first pass, the twiddle tables)
    if (tile + step < hi) fetch(tile + step);
    float2 v[E];
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q * R0 + r] = buf[pidx(t + q * T + r * (L / R0))];
    wave_lds_fence();
    LineStockham<L, E, DIR, false, 1, 0, Rs...>::run(v, t, buf, tab, wave_lds_fence);
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < E / Rl; ++q)
#pragma unroll
      for (int r = 0; r < Rl; ++r) {
        const int j = q + r * (E / Rl);  // k1 = t + T*j
        buf[pidx(t + T * j)] = c_mul(v[q * Rl + r], c_mul(wt, readlane2(wu, j)));
      }
    wait_vm();
    lds_barrier();
    float2* __restrict__ dst = a.out + (size_t)b * a.n + c0;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int j = row0 + k * (NT / C);
      dst[(size_t)j * N2 + col] = lds[col][pidx(j)];
    }
    lds_barrier();  // every read of the tile is done before the next one is written
  }
}

// Pass B, pipelined: tile = C adjacent rows (length L, contiguous) of one transform; the
// output (row k1, k2) goes to k1 + N1*k2 through LDS as runs of C values, streaming stores.
template <int L, int E, int C, int DIR, int... Rs>
__global__ void __launch_bounds__(C*(L / E), 1) row_pipe_kernel(LargeArgs a) {
  constexpr int T = L / E;
  constexpr int NT = C * T;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int LP = staged_stride<L>();
  constexpr int TWN = line_tab_size<L, false, Rs...>();
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  static_assert(T == 64, "one wave per line");
  __shared__ float2 lds[C][LP];
  __shared__ float2 tab[TWN > 0 ? TWN : 1];

  const int tid = threadIdx.x;
  const int groups = a.n1 / C;
  int tile, hi, step;
  xcd_range(groups * (int)a.batch, &tile, &hi, &step);
  const size_t N1 = (size_t)a.n1;
  const int line = __builtin_amdgcn_readfirstlane(tid / T);
  const int t = tid - line * T;
  float2* buf = &lds[line][0];

  float2 nx[E];
  auto fetch = [&](int tl_) {
    const size_t b = (size_t)(tl_ / groups);
    const size_t r0 = (size_t)(tl_ - (int)b * groups) * C;
    const float2* __restrict__ row = a.in + b * a.n + (r0 + (size_t)line) * (size_t)L;
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) nx[q * R0 + r] = row[t + q * T + r * (L / R0)];
  };
  LineTw<L, DIR, 1, 0, Rs...>::fill(tab, a.tw, (int)(a.n / L), tid, NT);
  if (tile < hi) fetch(tile);
  wait_vm();
  lds_barrier();  // twiddle tables
  for (; tile < hi; tile += step) {
    const int b = tile / groups;
    const size_t r0 = (size_t)(tile - b * groups) * C;
    float2 v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) v[k] = nx[k];
    if (tile + step < hi) fetch(tile + step);
    LineStockham<L, E, DIR, false, 1, 0, Rs...>::run(v, t, buf, tab, wave_lds_fence);
    const size_t k1 = r0 + (size_t)line;
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < E / Rl; ++q)
#pragma unroll
      for (int r = 0; r < Rl; ++r) {
        const int k2 = t + q * T + r * (L / Rl);
        buf[pidx(k2)] = post_large(v[q * Rl + r], k1 + N1 * (size_t)k2, a.n, a.hooks);
      }
    wait_vm();
    lds_barrier();
    float2* __restrict__ dst = a.out + (size_t)b * a.n + r0;
    const int col = tid % C;
    const int row0 = tid / C;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int k2 = row0 + k * (NT / C);
      const float2 z = lds[col][pidx(k2)];
      float* o = reinterpret_cast<float*>(dst + (size_t)k2 * N1 + col);
      __builtin_nontemporal_store(z.x, o);
      __builtin_nontemporal_store(z.y, o + 1);
    }
    lds_barrier();  // every read of the tile is done before the next one's exchanges
  }
}

using LineFn = hipError_t (*)(const LargeArgs&, int lines, hipStream_t);

template <int L, int E, int C, int DIR, int... Rs>
hipError_t col_launch(const LargeArgs& a, int lines, hipStream_t s) {
  if (lines % C || (size_t)(lines / C) * a.batch > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((col_pass_kernel<L, E, C, DIR, Rs...>), dim3((unsigned)((lines / C) * a.batch)), dim3(C * (L / E)),
                     0, s, a);
  return hipGetLastError();
}
template <int L, int E, int C, int DIR, int... Rs>
hipError_t row_launch(const LargeArgs& a, int lines, hipStream_t s) {
  if (lines % C || (size_t)(lines / C) * a.batch > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((row_pass_kernel<L, E, C, DIR, Rs...>), dim3((unsigned)((lines / C) * a.batch)), dim3(C * (L / E)),
                     0, s, a);
  return hipGetLastError();
}

template <int L, int E, int C, bool HT, int DIR, int... Rs>
hipError_t col_line_launch(const LargeArgs& a, int lines, hipStream_t s) {
  if (lines % C || (size_t)(lines / C) * a.batch > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((col_line_kernel<L, E, C, HT, DIR, Rs...>), dim3((unsigned)((lines / C) * a.batch)),
                     dim3(C * (L / E)), 0, s, a);
  return hipGetLastError();
}
template <int L, int E, int C, bool HT, bool NTS, int DIR, int... Rs>
hipError_t row_line_launch(const LargeArgs& a, int lines, hipStream_t s) {
  if (lines % C || (size_t)(lines / C) * a.batch > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((row_line_kernel<L, E, C, HT, NTS, DIR, Rs...>), dim3((unsigned)((lines / C) * a.batch)),
                     dim3(C * (L / E)), 0, s, a);
  return hipGetLastError();
}

// Persistent launches: one workgroup per CU (at most one per tile).
template <int L, int E, int C, int DIR, int... Rs>
hipError_t col_pipe_launch(const LargeArgs& a, int lines, hipStream_t s) {
  if (lines % C || (size_t)(lines / C) * a.batch > 0x7fffffffull) return hipErrorInvalidValue;
  const size_t tiles = (size_t)(lines / C) * a.batch;
  const unsigned grid = (unsigned)std::min<size_t>(tiles, (size_t)device_cus());
  hipLaunchKernelGGL((col_pipe_kernel<L, E, C, DIR, Rs...>), dim3(grid), dim3(C * (L / E)), 0, s, a);
  return hipGetLastError();
}
template <int L, int E, int C, int DIR, int... Rs>
hipError_t row_pipe_launch(const LargeArgs& a, int lines, hipStream_t s) {
  if (lines % C || (size_t)(lines / C) * a.batch > 0x7fffffffull) return hipErrorInvalidValue;
  const size_t tiles = (size_t)(lines / C) * a.batch;
  const unsigned grid = (unsigned)std::min<size_t>(tiles, (size_t)device_cus());
  hipLaunchKernelGGL((row_pipe_kernel<L, E, C, DIR, Rs...>), dim3(grid), dim3(C * (L / E)), 0, s, a);
  return hipGetLastError();
}

struct LineEntry {
  int l, e, c;
  const char* radices;
  LineFn col_fwd, col_inv, row_fwd, row_inv;
};

// Per line length, the first entry is the default; the others are autotuner candidates.
#define FFTFREE_LINE(L, E, C, NAME, ...)                                                           \
  {L, E, C, NAME, &col_launch<L, E, C, -1, __VA_ARGS__>, &col_launch<L, E, C, 1, __VA_ARGS__>,    \
   &row_launch<L, E, C, -1, __VA_ARGS__>, &row_launch<L, E, C, 1, __VA_ARGS__>}
#define FFTFREE_LINE_T(L, E, C, HT, NTS, NAME, ...)                                                  \
  {L, E, C, NAME, &col_line_launch<L, E, C, HT, -1, __VA_ARGS__>, &col_line_launch<L, E, C, HT, 1, __VA_ARGS__>, \
   &row_line_launch<L, E, C, HT, NTS, -1, __VA_ARGS__>, &row_line_launch<L, E, C, HT, NTS, 1, __VA_ARGS__>}
#define FFTFREE_LINE_P(L, E, C, NAME, ...)                                                          \
  {L, E, C, NAME, &col_pipe_launch<L, E, C, -1, __VA_ARGS__>, &col_pipe_launch<L, E, C, 1, __VA_ARGS__>,    \
   &row_pipe_launch<L, E, C, -1, __VA_ARGS__>, &row_pipe_launch<L, E, C, 1, __VA_ARGS__>}
const LineEntry kLines[] = {
    FFTFREE_LINE(64, 8, 32, "8x8", 8, 8),
    FFTFREE_LINE(128, 16, 32, "16x8", 16, 8),
    FFTFREE_LINE(256, 16, 16, "16x16", 16, 16),
    FFTFREE_LINE(512, 8, 8, "8x8x8", 8, 8, 8),
    FFTFREE_LINE_T(1024, 16, 8, false, true, "16x16x4 lds-tw", 16, 16, 4),
    FFTFREE_LINE_T(1024, 16, 8, true, true, "16x16x4 lds-half-tw", 16, 16, 4),
    FFTFREE_LINE_T(1024, 16, 16, false, true, "16x16x4 lds-tw", 16, 16, 4),
    FFTFREE_LINE(1024, 16, 8, "16x16x4", 16, 16, 4),
    FFTFREE_LINE(1024, 8, 8, "8x8x4x4", 8, 8, 4, 4),
    FFTFREE_LINE(1024, 4, 4, "4x4x4x4x4", 4, 4, 4, 4, 4),
    // write-back row stores: WRITE_SIZE 1.00x the chunk instead of 1.30x (L2 merges the
    // half lines of neighbouring workgroups), but 315.7 against 302.8 us per 2^20 x 16
    // round trip
    FFTFREE_LINE_T(1024, 16, 8, false, false, "16x16x4 lds-tw wb", 16, 16, 4),
    FFTFREE_LINE_P(1024, 16, 8, "16x16x4 pipelined", 16, 16, 4),
    FFTFREE_LINE_T(2048, 16, 8, false, true, "16x16x8 lds-tw", 16, 16, 8),
    FFTFREE_LINE_T(2048, 16, 4, true, false, "16x16x8 lds-half-tw wb", 16, 16, 8),
    FFTFREE_LINE(2048, 16, 4, "16x16x8", 16, 16, 8),
    FFTFREE_LINE(2048, 8, 4, "8x8x8x4", 8, 8, 8, 4),
    FFTFREE_LINE(4096, 16, 4, "16x16x16", 16, 16, 16),
};
#undef FFTFREE_LINE
#undef FFTFREE_LINE_T
#undef FFTFREE_LINE_P

int line_count(int l) {
  int c = 0;
  for (const auto& e : kLines) c += e.l == l;
  return c;
}

const LineEntry* line_entry(int l, int idx = 0) {
  for (const auto& e : kLines)
    if (e.l == l && idx-- == 0) return &e;
  return nullptr;
}

}  // namespace

void large_split(size_t n, int* n1, int* n2, int* s) {
  int lg = 0;
  while ((size_t(1) << lg) < n) ++lg;
  const int lg1 = lg / 2, lg2 = lg - lg1;
  *n1 = 1 << lg1;
  *n2 = 1 << lg2;
  *s = (lg + 1) / 2;
}

bool large_supported(size_t n) {
  if (n <= (size_t)c2c_max_lds_n()) return false;
  int n1, n2, s;
  large_split(n, &n1, &n2, &s);
  return line_entry(n1) && line_entry(n2);
}

// Variant v picks column entry v % count(N1) and row entry v / count(N1).
int large_variant_count(size_t n) {
  if (!large_supported(n)) return 0;
  int n1, n2, s;
  large_split(n, &n1, &n2, &s);
  return line_count(n1) * line_count(n2);
}

bool large_variant_desc(size_t n, int v, char* buf, size_t len) {
  if (v < 0 || v >= large_variant_count(n)) return false;
  int n1, n2, s;
  large_split(n, &n1, &n2, &s);
  const int k1 = line_count(n1);
  const LineEntry* c = line_entry(n1, v % k1);
  const LineEntry* r = line_entry(n2, v / k1);
  std::snprintf(buf, len, "columns %d: %s e=%d lines=%d | rows %d: %s e=%d lines=%d", n1, c->radices, c->e, c->c, n2,
                r->radices, r->e, r->c);
  return true;
}

hipError_t launch_c2c_large(const C2CArgs& a, int inverse, int variant, float2* tmp, hipStream_t s) {
  if (a.batch == 0) return hipSuccess;
  LargeArgs la;
  la.batch = a.batch;
  la.n = (size_t)a.n;
  large_split(la.n, &la.n1, &la.n2, &la.s);
  la.tw = a.tw;
  la.hooks = a;
  const int k1 = line_count(la.n1);
  if (variant >= k1 * line_count(la.n2)) return hipErrorInvalidValue;
  const LineEntry* e1 = line_entry(la.n1, variant < 0 ? 0 : variant % k1);
  const LineEntry* e2 = line_entry(la.n2, variant < 0 ? 0 : variant / k1);
  if (!e1 || !e2) return hipErrorInvalidValue;
  la.in = a.in;
  if (a.pre_mode == 2) {  // rebuild into out, then run the plain transform from there
    const size_t h = la.n / 2;
    hipLaunchKernelGGL(hermitian_rebuild_kernel, dim3((unsigned)((h + 256) / 256), (unsigned)a.batch), dim3(256), 0, s,
                       a.in, a.out, la.n, a.pre_polar);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    la.in = a.out;
    la.hooks.pre_mode = 0;
  }
  // pass A: in -> tmp (columns, length N1, N2 of them)
  la.out = tmp;
  hipError_t err = (inverse ? e1->col_inv : e1->col_fwd)(la, la.n2, s);
  if (err != hipSuccess) return err;
  // pass B: tmp -> out (rows, length N2, N1 of them)
  la.in = tmp;
  la.out = a.out;
  la.hooks.pre_mode = 0;
  return (inverse ? e2->row_inv : e2->row_fwd)(la, la.n1, s);
}

}  // namespace fftfree

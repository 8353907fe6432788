// fft_stft_impl.hpp -- the real-input STFT kernels and their launch templates, shared by
// the translation units that instantiate them (fft_stft.hip: the variant table, the
// generic and small-N kernels; fft_stft_lines*.hip: the line-aligned kernels;
// fft_stft_mf.hip: the multi-frame kernels; fft_ragged.hip: the ragged batch). Kernels and
// launchers have internal linkage: each translation unit owns the instantiations it names.
//
// stft_r2c_lines_kernel  the default at N = 1024, 2048, 4096: fused framing + analysis window +
//                        zero-pad (fft_cffi.cpp:766-791), the real-input DFT of size N
//                        as a complex DFT of size N/2 plus a split step, the
//                        post-transform hook (eigen_fft.hpp:2535-2596) and the
//                        three-plane output copy (fft_cffi.cpp:811-825) in one launch,
//                        with every plane written as whole 128-byte lines.
// stft_r2c_mf_kernel     the same fused path with per-iteration plane runs (other N,
//                        full spectra).
// stft_r2c_kernel        the general fused path: any window <= N, hop and alignment.
//
// Data layout in HBM: PCM is one float array; spectra planes are frame-major
// (frame f, bin k at f*bins + k). One workgroup owns whole frames: nothing crosses
// workgroups, so no atomics.
#pragma once

#include "fft_device.hpp"
#include "fft_kernels.hpp"
#include "fft_stft_tile.hpp"

#include <cmath>
#include <cstring>

namespace fftfree {

using StftFn = hipError_t (*)(const StftArgs&, hipStream_t);

// One row of the real-input variant table: N, its place among the variants of N (the
// first is the default family), the kernel shape, forward and inverse launchers, and for
// multi-frame variants their resident workgroups on the current device.
struct StftEntry {
  int n;
  int order;
  VariantInfo info;
  StftFn fwd, inv;
  size_t (*slots)();
};

// Compute units of the current device, cached per device.
int device_cus();
// Workgroups of kernel k that the current device holds at once (CUs x occupancy at 256
// threads), cached per device.
size_t resident_slots(const void* k);
// Dynamic LDS bytes that cap kernel k at `wpc` workgroups per CU (0: no cap).
size_t lds_cap_bytes(const void* k, int wpc);

// The table rows each kernel translation unit contributes.
const StftEntry* stft_lines_entries(int* count);
const StftEntry* stft_lines_exch_entries(int* count);
const StftEntry* stft_lines_order_entries(int* count);
const StftEntry* stft_lines_2048_entries(int* count);
const StftEntry* stft_lines_4096_entries(int* count);
const StftEntry* stft_mf_entries(int* count);

using namespace dev;

namespace {


// Second radix of a list.
template <int R0, int R1, int... Rs>
struct SecondOf {
  static constexpr int value = R1;
};

// Smallest power of two >= x (LDS ring sizes).
constexpr int ring_floats(int x) {
  int r = 1;
  while (r < x) r *= 2;
  return r;
}
// Line-aligned store ranges of the lines kernel: iteration it of a K-iteration tile
// (RUN floats per plane per iteration) stores the tile region's floats [lo, hi); hi is
// the last 32-float line end its frames complete, the tile's end for the last iteration.
__host__ __device__ constexpr int lines_hi(int run, int k, int it) {
  return it == k - 1 ? k * run : ((it + 1) * run / 32) * 32;
}
template <int RUN, int K>
__host__ __device__ constexpr int lines_hi(int it) {
  return lines_hi(RUN, K, it);
}
template <int RUN, int K>
__host__ __device__ constexpr int lines_lo(int it) {
  return it == 0 ? 0 : lines_hi(RUN, K, it - 1);
}
template <int RUN, int K, int NT>
__host__ __device__ constexpr int lines_rounds(int it) {
  return ((lines_hi<RUN, K>(it) - lines_lo<RUN, K>(it)) / 4 + NT - 1) / NT;
}
template <int RUN, int K, int NT>
__host__ __device__ constexpr int lines_max_rounds() {
  int m = 0;
  for (int it = 0; it < K; ++it) m = lines_rounds<RUN, K, NT>(it) > m ? lines_rounds<RUN, K, NT>(it) : m;
  return m;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// Buffer descriptor of the `bytes` bytes of a plane from float `first` on. A skipped
// plane (nullptr) gets a zero-size descriptor: the hardware's range check drops every
// store to it.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(float* p, size_t first, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(p ? static_cast<void*>(p + first) : nullptr, 0, p ? bytes : 0,
                                           0x00020000);
}
// Byte offset past every descriptor: a lane given it stores nothing, so a partial round
// is the same instruction stream as a full one.
constexpr int kDropOffset = (int)0x80000000u;
// Streaming (non-temporal) 16-byte store at byte offset `off` of r.
__device__ __forceinline__ void store_nt16(f32x4 v, __amdgpu_buffer_rsrc_t r, int off) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, off, 0, 2);
}

template <int L, int E, int FPW, int IP, int MODE, int HALF, int DIR, int... Rs>
__global__ void __launch_bounds__(FPW*(L / E)) stft_r2c_kernel(StftArgs a) {
  using Tile = StftTile<L, E, FPW, IP, MODE, HALF, DIR, Rs...>;
  __shared__ float2 lds[Tile::kLds];
  __shared__ float stage_extra[Tile::kExtra ? Tile::kExtra : 1];
  Tile::run(a, (size_t)xcd_tile(blockIdx.x, gridDim.x), lds, stage_extra);
}

// Multi-frame fused real-input STFT for transforms that fit one wave (T = L/E <= 64).
// A workgroup of NT = 256 threads runs P = NT/T transform slots; slot s transforms the K
// consecutive frames f0 + s*K + i, i < K, one after another. Per thread:
//  * the window taps of its elements, every pass twiddle and the split twiddles are
//    loaded once into registers and reused for all K frames;
//  * the P frames of one iteration are consecutive; the samples they span are loaded
//    once with 16-byte loads (one iteration ahead) and staged in LDS;
//  * at iteration i the P slots hold P consecutive frames; their bins are staged in LDS
//    and written as one contiguous run per plane with 16-byte stores, the same number
//    of stores in every wave and iteration, so a wait for prefetched samples never
//    covers more than the stores of the last few iterations (CDNA's vector-memory
//    counter covers loads and stores, in issue order, and saturates at 63);
//  * those stores are issued during the next iteration, between its FFT passes, so each
//    wave keeps a steady store stream instead of computing and storing in turns.
// Whole tiles of K*P regular frames (every frame inside the PCM) run as above; they
// need the window to span N, hop a multiple of 4 and at most N/4, and 16-byte aligned
// samples and planes. The frames after the last whole tile (a few dozen) go to
// workgroup 0, which is dispatched first: it walks them P at a time with per-frame
// bounds, zero fill and direct plane stores, while the tiles run.
template <int L, int E, int K, int PL, int MODE, int HALF, int DIR, int... Rs>
__global__ void __launch_bounds__(256, 3) stft_r2c_mf_kernel(StftArgs a) {
  constexpr int T = L / E;
  constexpr int N = 2 * L;
  constexpr int NT = 256;
  constexpr int P = NT / T;            // transform slots per workgroup
  constexpr int NP = sizeof...(Rs);
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int LP = padded_len<L>();
  constexpr int JF = (L / 2) / T;      // split steps covering p in [0, L/2)
  constexpr int BINS = HALF ? L + 1 : N;
  constexpr bool kB = mode_has_b<MODE>();
  constexpr int ROW = L + 1;           // staged bins per frame
  constexpr int STAGE = (kB ? 2 : 1) * P * ROW;
  constexpr int RUN = P * BINS;        // floats per plane per iteration
  constexpr int N4 = RUN / 4;
  constexpr int C4 = (N4 + NT - 1) / NT;
  // PL = 0: each iteration's samples (P frames) are loaded a few iterations ahead into a
  // double buffer; PL = 1: the samples of the whole tile (P*K frames) are loaded once, at
  // the start, so no load (and no wait for one) follows the first store.
  constexpr int SPAN = (PL ? P * K - 1 : P - 1) * (N / 4) + N;  // hop <= N/4
  constexpr int SPAN4 = SPAN / 4;
  constexpr int SR = (SPAN4 + NT - 1) / NT;    // 16-byte loads per thread per load phase
  constexpr int NSPAN = PL ? 1 : 2;
  constexpr int NSTAGE = PL ? 1 : 2;
  static_assert(T <= 64 && (64 % T) == 0, "one transform per wave or less");
  static_assert(JF * T == L / 2, "T must divide L/2");
  static_assert(RUN % 4 == 0 && SPAN % 4 == 0, "runs must be whole 16-byte words");
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  __shared__ float2 lds[P][LP];
  __shared__ __attribute__((aligned(16))) float stage[NSTAGE][STAGE];
  __shared__ __attribute__((aligned(16))) float span[NSPAN][SPAN];

  const int tid = threadIdx.x;
  const int slot = tid / T;
  const int t = tid & (T - 1);
  float2* buf = &lds[slot][0];
  const int has_rest = a.frames > a.lead ? 1 : 0;
  const size_t tile = (size_t)xcd_tile((int)blockIdx.x - has_rest, (int)gridDim.x - has_rest);
  size_t f0 = tile * (size_t)(K * P);  // first frame of the tile in hand
  const int hop = a.hop;
  const float* __restrict__ win = a.win;
  const float2* __restrict__ w2 = reinterpret_cast<const float2*>(win);
  constexpr float scale = DIR > 0 ? 1.0f / (float)N : 1.0f;

  // per-thread constants: window taps, pass twiddles, split twiddles
  float2 wr[E];
  const float2* __restrict__ wsrc = win ? w2 : a.tw;
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const float2 w = wsrc[t + q * T + r * (L / R0)];  // unconditional load, then a select
      wr[q * R0 + r] = win ? w : make_float2(1.0f, 1.0f);
    }
  float2 twr[E * NP];
  StockhamPlan<L, E, DIR, Rs...>::load_twiddles(twr, t, a.tw, 2);
  float2 wsp[JF];
#pragma unroll
  for (int j = 0; j < JF; ++j) wsp[j] = load_tw<DIR>(a.tw, t + j * T);
  const float2 wmid = load_tw<DIR>(a.tw, L / 2);

  // frames after the last whole tile: P at a time with per-frame bounds, zero fill and
  // direct plane stores
  auto do_rest = [&]() {
    for (size_t fb = a.lead; fb < a.frames; fb += P) {
      const size_t f = fb + slot;
      const bool active = f < a.frames;
      const size_t start = (f + a.frame_offset) * (size_t)hop;
      size_t cnt = 0;
      if (active && start < a.pcm_len) {
        const size_t rem = a.pcm_len - start;
        cnt = rem < (size_t)a.window ? rem : (size_t)a.window;
      }
      const float* __restrict__ src = a.pcm + (cnt ? start : 0);
      float2 v[E];
#pragma unroll
      for (int q = 0; q < E / R0; ++q)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const int n = t + q * T + r * (L / R0);
          v[q * R0 + r] = make_float2(frame_sample(src, win, cnt, 2 * n), frame_sample(src, win, cnt, 2 * n + 1));
        }
      WaveStockham<L, E, DIR, 1, 0, Rs...>::run(v, t, buf, twr, [](int) {});
      store_natural<L, E, Rl>(v, t, buf);
      wave_lds_fence();
      if (active) {
        float* __restrict__ ore = a.o_re ? a.o_re + f * (size_t)BINS : nullptr;
        float* __restrict__ oim = a.o_im ? a.o_im + f * (size_t)BINS : nullptr;
        float* __restrict__ omg = a.o_mag ? a.o_mag + f * (size_t)BINS : nullptr;
        auto put1 = [&](int k, float2 X) {
          float A, B;
          stage_vals<MODE>(X, A, B);
          if (ore) ore[k] = plane_value<MODE, 0>(A, B);
          if (oim) oim[k] = plane_value<MODE, 1>(A, B);
          if (omg) omg[k] = plane_value<MODE, 2>(A, B);
        };
        auto put = [&](int k, float2 X) {
          if (DIR > 0) X = c_scale(X, scale);
          put1(k, X);
          if constexpr (!HALF) {
            if (k > 0 && k < L) put1(N - k, c_conj(X));
          }
        };
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
          if (p == 0) {
            X1.y = 0.0f;
            X2.y = 0.0f;
          }
          put(p, X1);
          put(L - p, X2);
        }
        const float2 zm = buf[pidx(L / 2)];
        const float2 zc = c_conj(zm);
        const float2 e = c_scale(c_add(zm, zc), 0.5f);
        const float2 d = c_sub(zm, zc);
        put(L / 2, c_add(e, c_mul(wmid, make_float2(0.5f * d.y, -0.5f * d.x))));
      }
      wave_lds_fence();
    }
  };

  // iteration it covers frames f0 + it*P .. +P-1, i.e. samples [s, s + (P-1)*hop + N)
  // of the PCM (the host guarantees 16-byte alignment and that they lie inside it);
  // SR 16-byte loads per thread, spare lanes re-read the last word
  const int span4 = ((PL ? P * K - 1 : P - 1) * hop + N) / 4;
  auto load_span = [&](int it, f32x4* dst) {
    const f32x4* __restrict__ src =
        reinterpret_cast<const f32x4*>(a.pcm + (f0 + (size_t)it * P + a.frame_offset) * (size_t)hop);
#pragma unroll
    for (int c = 0; c < SR; ++c) {
      const int idx = tid + c * NT;
      dst[c] = src[idx < span4 ? idx : span4 - 1];
    }
  };
  auto stash_span = [&](int it, const f32x4* val) {
#pragma unroll
    for (int c = 0; c < SR; ++c) {
      const int idx = tid + c * NT;
      if (idx < span4) reinterpret_cast<f32x4*>(&span[it & (NSPAN - 1)][0])[idx] = val[c];
    }
  };

  // Spans are requested PF iterations ahead of use. A wait for a load also waits for
  // every store issued before it, so the distance lets each wave keep PF iterations of
  // stores in flight instead of draining them at every iteration.
  constexpr int PF = PL ? 1 : (K - 1 < 3 ? (K - 1 > 0 ? K - 1 : 1) : 3);
  f32x4 nxt[PF][SR];

  // Store round c (c < C4) of iteration it's staged run: the P frames of an iteration
  // are one contiguous, 16-byte-aligned run per plane; spare lanes of the last round
  // rewrite the run's final word, so every wave issues the same stores.
  auto store_round = [&](int it, int c) {
    const float* sA = &stage[it & (NSTAGE - 1)][0];
    const float* sB = sA + P * ROW;
    const size_t base = (f0 + (size_t)it * P) * (size_t)BINS;
    int idx = tid + c * NT;
    idx = idx < N4 ? idx : N4 - 1;
    f32x4 A4, B4 = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (HALF) {
      A4 = *reinterpret_cast<const f32x4*>(sA + 4 * idx);
      if constexpr (kB) B4 = *reinterpret_cast<const f32x4*>(sB + 4 * idx);
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int g = 4 * idx + u;
        const int fr = g / BINS;
        const int k = g - fr * BINS;
        const int jj = fr * ROW + (k <= L ? k : N - k);
        A4[u] = sA[jj];
        if constexpr (kB) B4[u] = k <= L ? sB[jj] : -sB[jj];
      }
    }
    f32x4 o0, o1, o2;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      o0[u] = plane_value<MODE, 0>(A4[u], B4[u]);
      o1[u] = plane_value<MODE, 1>(A4[u], B4[u]);
      o2[u] = plane_value<MODE, 2>(A4[u], B4[u]);
    }
    // a skipped plane lands in the sink (N floats, never read)
    const int sidx = idx & (N / 4 - 1);
    // planes are written once and not read back here: streaming (non-temporal) stores
    __builtin_nontemporal_store(o0, reinterpret_cast<f32x4*>(a.o_re ? a.o_re + base : a.sink) + (a.o_re ? idx : sidx));
    __builtin_nontemporal_store(o1, reinterpret_cast<f32x4*>(a.o_im ? a.o_im + base : a.sink) + (a.o_im ? idx : sidx));
    __builtin_nontemporal_store(o2, reinterpret_cast<f32x4*>(a.o_mag ? a.o_mag + base : a.sink) + (a.o_mag ? idx : sidx));
  };
  // rounds of the previous iteration spread over the NP passes of the current one
  constexpr int PER_PASS = (C4 + NP - 1) / NP;

  // The K iterations of the tile at f0; its samples are already in the span buffer when
  // PL == 2 (the persistent loop stashes them).
  auto run_tile = [&]() {
  if constexpr (PL != 2) {
    load_span(0, nxt[0]);
    stash_span(0, nxt[0]);
    __syncthreads();
  }
  if constexpr (!PL) {
#pragma unroll
    for (int j = 1; j <= PF && j < K; ++j) load_span(j, nxt[j % PF]);
  }
#pragma unroll
  for (int i = 0; i < K; ++i) {
    // this slot's frame from the staged span, times the window, first-pass order
    float2 v[E];
    const float2* fs = reinterpret_cast<const float2*>(PL ? &span[0][(i * P + slot) * hop] : &span[i & 1][slot * hop]);
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int e = q * R0 + r;
        const float2 x = fs[t + q * T + r * (L / R0)];
        v[e] = make_float2(x.x * wr[e].x, x.y * wr[e].y);
      }

    WaveStockham<L, E, DIR, 1, 0, Rs...>::run(v, t, buf, twr, [&](int pass) {
      if (i > 0) {
#pragma unroll
        for (int c = pass * PER_PASS; c < (pass + 1) * PER_PASS && c < C4; ++c) store_round(i - 1, c);
      }
    });
    store_natural<L, E, Rl>(v, t, buf);
    if constexpr (PL) {
      __syncthreads();  // the one staging buffer: every store round of i-1 has read it
    } else {
      wave_lds_fence();
    }

    // split into this iteration's staging buffer (double-buffered across iterations)
    float* rowA = &stage[i & (NSTAGE - 1)][0] + slot * ROW;
    float* rowB = &stage[i & (NSTAGE - 1)][0] + P * ROW + slot * ROW;
    auto put = [&](int k, float2 X) {
      if (DIR > 0) X = c_scale(X, scale);
      float A, B;
      stage_vals<MODE>(X, A, B);
      rowA[k] = A;
      if constexpr (kB) rowB[k] = B;
    };
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
      put(p, X1);
      put(L - p, X2);
    }
    {  // bin L/2: every lane computes and writes the same value
      const float2 zm = buf[pidx(L / 2)];
      const float2 zc = c_conj(zm);
      const float2 e = c_scale(c_add(zm, zc), 0.5f);
      const float2 d = c_sub(zm, zc);
      put(L / 2, c_add(e, c_mul(wmid, make_float2(0.5f * d.y, -0.5f * d.x))));
    }
    wave_lds_fence();
    // next iteration's samples: stash what was prefetched, request the one PF later
    if constexpr (!PL) {
      if (i + 1 < K) {
        stash_span(i + 1, nxt[(i + 1) % PF]);
        if (i + 1 + PF < K) load_span(i + 1 + PF, nxt[(i + 1 + PF) % PF]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < C4; ++c) store_round(K - 1, c);
  };

  if constexpr (PL == 2) {
    // Persistent: workgroup of rank r (consecutive ranks share an XCD) runs tiles
    // [r*T/G, (r+1)*T/G) and requests each tile's samples while it runs the one before,
    // so the wait for them covers only stores a whole tile old. The rank the host picks
    // (a tile short) also runs the frames after the last tile.
    const size_t G = gridDim.x;
    const size_t T = a.lead / (size_t)(K * P);
    const size_t r = (size_t)xcd_tile(blockIdx.x, gridDim.x);
    const size_t tb = r * T / G, te = (r + 1) * T / G;
    f32x4 ahead[SR];
    if (tb < te) {
      f0 = tb * (size_t)(K * P);
      load_span(0, ahead);
      stash_span(0, ahead);
    }
    __syncthreads();
    for (size_t j = tb; j < te; ++j) {
      // unconditional request and stash (the last tile re-requests itself), so that no
      // load is still pending on the loop's back edge
      f0 = (j + 1 < te ? j + 1 : j) * (size_t)(K * P);
      load_span(0, ahead);
      f0 = j * (size_t)(K * P);
      run_tile();
      __syncthreads();  // every frame of tile j has been read from the span
      stash_span(0, ahead);
      __syncthreads();
    }
    if (r == (size_t)a.rest_rank && a.frames > a.lead) do_rest();
  } else {
    if (has_rest && blockIdx.x == 0) {
      do_rest();
      return;
    }
    run_tile();
  }
}

// Line-aligned multi-frame real-input STFT for transforms that fit one wave (T = L/E <=
// 64), half spectrum. The default at N = 1024.
//
// A frame-major plane row of `bins` floats is rarely a whole number of 128-byte lines
// (513 floats are 16 lines and 4 bytes), and a store that covers part of a line costs
// HBM time close to a whole line's: three planes written as 4-frame runs that start and
// end inside lines move at about 3.9 TB/s, the same bytes written as whole lines at
// 5.3-6.0 TB/s. So a workgroup owns tiles of K*P consecutive frames whose plane region
// starts on a line (at 513 bins, 32 frames are exactly 513 lines) and writes every
// plane as whole lines:
//  * P = 256/T slots transform P consecutive frames per iteration, one frame per slot;
//  * each slot loads its frame's samples from global memory straight into registers,
//    one iteration ahead (neighbouring frames share 3/4 of their samples; the repeats
//    hit L1/L2), times window taps pre-scaled by 1/2 (exact), so that the split needs
//    no halving;
//  * the window taps, the split twiddles and the second pass's twiddles live in LDS,
//    not registers, so that four workgroups fit per CU without spills;
//  * the split writes the bins into an LDS ring indexed by the float's offset inside
//    the tile's plane region (mod RING);
//  * iteration i stores the region's floats [lo_i, hi_i): hi_i is the last line end its
//    frames complete, so the few floats past it wait in the ring for iteration i+1, and
//    the last iteration flushes to the tile's end;
//  * the stores are buffer stores through one descriptor per plane (a skipped plane's
//    descriptor is empty), issued between the FFT passes of the next iteration, and a
//    lane past a round's end gets an offset the range check drops, so every wave issues
//    the same straight-line stores and each wait for the next frame's samples covers
//    only stores issued before their loads (CDNA's vector-memory counter counts loads
//    and stores in issue order).
// OPT bit 0 deals tiles in launch order (consecutive tiles on different XCDs) instead
// of contiguous runs per XCD; bit 1 makes pass p of an iteration write plane p's
// rounds, so each wave has one write stream at a time. Both together measured fastest
// at N = 1024 (65533 frames: 87-88 us against 92 us for neither). Bit 2 keeps the window
// taps, the split twiddles and the second pass's twiddles in registers instead of LDS:
// about a quarter less LDS traffic per frame for more registers (at most 3 workgroups
// per CU); it measured no faster. The default runs 4 workgroups per CU (87 us; 3 per
// CU: 88.5 us, 2 per CU: 95 us), tiles of 32 frames (64-frame tiles: 92 us), dealt in
// launch order in groups of 8 consecutive tiles per XCD with the even tiles walked last
// frames first (bits 5..7): a tile shares 768 samples with each neighbour, and walked so,
// both read them at about the same time through one L2, which takes the samples' reads
// from 1.10x to 1.02x at the same time per launch.
// Whole tiles of regular frames (every frame inside the PCM) run as above; they need the
// window to span N, an even hop, 8-byte aligned samples and 16-byte aligned window and
// planes. The frames after the last whole tile (a few dozen) go to workgroup 0,
// dispatched first, which walks them P at a time with per-frame bounds, zero fill and
// direct stores.
// The kernel body is a device routine over a virtual block index in [0, vgrid) (tiles in
// order, block 0 the rest when there is one), so that persistent kernels (ragged batches)
// can run it for work items of many requests.
template <int L, int E, int K, int OPT, int MODE, int DIR, int... Rs>
__device__ __forceinline__ void stft_lines_body(const StftArgs& a, int vblock, int vgrid) {
  constexpr int T = L / E;
  constexpr int N = 2 * L;
  constexpr int NT = 256;
  constexpr int P = NT / T;            // frames per iteration
  constexpr int NP = sizeof...(Rs);
  constexpr int R0 = FirstOf<Rs...>::value;
  constexpr int Rl = LastOf<Rs...>::value;
  constexpr int LP = padded_len<L>();
  constexpr int JF = (L / 2) / T;      // split steps covering p in [0, L/2)
  constexpr int BINS = L + 1;
  constexpr bool kB = mode_has_b<MODE>();
  constexpr int RUN = P * BINS;        // floats per plane per iteration
  // LDS ring of the split's output: a power of two (index masks) unless that would double
  // a large ring (N = 2048: 4100 + 32 floats per plane), which is then sized to fit and
  // indexed with one conditional wrap
  constexpr bool kPow2Ring = ring_floats(RUN + 32) < 2 * 4096;
  constexpr int RING = kPow2Ring ? ring_floats(RUN + 32) : ((RUN + 32 + 3) / 4) * 4;
  constexpr bool kRegTabs = (OPT & 4) != 0;
  // bit 8 (with bit 2): the window taps stay in LDS, only the twiddles go to registers;
  // bit 9 (without bit 2): the window taps go to registers, the twiddles stay in LDS
  constexpr bool kRegWin = (kRegTabs && (OPT & 256) == 0) || (!kRegTabs && (OPT & 512) != 0);
  // OPT bit 3: the exchanges between passes in registers (DPP / permlane swaps); with bit
  // 4 only the first one (lane bits 3..5), the second through LDS
  constexpr bool kRegX = (OPT & 8) != 0;
  constexpr bool kLoLds = (OPT & 16) != 0;
  static_assert((T <= 64 && (64 % T) == 0) || T == 128, "a transform on one wave or less, or on two");
  static_assert(JF * T == L / 2, "T must divide L/2");
  // store ranges end on line ends, except the tile's last, which ends at K * RUN
  static_assert((K * RUN) % 4 == 0, "tiles must be whole 16-byte words");
  static_assert(product<Rs...>() == L, "radices must multiply to L");
  static_assert(!kRegX || (L == 512 && E == 8 && NP == 3 && R0 == 8 && Rl == 8 && SecondOf<Rs...>::value == 8),
                "register exchanges: the 8x8x8 transform of one wave");
  static_assert(!kLoLds || (kRegX && !kRegTabs), "bit 4 refines bit 3; pass 1's twiddles come from LDS");
  __shared__ float2 lds[P][LP];
  __shared__ __attribute__((aligned(16))) float ring[kB ? 2 : 1][RING];
  // window taps x 1/2, as (x[2n], x[2n+1]) pairs
  __shared__ __attribute__((aligned(16))) float2 wtab[kRegWin ? 2 : L];
  __shared__ float2 stw[kRegTabs ? 1 : L / 2 + 1];  // split twiddles, p = 0 .. L/2
  // second-pass twiddles: they depend on t mod R0 only, so (LDS tables) they are reloaded
  // before that pass of every frame instead of living in registers across the kernel
  constexpr int R1 = SecondOf<Rs...>::value;
  constexpr int NS1 = R0;  // the second pass's input span
  __shared__ float2 tw1[kRegTabs ? 1 : NS1][R1 - 1];

  const int tid = threadIdx.x;
  // with a transform on one wave or more the slot is wave-uniform: a scalar, so that every
  // slot-derived LDS and global address is scalar base + lane offset
  const int slot = T >= 64 ? __builtin_amdgcn_readfirstlane(tid / T) : tid / T;
  const int t = tid & (T - 1);
  // the butterfly index of pass 1 (tb1) and pass 2 (tb2), which set their twiddles: t, or
  // with register exchanges the lane's digits rotated
  const int tb1 = kRegX ? rot_lane(t) : t;
  const int tb2 = kRegX && !kLoLds ? rot_lane(t) : t;
  float2* buf = &lds[slot][0];
  const int has_rest = a.frames > a.lead ? 1 : 0;
  // workgroup 0 takes the rest (when there is one); the others walk the tiles.
  // OPT bit 0: tiles in launch order (consecutive tiles on consecutive workgroups, so on
  // different XCDs) instead of contiguous runs per XCD; bits 5..6 = log2 G > 0: launch
  // order in groups of G consecutive tiles per XCD, so that the samples two neighbouring
  // tiles share are fetched once, through one L2
  constexpr int kGroup = 1 << ((OPT >> 5) & 3);
  const long tile = (long)((OPT & 1) ? xcd_group_tile<kGroup>(vblock, vgrid) : xcd_tile(vblock, vgrid)) - has_rest;
  const int hop = a.hop;
  const float* __restrict__ win = a.win;
  constexpr float scale = DIR > 0 ? 1.0f / (float)N : 1.0f;

  // per-thread constants: pass twiddles (registers)
  float2 twr[E * NP];
  StockhamPlan<L, E, DIR, Rs...>::load_twiddles(twr, tb2, a.tw, 2);

  // window taps x 1/2 and split twiddles in registers (bit 2), else in LDS tables
  float2 wreg[kRegWin ? E : 1];
  float2 wsp[kRegTabs ? JF : 1];
  float2 wmid = make_float2(0.0f, 0.0f);
  if constexpr (kRegWin) {
    const float2* __restrict__ w2 = reinterpret_cast<const float2*>(win ? win : reinterpret_cast<const float*>(a.tw));
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const float2 x = w2[t + q * T + r * (L / R0)];
        wreg[q * R0 + r] = win ? make_float2(0.5f * x.x, 0.5f * x.y) : make_float2(0.5f, 0.5f);
      }
  }
  if constexpr (kRegTabs) {
#pragma unroll
    for (int j = 0; j < JF; ++j) wsp[j] = load_tw<DIR>(a.tw, t + j * T);
    wmid = load_tw<DIR>(a.tw, L / 2);
  } else {
    // split and second-pass twiddles into LDS (read per frame there, so that four
    // workgroups fit per CU)
    for (int p = tid; p <= L / 2; p += NT) stw[p] = load_tw<DIR>(a.tw, p);
    if (tid < NS1 * (R1 - 1)) {
      const int m = tid / (R1 - 1), r = tid % (R1 - 1) + 1;
      tw1[m][r - 1] = load_tw<DIR>(a.tw, m * (L / R1 / NS1) * 2 * r);
    }
  }
  if constexpr (!kRegWin) {
    // window taps times 1/2 into LDS (1/2 when windows are off), 16-byte loads; held
    // there rather than in registers so that four workgroups fit per CU
    static_assert(L % 2 == 0, "whole 16-byte words of taps");
#pragma unroll
    for (int w = tid; w < L / 2; w += NT) {  // 16-byte word: taps 4w .. 4w+3
      // unconditional load (the twiddle table stands in for a missing window), then a
      // select: a conditional load would become a branch and a wait
      const f32x4 x = reinterpret_cast<const f32x4*>(win ? win : reinterpret_cast<const float*>(a.tw))[w];
      const f32x4 h = {0.5f, 0.5f, 0.5f, 0.5f};
      reinterpret_cast<f32x4*>(&wtab[0])[w] = win ? x * h : h;
    }
  }
  __syncthreads();

  // One frame from its raw samples `cur` (first-pass order): window, the passes (the
  // second pass's twiddles refreshed from LDS in pass 0, then during(pass)), the
  // natural-order store into buf, a workgroup barrier, and the split, which hands bin k
  // to put(k, X). after_window() runs once the samples are consumed (the next frame's
  // loads go there). The tile frames and the rest frames both run this one body, so a
  // frame's bins do not depend on which path (or which launch size) produced them.
  auto spectrum = [&](float2* cur, auto&& after_window, auto&& during, auto&& put) {
    float2 v[E];
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int e = q * R0 + r;
        float2 w;
        if constexpr (kRegWin)
          w = wreg[e];
        else
          w = wtab[t + q * T + r * (L / R0)];
        v[e] = make_float2(cur[e].x * w.x, cur[e].y * w.y);
      }
    after_window();
    auto hook = [&](int pass) {
      if (!kRegTabs && pass == 0) {  // the second pass's twiddles (slots E .. of twr), fresh from LDS
#pragma unroll
        for (int q = 0; q < E / R1; ++q)
#pragma unroll
          for (int r = 1; r < R1; ++r) twr[E + q * (R1 - 1) + r - 1] = tw1[(tb1 + q * T) & (NS1 - 1)][r - 1];
      }
      during(pass);
    };
    if constexpr (kRegX) {
      WaveRegs888<DIR, kLoLds>::run(v, t, buf, twr, hook);
      if constexpr (kLoLds)
        store_natural<L, E, Rl>(v, t, buf);
      else
        store_natural_rot(v, t, buf);
    } else {
      WaveStockham<L, E, DIR, 1, 0, Rs...>::run(v, t, buf, twr, hook);
      store_natural<L, E, Rl>(v, t, buf);
    }
    __syncthreads();  // tiles: every wave has read the ring for the previous iteration's stores
    auto emit = [&](int k, float2 X) {
      if (DIR > 0) X = c_scale(X, scale);
      put(k, X);
    };
#pragma unroll
    for (int j = 0; j < JF; ++j) {
      const int p = t + j * T;
      const float2 zp = buf[pidx(p)];
      const float2 zc = c_conj(buf[pidx((L - p) & (L - 1))]);
      const float2 e = c_add(zp, zc);
      const float2 d = c_sub(zp, zc);
      float2 wp;
      if constexpr (kRegTabs)
        wp = wsp[j];
      else
        wp = stw[p];
      const float2 wo = c_mul(wp, make_float2(d.y, -d.x));
      float2 X1 = c_add(e, wo);
      float2 X2 = c_conj(c_sub(e, wo));
      if (p == 0) {  // bins 0 and L are real
        X1.y = 0.0f;
        X2.y = 0.0f;
      }
      emit(p, X1);
      emit(L - p, X2);
    }
    {  // bin L/2, emitted by one lane
      const float2 zm = buf[pidx(L / 2)];
      const float2 zc = c_conj(zm);
      const float2 e = c_add(zm, zc);
      const float2 d = c_sub(zm, zc);
      float2 wm;
      if constexpr (kRegTabs)
        wm = wmid;
      else
        wm = stw[L / 2];
      const float2 X = c_add(e, c_mul(wm, make_float2(d.y, -d.x)));
      if (t == 0) emit(L / 2, X);
    }
    __syncthreads();  // tiles: the ring holds this iteration; buf is free for the next frame
  };

  // Frames [a.lead, a.frames), P at a time (the loop is uniform over the workgroup):
  // per-frame bounds, zero fill past the PCM, direct plane stores.
  auto rest_frames = [&]() {
    for (size_t fb = a.lead; fb < a.frames; fb += P) {
      const size_t f = fb + slot;
      const bool active = f < a.frames;
      const size_t start = (f + a.frame_offset) * (size_t)hop;
      size_t cnt = 0;
      if (active && start < a.pcm_len) {
        const size_t rem = a.pcm_len - start;
        cnt = rem < (size_t)N ? rem : (size_t)N;
      }
      const float* __restrict__ src = a.pcm + (cnt ? start : 0);
      float2 cur[E];
#pragma unroll
      for (int q = 0; q < E / R0; ++q)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const int m = t + q * T + r * (L / R0);
          cur[q * R0 + r] = make_float2((size_t)(2 * m) < cnt ? src[2 * m] : 0.0f,
                                        (size_t)(2 * m + 1) < cnt ? src[2 * m + 1] : 0.0f);
        }
      float* __restrict__ ore = a.o_re ? a.o_re + f * (size_t)BINS : nullptr;
      float* __restrict__ oim = a.o_im ? a.o_im + f * (size_t)BINS : nullptr;
      float* __restrict__ omg = a.o_mag ? a.o_mag + f * (size_t)BINS : nullptr;
      spectrum(
          cur, [] {}, [](int) {},
          [&](int k, float2 X) {
            float A, B;
            stage_vals<MODE>(X, A, B);
            if (!active) return;
            if (ore) ore[k] = plane_value<MODE, 0>(A, B);
            if (oim) oim[k] = plane_value<MODE, 1>(A, B);
            if (omg) omg[k] = plane_value<MODE, 2>(A, B);
          });
    }
  };

  if (tile < 0) {  // workgroup 0: the frames after the last whole tile
    rest_frames();
    return;
  }

  // samples of iteration it's frame for this slot in tile tl, first-pass order (8-byte
  // loads: the host guarantees an even hop and 8-byte aligned samples)
  auto load_frame = [&](long tl, int it, float2* dst) {
    const size_t f0 = (size_t)tl * (size_t)(K * P);
    const float2* __restrict__ src =
        reinterpret_cast<const float2*>(a.pcm + (f0 + (size_t)(it * P + slot) + a.frame_offset) * (size_t)hop);
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) dst[q * R0 + r] = src[t + q * T + r * (L / R0)];
  };

  // A tile's plane regions as buffer descriptors; store round c of iteration it writes
  // the region's floats lo + 4*(tid + c*NT), whole 16-byte words; lanes past the
  // iteration's range (only in its last round) get the dropped offset.
  constexpr int kTileBytes = K * RUN * 4;
  struct Regions {
    __amdgpu_buffer_rsrc_t r[3];
  };
  auto regions = [&](long tl) {
    const size_t region = (size_t)tl * (size_t)(K * P) * (size_t)BINS;  // tile's first float in each plane
    Regions g;
    g.r[0] = plane_rsrc(a.o_re, region, kTileBytes);
    g.r[1] = plane_rsrc(a.o_im, region, kTileBytes);
    g.r[2] = plane_rsrc(a.o_mag, region, kTileBytes);
    return g;
  };
  // WHICH < 0: all three planes; else that plane only
  // OPT bit 7: tiles of even index run their iterations last frames first, so that the
  // samples they share with a neighbouring tile are read by both at about the same time
  // (and once through L2) rather than a tile's duration apart. Iteration it of such a tile
  // stores the mirror image of the forward range, [K*RUN - hi(it), K*RUN - lo(it)):
  // whole lines, as K*RUN is.
  constexpr bool kAlt = (OPT & 128) != 0;
  static_assert(!kAlt || (K * RUN) % 32 == 0, "mirrored store ranges need whole-line tiles");
  const bool rev = kAlt && tile >= 0 && (tile & 1) == 0;
  auto store_round = [&](const Regions& rg, int it, int c, int which) {
    const int lo = rev ? K * RUN - lines_hi<RUN, K>(it) : lines_lo<RUN, K>(it);
    const int n4 = (lines_hi<RUN, K>(it) - lines_lo<RUN, K>(it)) / 4;
    const int j = tid + c * NT;
    const int g = lo + 4 * j;
    // indexed in 16-byte words, so that the compiler emits one 16-byte LDS read (its
    // lanes cover consecutive banks) instead of 4-byte pairs that conflict
    int w4;
    if constexpr (kPow2Ring) {
      w4 = (g >> 2) & (RING / 4 - 1);
    } else {  // lo's word mod RING/4 (uniform) + j; lanes past the range (their stores
              // are dropped) read word w0, so every address stays inside the ring
      const int w0 = (lo >> 2) % (RING / 4);
      w4 = w0 + (j < n4 ? j : 0);  // n4 <= (RUN + 31) / 4 < RING / 4
      if (w4 >= RING / 4) w4 -= RING / 4;
    }
    const f32x4 A4 = reinterpret_cast<const f32x4*>(&ring[0][0])[w4];
    f32x4 B4 = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (kB) B4 = reinterpret_cast<const f32x4*>(&ring[kB ? 1 : 0][0])[w4];
    f32x4 o0, o1, o2;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      o0[u] = plane_value<MODE, 0>(A4[u], B4[u]);
      o1[u] = plane_value<MODE, 1>(A4[u], B4[u]);
      o2[u] = plane_value<MODE, 2>(A4[u], B4[u]);
    }
    const int off = (c * NT + NT <= n4 || j < n4) ? 4 * g : kDropOffset;
    if (which < 0 || which == 0) store_nt16(o0, rg.r[0], off);
    if (which < 0 || which == 1) store_nt16(o1, rg.r[1], off);
    if (which < 0 || which == 2) store_nt16(o2, rg.r[2], off);
  };
  // iteration it's store rounds during pass `pass` of the next iteration
  auto stores_in_pass = [&](const Regions& rg, int it, int pass) {
    constexpr int C_PREV = lines_max_rounds<RUN, K, NT>();
    const int rounds = lines_rounds<RUN, K, NT>(it);
    if constexpr ((OPT & 2) && NP == 3) {
      // OPT bit 1: pass p writes every round of plane p (one write stream at a time)
#pragma unroll
      for (int c = 0; c < C_PREV; ++c)
        if (c < rounds) store_round(rg, it, c, pass);
    } else {
      const int per = (rounds + NP - 1) / NP;
#pragma unroll
      for (int c = 0; c < C_PREV; ++c)
        if (c < rounds && c / per == pass) store_round(rg, it, c, -1);
    }
  };

  const Regions rg = regions(tile);
  float2 cur[E];
  load_frame(tile, rev ? K - 1 : 0, cur);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    // iteration i transforms the frames of position pos; bin k of slot s lands in the ring
    // at float (pos*P + s)*BINS + k
    const int pos = rev ? K - 1 - i : i;
    const int rbase = kPow2Ring ? (pos * P + slot) * BINS : ((pos * P + slot) * BINS) % RING;
    spectrum(
        cur,
        [&] {
          if (i + 1 < K) load_frame(tile, rev ? K - 2 - i : i + 1, cur);
        },
        [&](int pass) {  // the previous iteration's store rounds, spread over the passes
          if (i > 0) stores_in_pass(rg, i - 1, pass);
        },
        [&](int k, float2 X) {
          float A, B;
          stage_vals<MODE>(X, A, B);
          int g;
          if constexpr (kPow2Ring) {
            g = (rbase + k) & (RING - 1);
          } else {  // rbase < RING, k <= L < RING
            g = rbase + k;
            if (g >= RING) g -= RING;
          }
          ring[0][g] = A;
          if constexpr (kB) ring[kB ? 1 : 0][g] = B;
        });
  }
  {
    constexpr int rounds = lines_rounds<RUN, K, NT>(K - 1);
#pragma unroll
    for (int c = 0; c < rounds; ++c) store_round(rg, K - 1, c, -1);
  }
}

// Waves per SIMD the lines kernel is compiled for: 4 (3 with register tables); at N = 2048
// two, which its LDS (frame buffers and ring, about 68 KB per workgroup) allows anyway.
template <int L, int OPT>
constexpr int lines_waves_per_eu() {
  return L >= 1024 ? 2 : ((OPT & 4) ? 3 : 4);
}

template <int L, int E, int K, int OPT, int MODE, int DIR, int... Rs>
__global__ void __launch_bounds__(256, (lines_waves_per_eu<L, OPT>())) stft_r2c_lines_kernel(StftArgs a) {
  stft_lines_body<L, E, K, OPT, MODE, DIR, Rs...>(a, (int)blockIdx.x, (int)gridDim.x);
}

// The blocks of a single-request launch of the line-aligned kernel: whole tiles of
// regular frames (frames fully inside the PCM) plus, when at most 2*P*K frames remain, the
// rest in block 0. Returns whether the request qualifies (window = N, even hop, aligned
// samples, window and planes); `lead` = frames in whole tiles, `blocks` = blocks of the
// launch, `take` = frames the launch covers (the rest needs another path).
template <int L, int E, int K>
bool lines_layout(const StftArgs& a, size_t* lead, size_t* blocks, size_t* take) {
  constexpr int N = 2 * L;
  constexpr int P = 256 / (L / E);
  auto al = [](const void* q, uintptr_t m) { return (((uintptr_t)q) & (m - 1)) == 0; };
  const bool eligible = a.bins == L + 1 && a.window == N && a.hop > 0 && (a.hop % 2) == 0 && al(a.pcm, 8) &&
                        al(a.win, 16) && al(a.o_re, 16) && al(a.o_im, 16) && al(a.o_mag, 16);
  size_t nreg = 0;
  if (eligible && a.pcm_len >= (size_t)N) {
    const size_t fit = (a.pcm_len - N) / (size_t)a.hop + 1;  // frames fully inside the PCM
    nreg = fit > a.frame_offset ? fit - a.frame_offset : 0;
    if (nreg > a.frames) nreg = a.frames;
    nreg -= nreg % ((size_t)P * K);  // whole tiles only
  }
  *take = eligible && a.frames - nreg <= 2 * (size_t)P * K ? a.frames : nreg;
  *lead = nreg;
  *blocks = nreg / ((size_t)P * K) + (*take > nreg ? 1 : 0);
  return eligible;
}

template <int L, int E, int FPW, int IP, int DIR, int... Rs>
hipError_t stft_launch_impl(const StftArgs& a, hipStream_t s) {
  const size_t grid = (a.frames + FPW - 1) / FPW;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  const dim3 g((unsigned)grid), blk(FPW * (L / E));
  const bool half = a.bins == L + 1;
#define FFTFREE_STFT_GO(MODE)                                                                     \
  do {                                                                                            \
    if (half)                                                                                     \
      hipLaunchKernelGGL((stft_r2c_kernel<L, E, FPW, IP, MODE, 1, DIR, Rs...>), g, blk, 0, s, a); \
    else                                                                                          \
      hipLaunchKernelGGL((stft_r2c_kernel<L, E, FPW, IP, MODE, 0, DIR, Rs...>), g, blk, 0, s, a); \
  } while (0)
  switch (a.out_mode) {
    case kOutComplex: FFTFREE_STFT_GO(kOutComplex); break;
    case kOutC2RClear: FFTFREE_STFT_GO(kOutC2RClear); break;
    case kOutMagnitude: FFTFREE_STFT_GO(kOutMagnitude); break;
    default: FFTFREE_STFT_GO(kOutPolar); break;
  }
#undef FFTFREE_STFT_GO
  return hipGetLastError();
}

// Multi-frame launch: whole tiles of regular frames plus a short remainder go to one
// stft_r2c_mf_kernel launch; anything it cannot take goes to stft_r2c_kernel.
template <int L, int E, int K, int PL, int DIR, int... Rs>
hipError_t stft_mf_launch_impl(const StftArgs& a, hipStream_t s) {
  constexpr int N = 2 * L;
  constexpr int P = 256 / (L / E);
  size_t nreg = 0;
  const bool aligned = ((((uintptr_t)a.pcm) & 15) == 0) && (a.hop % 4) == 0 && a.hop <= N / 4;
  auto al16 = [](const float* q) { return (((uintptr_t)q) & 15) == 0; };
  const bool planes_ok = al16(a.o_re) && al16(a.o_im) && al16(a.o_mag) && ((P * a.bins) % 4 == 0);
  const bool eligible = a.window == N && aligned && planes_ok;
  if (eligible && a.pcm_len >= (size_t)N) {
    const size_t fit = (a.pcm_len - N) / (size_t)a.hop + 1;  // frames fully inside the PCM
    nreg = fit > a.frame_offset ? fit - a.frame_offset : 0;
    if (nreg > a.frames) nreg = a.frames;
    nreg -= nreg % ((size_t)P * K);  // whole workgroups only
  }
  // a short remainder rides along in workgroup 0 of the same launch
  const size_t take = eligible && a.frames - nreg <= 2 * (size_t)P * K ? a.frames : nreg;
  if (take > 0) {
    if (!a.sink) return hipErrorInvalidValue;
    StftArgs r = a;
    r.frames = take;
    r.lead = nreg;
    size_t grid = nreg / ((size_t)P * K) + (take > nreg ? 1 : 0);
    const bool half = a.bins == L + 1;
    if constexpr (PL == 2) {
      // one resident round: as many workgroups as the chip holds at once, or fewer
      const size_t tiles = nreg / ((size_t)P * K);
      const void* k = half ? reinterpret_cast<const void*>(&stft_r2c_mf_kernel<L, E, K, PL, kOutMagnitude, 1, DIR, Rs...>)
                           : reinterpret_cast<const void*>(&stft_r2c_mf_kernel<L, E, K, PL, kOutMagnitude, 0, DIR, Rs...>);
      const size_t slots = resident_slots(k);
      grid = tiles < slots ? tiles : slots;
      if (grid == 0) grid = 1;
      // the rest rides with a rank that runs the smaller tile count
      r.rest_rank = 0;
      for (size_t q = 0; q < grid; ++q)
        if ((q + 1) * tiles / grid - q * tiles / grid == tiles / grid) {
          r.rest_rank = (int)q;
          break;
        }
    }
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    const dim3 g((unsigned)grid), blk(256);
#define FFTFREE_MF_GO(MODE)                                                                    \
  do {                                                                                         \
    if (half)                                                                                  \
      hipLaunchKernelGGL((stft_r2c_mf_kernel<L, E, K, PL, MODE, 1, DIR, Rs...>), g, blk, 0, s, r); \
    else                                                                                       \
      hipLaunchKernelGGL((stft_r2c_mf_kernel<L, E, K, PL, MODE, 0, DIR, Rs...>), g, blk, 0, s, r); \
  } while (0)
    switch (a.out_mode) {
      case kOutComplex: FFTFREE_MF_GO(kOutComplex); break;
      case kOutC2RClear: FFTFREE_MF_GO(kOutC2RClear); break;
      case kOutMagnitude: FFTFREE_MF_GO(kOutMagnitude); break;
      default: FFTFREE_MF_GO(kOutPolar); break;
    }
#undef FFTFREE_MF_GO
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (take == a.frames) return hipSuccess;
  StftArgs rest = a;
  rest.frames = a.frames - take;
  rest.frame_offset = a.frame_offset + take;
  const size_t skip = take * (size_t)a.bins;
  if (rest.o_re) rest.o_re += skip;
  if (rest.o_im) rest.o_im += skip;
  if (rest.o_mag) rest.o_mag += skip;
  return stft_launch_impl<L, E, P, 1, DIR, Rs...>(rest, s);
}

// Resident workgroups of a multi-frame variant (its magnitude, half-spectrum kernel).
template <int L, int E, int K, int WPC, int OPT, int... Rs>
size_t lines_slots() {
  const size_t slots =
      resident_slots(reinterpret_cast<const void*>(&stft_r2c_lines_kernel<L, E, K, OPT, kOutMagnitude, -1, Rs...>));
  const size_t cap = (size_t)device_cus() * (WPC > 0 ? WPC : 1024);
  return cap < slots ? cap : slots;
}
// Line-aligned launch: whole tiles of regular frames plus a short remainder go to one
// stft_r2c_lines_kernel launch; anything it cannot take goes to stft_r2c_kernel. Full
// spectra (bins = N) go to the tile-preload multi-frame kernel (their rows are whole
// lines already).
template <int L, int E, int K, int WPC, int OPT, int DIR, int... Rs>
hipError_t stft_lines_launch_impl(const StftArgs& a, hipStream_t s) {
  constexpr int P = 256 / (L / E);
  if (a.bins != L + 1) {
    if constexpr (L / E <= 64)
      return stft_mf_launch_impl<L, E, 6, 1, DIR, Rs...>(a, s);
    else
      return stft_launch_impl<L, E, P, 1, DIR, Rs...>(a, s);
  }
  // whole tiles, and a short remainder riding along in workgroup 0 of the same launch
  size_t nreg = 0, grid = 0, take = 0;
  lines_layout<L, E, K>(a, &nreg, &grid, &take);
  if (take > 0) {
    StftArgs r = a;
    r.frames = take;
    r.lead = nreg;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    const dim3 g((unsigned)grid), blk(256);
#define FFTFREE_LINES_GO(MODE)                                                                        \
  do {                                                                                                \
    auto* k = &stft_r2c_lines_kernel<L, E, K, OPT, MODE, DIR, Rs...>;                                       \
    hipLaunchKernelGGL(k, g, blk, lds_cap_bytes(reinterpret_cast<const void*>(k), WPC), s, r);         \
  } while (0)
    switch (a.out_mode) {
      case kOutComplex: FFTFREE_LINES_GO(kOutComplex); break;
      case kOutC2RClear: FFTFREE_LINES_GO(kOutC2RClear); break;
      case kOutMagnitude: FFTFREE_LINES_GO(kOutMagnitude); break;
      default: FFTFREE_LINES_GO(kOutPolar); break;
    }
#undef FFTFREE_LINES_GO
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (take == a.frames) return hipSuccess;
  StftArgs rest = a;
  rest.frames = a.frames - take;
  rest.frame_offset = a.frame_offset + take;
  const size_t skip = take * (size_t)a.bins;
  if (rest.o_re) rest.o_re += skip;
  if (rest.o_im) rest.o_im += skip;
  if (rest.o_mag) rest.o_mag += skip;
  return stft_launch_impl<L, E, P, 1, DIR, Rs...>(rest, s);
}

template <int L, int E, int K, int PL, int... Rs>
size_t mf_slots() {
  return resident_slots(reinterpret_cast<const void*>(&stft_r2c_mf_kernel<L, E, K, PL, kOutMagnitude, 1, -1, Rs...>));
}

}  // namespace


#define FFTFREE_STFT(ORD, L, E, FPW, IP, RS, ...)                                         \
  {2 * (L), ORD, {L, E, FPW, RS}, &stft_launch_impl<L, E, FPW, IP, -1, __VA_ARGS__>,           \
   &stft_launch_impl<L, E, FPW, IP, 1, __VA_ARGS__>, nullptr}
#define FFTFREE_STFT_MF(ORD, L, E, K, RS, ...)                                           \
  {2 * (L), ORD, {L, E, K, RS}, &stft_mf_launch_impl<L, E, K, 0, -1, __VA_ARGS__>,             \
   &stft_mf_launch_impl<L, E, K, 0, 1, __VA_ARGS__>, &mf_slots<L, E, K, 0, __VA_ARGS__>}
#define FFTFREE_STFT_PL(ORD, L, E, K, RS, ...)                                           \
  {2 * (L), ORD, {L, E, K, RS}, &stft_mf_launch_impl<L, E, K, 1, -1, __VA_ARGS__>,             \
   &stft_mf_launch_impl<L, E, K, 1, 1, __VA_ARGS__>, &mf_slots<L, E, K, 1, __VA_ARGS__>}
#define FFTFREE_STFT_LN(ORD, L, E, K, WPC, OPT, RS, ...)                                 \
  {2 * (L), ORD, {L, E, K, RS}, &stft_lines_launch_impl<L, E, K, WPC, OPT, -1, __VA_ARGS__>,   \
   &stft_lines_launch_impl<L, E, K, WPC, OPT, 1, __VA_ARGS__>,                            \
   &lines_slots<L, E, K, WPC, OPT, __VA_ARGS__>}

}  // namespace fftfree

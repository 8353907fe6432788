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
#include "fft_stft_tile.hpp"

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

namespace fftfree {
using namespace dev;

namespace {

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
#pragma unroll
  for (int q = 0; q < E / R0; ++q)
#pragma unroll
    for (int r = 0; r < R0; ++r)
      wr[q * R0 + r] = win ? w2[t + q * T + r * (L / R0)] : make_float2(1.0f, 1.0f);
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

// Workgroups of kernel k that the current device holds at once (CUs x occupancy at 256
// threads), cached per device.
size_t resident_slots(const void* k) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find({k, dev});
  if (it != cache.end()) return it->second;
  int cus = 0, occ = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, 256, 0) != hipSuccess || occ <= 0) occ = 1;
  const size_t slots = (size_t)cus * (size_t)occ;
  cache[{k, dev}] = slots;
  return slots;
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

#define FFTFREE_STFT(L, E, FPW, IP, RS, ...)                                              \
  {2 * (L), {L, E, FPW, RS}, &stft_launch_impl<L, E, FPW, IP, -1, __VA_ARGS__>,           \
   &stft_launch_impl<L, E, FPW, IP, 1, __VA_ARGS__>}
#define FFTFREE_STFT_MF(L, E, K, RS, ...)                                                \
  {2 * (L), {L, E, K, RS}, &stft_mf_launch_impl<L, E, K, 0, -1, __VA_ARGS__>,             \
   &stft_mf_launch_impl<L, E, K, 0, 1, __VA_ARGS__>}
#define FFTFREE_STFT_PL(L, E, K, RS, ...)                                                \
  {2 * (L), {L, E, K, RS}, &stft_mf_launch_impl<L, E, K, 1, -1, __VA_ARGS__>,             \
   &stft_mf_launch_impl<L, E, K, 1, 1, __VA_ARGS__>}
#define FFTFREE_STFT_PP(L, E, K, RS, ...)                                                \
  {2 * (L), {L, E, K, RS}, &stft_mf_launch_impl<L, E, K, 2, -1, __VA_ARGS__>,             \
   &stft_mf_launch_impl<L, E, K, 2, 1, __VA_ARGS__>}
#define FFTFREE_C2C(L, E, FPW, RS, ...)                                                   \
  {L, {L, E, FPW, RS}, &c2c_launch_impl<L, E, FPW, -1, __VA_ARGS__>,                     \
   &c2c_launch_impl<L, E, FPW, 1, __VA_ARGS__>}

// Real-input variants per N (first one per N is the default).
const StftEntry kStft[] = {
    FFTFREE_STFT(8, 8, 64, 0, "8", 8),
    FFTFREE_STFT(16, 16, 64, 0, "16", 16),
    FFTFREE_STFT(32, 8, 64, 1, "8x4 inplace", 8, 4),
    FFTFREE_STFT(64, 8, 32, 1, "8x8 inplace", 8, 8),
    FFTFREE_STFT(128, 16, 32, 1, "16x8 inplace", 16, 8),
    FFTFREE_STFT_MF(256, 16, 4, "16x16 multiframe k=4", 16, 16),
    FFTFREE_STFT(256, 16, 16, 1, "16x16 inplace", 16, 16),
    FFTFREE_STFT_PL(512, 8, 6, "8x8x8 tile-preload k=6", 8, 8, 8),
    FFTFREE_STFT_PL(512, 8, 5, "8x8x8 tile-preload k=5", 8, 8, 8),
    FFTFREE_STFT_PL(512, 8, 4, "8x8x8 tile-preload k=4", 8, 8, 8),
    FFTFREE_STFT_PP(512, 8, 4, "8x8x8 persistent k=4", 8, 8, 8),
    FFTFREE_STFT_MF(512, 8, 11, "8x8x8 multiframe k=11", 8, 8, 8),
    FFTFREE_STFT_MF(512, 8, 6, "8x8x8 multiframe k=6", 8, 8, 8),
    FFTFREE_STFT_MF(512, 8, 8, "8x8x8 multiframe k=8", 8, 8, 8),
    FFTFREE_STFT_MF(512, 8, 4, "8x8x8 multiframe k=4", 8, 8, 8),
    FFTFREE_STFT(512, 8, 4, 1, "8x8x8 inplace", 8, 8, 8),
    FFTFREE_STFT(512, 8, 4, 0, "8x8x8 pingpong", 8, 8, 8),
    FFTFREE_STFT(512, 8, 8, 1, "8x8x8 inplace", 8, 8, 8),
    FFTFREE_STFT(512, 16, 8, 1, "16x16x2 inplace", 16, 16, 2),
    FFTFREE_STFT_PL(512, 8, 2, "8x8x8 small-batch tile-preload k=2", 8, 8, 8),
    FFTFREE_STFT_PL(512, 8, 1, "8x8x8 small-batch tile-preload k=1", 8, 8, 8),
    FFTFREE_STFT_MF(1024, 16, 8, "16x16x4 multiframe k=8", 16, 16, 4),
    FFTFREE_STFT(1024, 16, 4, 1, "16x16x4 inplace", 16, 16, 4),
    FFTFREE_STFT(1024, 8, 2, 1, "8x8x8x2 inplace", 8, 8, 8, 2),
    FFTFREE_STFT(2048, 16, 2, 1, "16x16x8 inplace", 16, 16, 8),
    FFTFREE_STFT(2048, 8, 1, 1, "8x8x8x4 inplace", 8, 8, 8, 4),
    FFTFREE_STFT(4096, 16, 1, 1, "16x16x16 inplace", 16, 16, 16),
};

const C2CEntry kC2C[] = {
    FFTFREE_C2C(8, 8, 64, "8", 8),
    FFTFREE_C2C(16, 16, 64, "16", 16),
    FFTFREE_C2C(32, 8, 64, "8x4", 8, 4),
    FFTFREE_C2C(64, 8, 32, "8x8", 8, 8),
    FFTFREE_C2C(128, 16, 32, "16x8", 16, 8),
    FFTFREE_C2C(256, 16, 16, "16x16", 16, 16),
    FFTFREE_C2C(512, 8, 4, "8x8x8", 8, 8, 8),
    FFTFREE_C2C(1024, 8, 2, "8x8x8x2", 8, 8, 8, 2),  // 13% faster at 65533 transforms, equal at 64
    FFTFREE_C2C(1024, 16, 4, "16x16x4", 16, 16, 4),
    FFTFREE_C2C(2048, 16, 2, "16x16x8", 16, 16, 8),
    FFTFREE_C2C(4096, 16, 1, "16x16x16", 16, 16, 16),
};

#undef FFTFREE_STFT
#undef FFTFREE_C2C
#undef FFTFREE_STFT_MF
#undef FFTFREE_STFT_PL
#undef FFTFREE_STFT_PP

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

// Default kernel variant for an STFT: among the leading table entries of N that share
// the first entry's kind (same description up to "k="), the tile size K whose whole
// tiles fill the resident workgroup slots best: fewest rounds x K, ties to the larger K.
// Tiles are P = 256/(L/E) frames by K; the multi-frame kernels hold 3 workgroups per CU.
// When that choice leaves more than half of the slots empty (a small batch), the
// "small-batch" entries of N (the same kernel with K = 2, 1) compete on the same cost.
int default_stft_variant(const StftArgs& a) {
  const int count = count_variants(kStft, a.n);
  const StftEntry* first = find_variant(kStft, a.n, 0);
  if (!first || count < 2) return 0;
  const char* k0 = std::strstr(first->info.radices, "k=");
  if (!k0) return 0;
  const size_t stem = (size_t)(k0 - first->info.radices);
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  const double slots = 3.0 * (cus > 0 ? cus : 256);
  int best = 0;
  double best_cost = 1e300;
  for (int i = 0; i < count; ++i) {
    const StftEntry* e = find_variant(kStft, a.n, i);
    if (std::strncmp(e->info.radices, first->info.radices, stem + 2) != 0) break;
    const int P = 256 / (e->info.l / e->info.e);
    const double tiles = std::floor((double)a.frames / (double)(P * e->info.fpw));
    const double cost = std::ceil(tiles / slots) * e->info.fpw;
    if (cost < best_cost || (cost == best_cost && e->info.fpw > find_variant(kStft, a.n, best)->info.fpw)) {
      best_cost = cost;
      best = i;
    }
  }
  // small batches: partial tiles count too (the launcher runs a remainder in workgroup 0)
  auto small_cost = [&](const StftEntry* e) {
    const int P = 256 / (e->info.l / e->info.e);
    const double tiles = std::ceil((double)a.frames / (double)(P * e->info.fpw));
    return std::ceil(tiles / slots) * e->info.fpw;
  };
  const StftEntry* b = find_variant(kStft, a.n, best);
  const double best_tiles = std::floor((double)a.frames / (double)(256 / (b->info.l / b->info.e) * b->info.fpw));
  if (best_tiles * 2.0 >= slots) return best;
  const int chosen = best;
  best_cost = small_cost(b);
  for (int i = 0; i < count; ++i) {
    const StftEntry* e = find_variant(kStft, a.n, i);
    if (!std::strstr(e->info.radices, "small-batch")) continue;
    const double cost = small_cost(e);
    if (cost < best_cost || (cost == best_cost && best != chosen && e->info.fpw > find_variant(kStft, a.n, best)->info.fpw)) {
      best_cost = cost;
      best = i;
    }
  }
  return best;
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
  if (variant < 0) variant = default_stft_variant(a);
  const StftEntry* e = find_variant(kStft, a.n, variant);
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

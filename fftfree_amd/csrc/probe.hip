// probe.hip -- HBM bandwidth probes for the STFT's memory pattern (measurement only; the
// library proper does not use them). fft_probe_planes moves exactly the bytes the
// line-aligned STFT kernel moves -- each tile's span of samples read, each tile's region
// of the three planes written as whole 16-byte words -- with no transform in between, so
// its time is the floor that memory alone sets for that pattern; fft_probe_large does the
// same for the two passes of the 2^20-point transform. Built into its own library,
// libfft_probe.so, loaded by bench.py's write_probe and large_probe workloads.
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct ProbeArgs {
  const float* pcm;
  size_t tile_hop;  // floats between the first samples of consecutive tiles
  int span4;        // 16-byte words of samples per tile
  float* plane[3];  // any may be nullptr (not written)
  int tile4;        // 16-byte words per plane per tile
  int tiles;
  int flags;
};

enum : int {
  kNonTemporal = 1,  // streaming stores
  kInterleave = 2,   // each round stores one word of every plane (else plane after plane)
  kNoReads = 4,
  kNoWrites = 8,
  kBuffer = 16,      // buffer stores through a descriptor per plane region (as the STFT kernel)
  // bits 8-12: the buffer stores' cache-policy operand when kPolicy is set (else 2 = nt
  // with kNonTemporal, 0 without)
  kPolicy = 32,
};

// Buffer store with a cache-policy operand from a small fixed set (it must be a constant).
__device__ __forceinline__ void store_policy(f32x4 v, __amdgpu_buffer_rsrc_t r, int off, int policy) {
  const u32x4 d = __builtin_bit_cast(u32x4, v);
  switch (policy) {
    case 1: __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 1); break;
    case 2: __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 2); break;
    case 3: __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 3); break;
    case 16: __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 16); break;
    case 17: __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 17); break;
    case 18: __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 18); break;
    case 19: __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 19); break;
    default: __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 0); break;
  }
}

__global__ void __launch_bounds__(256) probe_planes_kernel(ProbeArgs a) {
  extern __shared__ float pad[];  // only there to cap the workgroups per CU
  const int tid = threadIdx.x;
  float acc = 0.0f;
  const f32x4 val = {1.0f, 2.0f, 3.0f, (float)tid};
  for (int tile = blockIdx.x; tile < a.tiles; tile += gridDim.x) {
    if (!(a.flags & kNoReads)) {
      const f32x4* __restrict__ src = reinterpret_cast<const f32x4*>(a.pcm + (size_t)tile * a.tile_hop);
      for (int i = tid; i < a.span4; i += 256) {
        const f32x4 v = src[i];
        acc += (v.x + v.y) + (v.z + v.w);
      }
    }
    if (a.flags & kNoWrites) continue;
    const size_t base = (size_t)tile * (size_t)a.tile4;
    if (a.flags & kBuffer) {
      __amdgpu_buffer_rsrc_t r[3];
#pragma unroll
      for (int p = 0; p < 3; ++p)
        r[p] = __builtin_amdgcn_make_buffer_rsrc(a.plane[p] ? static_cast<void*>(a.plane[p] + 4 * base) : nullptr, 0,
                                                 a.plane[p] ? a.tile4 * 16 : 0, 0x00020000);
      const int policy = (a.flags & kPolicy) ? ((a.flags >> 8) & 31) : ((a.flags & kNonTemporal) ? 2 : 0);
      if (a.flags & kInterleave) {
        for (int i = tid; i < a.tile4; i += 256)
#pragma unroll
          for (int p = 0; p < 3; ++p) store_policy(val, r[p], 16 * i, policy);
      } else {
#pragma unroll
        for (int p = 0; p < 3; ++p)
          for (int i = tid; i < a.tile4; i += 256) store_policy(val, r[p], 16 * i, policy);
      }
      continue;
    }
    auto put = [&](int p, int i) {
      f32x4* dst = reinterpret_cast<f32x4*>(a.plane[p]) + base + i;
      if (a.flags & kNonTemporal)
        __builtin_nontemporal_store(val, dst);
      else
        *dst = val;
    };
    if (a.flags & kInterleave) {
      for (int i = tid; i < a.tile4; i += 256)
#pragma unroll
        for (int p = 0; p < 3; ++p)
          if (a.plane[p]) put(p, i);
    } else {
#pragma unroll
      for (int p = 0; p < 3; ++p)
        if (a.plane[p])
          for (int i = tid; i < a.tile4; i += 256) put(p, i);
    }
  }
  // keeps the loads; never true for finite samples below 1e30
  if (acc == 1e38f && a.plane[0]) a.plane[0][tid] = acc;
  (void)pad;
}

// The access patterns of the two-pass large transform's passes at N = 1024 x 1024, with no
// transform: a workgroup of 512 threads moves a tile of C lines x 1024 values (float2).
//   MODE 0 (pass A): the C adjacent columns read (runs of C values, stride 1024) and
//                    written back in place in `out` (the same runs);
//   MODE 1 (pass B): C rows read contiguously, written as runs of C values, stride 1024;
//   MODE 2:          the tile read and written contiguously (a copy).
// Tile t covers transform t / (1024 / C) of the launch; every index stays inside the
// launch's `batch` transforms of 2^20 values.
constexpr int kLargeSide = 1024;
constexpr size_t kLargeN = (size_t)kLargeSide * kLargeSide;

template <int C, int MODE, bool NTS>
__global__ void __launch_bounds__(512) probe_large_kernel(const float2* __restrict__ in, float2* __restrict__ out) {
  extern __shared__ float pad[];  // only there to cap the workgroups per CU
  constexpr int NT = 512;
  constexpr int E = kLargeSide * C / NT;
  constexpr int groups = kLargeSide / C;
  const int tid = threadIdx.x;
  const size_t b = blockIdx.x / groups;
  const size_t c0 = (size_t)(blockIdx.x % groups) * C;
  const int col = tid % C, row0 = tid / C;
  float2 x[E];
  if (MODE == 0) {
#pragma unroll
    for (int k = 0; k < E; ++k) x[k] = in[b * kLargeN + (size_t)(row0 + k * (NT / C)) * kLargeSide + c0 + col];
  } else {
    const float2* __restrict__ src = in + b * kLargeN + c0 * kLargeSide;
#pragma unroll
    for (int k = 0; k < E; ++k) x[k] = src[tid + k * NT];
  }
  float2* __restrict__ dst = out + b * kLargeN;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const size_t idx = MODE == 2 ? c0 * kLargeSide + tid + k * NT : (size_t)(row0 + k * (NT / C)) * kLargeSide + c0 + col;
    if (NTS) {
      float* o = reinterpret_cast<float*>(dst + idx);
      __builtin_nontemporal_store(x[k].x, o);
      __builtin_nontemporal_store(x[k].y, o + 1);
    } else {
      dst[idx] = x[k];
    }
  }
  (void)pad;
}

template <int C, int MODE, bool NTS>
int launch_large_probe(const float2* in, float2* out, int batch, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((probe_large_kernel<C, MODE, NTS>), dim3((unsigned)(batch * (kLargeSide / C))), dim3(512), lds, s,
                     in, out);
  return (int)hipGetLastError();
}

}  // namespace

extern "C" {

// One launch of the large-transform pattern probe over `batch` transforms of 2^20 float2
// (in and out each hold batch * 2^20 values): mode 0 (pass A), 1 (pass B) or 2 (copy),
// c = 8 or 16 lines per workgroup, nts = streaming stores, wpc = at most that many
// workgroups per CU (0: no cap). Returns the launch status.
int fft_probe_large(const void* in, void* out, int batch, int mode, int c, int nts, int wpc, void* stream) {
  if (batch <= 0 || mode < 0 || mode > 2 || (c != 8 && c != 16)) return (int)hipErrorInvalidValue;
  size_t lds = 0;
  if (wpc > 0) {
    int dev = 0, cu_lds = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
        cu_lds > 0)
      lds = (size_t)cu_lds / (size_t)(wpc + 1) + 512;
  }
  const float2* i = static_cast<const float2*>(in);
  float2* o = static_cast<float2*>(out);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int key = mode * 4 + (c == 16 ? 2 : 0) + (nts ? 1 : 0);
  switch (key) {
    case 0: return launch_large_probe<8, 0, false>(i, o, batch, lds, s);
    case 1: return launch_large_probe<8, 0, true>(i, o, batch, lds, s);
    case 2: return launch_large_probe<16, 0, false>(i, o, batch, lds, s);
    case 3: return launch_large_probe<16, 0, true>(i, o, batch, lds, s);
    case 4: return launch_large_probe<8, 1, false>(i, o, batch, lds, s);
    case 5: return launch_large_probe<8, 1, true>(i, o, batch, lds, s);
    case 6: return launch_large_probe<16, 1, false>(i, o, batch, lds, s);
    case 7: return launch_large_probe<16, 1, true>(i, o, batch, lds, s);
    case 8: return launch_large_probe<8, 2, false>(i, o, batch, lds, s);
    case 9: return launch_large_probe<8, 2, true>(i, o, batch, lds, s);
    case 10: return launch_large_probe<16, 2, false>(i, o, batch, lds, s);
    default: return launch_large_probe<16, 2, true>(i, o, batch, lds, s);
  }
}


// Runs the probe over `tiles` tiles on `grid` workgroups of 256 threads (tile t on
// workgroup t mod grid), at most `wpc` workgroups per CU (0: no cap). The caller
// guarantees that tile t's samples [t*tile_hop, t*tile_hop + 4*span4) lie inside pcm and
// that each plane holds tiles*4*tile4 floats, 16-byte aligned. Returns the launch status.
int fft_probe_planes(const float* pcm, size_t tile_hop, int span4, float* p0, float* p1, float* p2, int tile4,
                     int tiles, int grid, int wpc, int flags, void* stream) {
  if (tiles <= 0 || grid <= 0 || span4 < 0 || tile4 <= 0) return (int)hipErrorInvalidValue;
  ProbeArgs a{pcm, tile_hop, span4, {p0, p1, p2}, tile4, tiles, flags};
  size_t lds = 0;
  if (wpc > 0) {
    int dev = 0, cu_lds = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
        cu_lds > 0)
      lds = (size_t)cu_lds / (size_t)(wpc + 1) + 512;
  }
  hipLaunchKernelGGL(probe_planes_kernel, dim3((unsigned)grid), dim3(256), lds, static_cast<hipStream_t>(stream), a);
  return (int)hipGetLastError();
}

}  // extern "C"

// fft_ragged.hip -- ragged batched STFT: many requests with different lengths, window
// sizes and hops in one persistent launch.
//
// Each request is an fft_execute_batched call of its own (frame rule fft_cffi.cpp:715-750,
// framing 766-791, post hook eigen_fft.hpp:2535-2596, output copy fft_cffi.cpp:811-825),
// described by a StftArgs block in device memory. The host cuts every request into work
// items of `ntiles` tiles (a tile = the FPW frames one 256-thread workgroup transforms
// at once, with the request's N, radix pattern and FPW from a configuration table),
// sorts the items by cost and splits them into 8 contiguous, cost-balanced chunks, one
// per XCD: the workgroups of XCD x (blockIdx % 8 == x, the dispatcher's round-robin
// order) stride through chunk x, so the neighbouring tiles of one request, which share
// samples, are read through the same L2. One launch, sized to the resident capacity of
// the chip, covers every request; the host captures it in a graph.
#include "fft_device.hpp"
#include "fft_kernels.hpp"
#include "fft_stft_tile.hpp"

namespace fftfree {
using namespace dev;

namespace {

// Tile shapes the ragged kernel can run: 256 threads = FPW transforms of L/E threads.
// Per N, the autotuner picks one of these and the tiles per work item.
#define FFTFREE_RAGGED_CFGS(X)                \
  X(0, 128, 8, 16, "8x8x2", 8, 8, 2)         \
  X(1, 128, 16, 32, "16x8", 16, 8)           \
  X(2, 256, 8, 8, "8x8x4", 8, 8, 4)          \
  X(3, 256, 16, 16, "16x16", 16, 16)         \
  X(4, 512, 8, 4, "8x8x8", 8, 8, 8)          \
  X(5, 512, 16, 8, "16x16x2", 16, 16, 2)     \
  X(6, 1024, 8, 2, "8x8x8x2", 8, 8, 8, 2)    \
  X(7, 1024, 16, 4, "16x16x4", 16, 16, 4)    \
  X(8, 2048, 8, 1, "8x8x8x4", 8, 8, 8, 4)    \
  X(9, 2048, 16, 2, "16x16x8", 16, 16, 8)

#define FFTFREE_RAGGED_INFO(ID, L, E, FPW, NAME, ...) {L, E, FPW, NAME},
const VariantInfo kRaggedCfgs[] = {FFTFREE_RAGGED_CFGS(FFTFREE_RAGGED_INFO)};
#undef FFTFREE_RAGGED_INFO
constexpr int kRaggedCfgCount = sizeof(kRaggedCfgs) / sizeof(kRaggedCfgs[0]);

// LDS of the largest tile shape (float2); every shape stages its bins inside it.
#define FFTFREE_RAGGED_LDS(ID, L, E, FPW, NAME, ...) \
  StftTile<L, E, FPW, 1, kOutPolar, 1, -1, __VA_ARGS__>::kLds,
constexpr int kRaggedLdsList[] = {FFTFREE_RAGGED_CFGS(FFTFREE_RAGGED_LDS)};
#undef FFTFREE_RAGGED_LDS
constexpr int max_lds() {
  int m = 0;
  for (int v : kRaggedLdsList) m = v > m ? v : m;
  return m;
}
constexpr int kRaggedLds = max_lds();

// Each tile shape is its own (not inlined) routine, so that register allocation sees
// one shape at a time.
template <int MODE, int ID, int L, int E, int FPW, int... Rs>
__device__ __noinline__ void ragged_tile(const StftArgs& a, size_t tile, float2* lds) {
  StftTile<L, E, FPW, 1, MODE, 1, -1, Rs...>::run(a, tile, lds, nullptr);
}

// Two kernels: family 0 holds the E=8 shapes, family 1 the E=16 ones. A kernel's register
// allocation is that of its largest shape, so keeping the lean shapes apart lets them
// run at a higher occupancy; the autotuner picks the family for the batch.
template <int MODE, int FAM>
__global__ void __launch_bounds__(256) stft_ragged_kernel(const StftArgs* __restrict__ reqs,
                                                          const RaggedItem* __restrict__ items,
                                                          const unsigned* __restrict__ chunks) {
  __shared__ float2 lds[kRaggedLds];
  const int x = blockIdx.x & 7;
  const unsigned per_xcd = (gridDim.x - x + 7) >> 3;  // workgroups dealt to this XCD
  for (unsigned i = chunks[x] + (blockIdx.x >> 3); i < chunks[x + 1]; i += per_xcd) {
    const RaggedItem w = items[i];
    const StftArgs a = reqs[w.req];
    for (unsigned k = 0; k < w.ntiles; ++k) {
      const size_t tile = (size_t)w.tile + k;
      switch (w.cfg) {
#define FFTFREE_RAGGED_CASE(ID, L, E, FPW, NAME, ...)                        \
  case ID:                                                                   \
    if constexpr ((E == 16) == (FAM == 1))                                    \
      ragged_tile<MODE, ID, L, E, FPW, __VA_ARGS__>(a, tile, lds);            \
    break;
        FFTFREE_RAGGED_CFGS(FFTFREE_RAGGED_CASE)
#undef FFTFREE_RAGGED_CASE
        default:
          break;
      }
      __syncthreads();  // the next tile reuses the LDS
    }
  }
}

template <int MODE>
const void* ragged_kernel_ptr(int family) {
  return family ? reinterpret_cast<const void*>(&stft_ragged_kernel<MODE, 1>)
                : reinterpret_cast<const void*>(&stft_ragged_kernel<MODE, 0>);
}

const void* ragged_kernel(int mode, int family) {
  switch (mode) {
    case kOutComplex: return ragged_kernel_ptr<kOutComplex>(family);
    case kOutMagnitude: return ragged_kernel_ptr<kOutMagnitude>(family);
    case kOutPolar: return ragged_kernel_ptr<kOutPolar>(family);
    default: return nullptr;
  }
}

}  // namespace

int ragged_cfg_count() { return kRaggedCfgCount; }
const VariantInfo* ragged_cfg(int id) { return id >= 0 && id < kRaggedCfgCount ? &kRaggedCfgs[id] : nullptr; }
int ragged_cfg_family(int id) { return id >= 0 && id < kRaggedCfgCount && kRaggedCfgs[id].e == 16 ? 1 : 0; }

hipError_t ragged_grid(int mode, int family, int* grid) {
  const void* k = ragged_kernel(mode, family);
  if (!k) return hipErrorInvalidValue;
  int dev = 0, cus = 0, occ = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, 256, 0);
  if (e != hipSuccess) return e;
  *grid = (cus > 0 ? cus : 1) * (occ > 0 ? occ : 1);
  return hipSuccess;
}

hipError_t launch_stft_ragged(const StftArgs* reqs, const RaggedItem* items, const unsigned* chunks, int mode,
                              int family, int grid, hipStream_t s) {
  const void* k = ragged_kernel(mode, family);
  if (!k || grid <= 0) return hipErrorInvalidValue;
  void* args[] = {&reqs, &items, &chunks};
  return hipLaunchKernel(k, dim3((unsigned)grid), dim3(256), args, 0, s);
}

}  // namespace fftfree

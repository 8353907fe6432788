// fft_ragged.hip -- ragged batched STFT: many requests with different lengths, window
// sizes and hops in one graph.
//
// Each request is an fft_execute_batched call of its own (frame rule fft_cffi.cpp:715-750,
// framing 766-791, post hook eigen_fft.hpp:2535-2596, output copy fft_cffi.cpp:811-825),
// described by a StftArgs block in device memory. The graph holds one launch per window
// size N in the batch, each compiled for that N alone (its own register allocation, no
// shape dispatch), the launches of different N on parallel branches. A launch runs one
// engine over the work items of its N:
//  * the line-aligned engine (N = 256 .. 4096): the body of stft_r2c_lines_kernel, whose
//    blocks are tiles of 32 or 64 frames whose plane regions are whole 128-byte lines,
//    plus a block for the frames after the last tile; one block per workgroup;
//  * the tile engine (every N; requests the line-aligned one does not take): StftTile,
//    FPW frames per tile with a radix pattern, persistent workgroups walking items of
//    `ntiles` blocks.
// The host cuts every request into work items, sorts the items of each N by cost and
// splits them into 8 contiguous, cost-balanced chunks, one per XCD: the workgroups of XCD
// x (blockIdx % 8 == x, the dispatcher's round-robin order) take chunk x in order, so the
// neighbouring tiles of one request, which share samples, are read through the same L2.
#include "fft_stft_impl.hpp"

namespace fftfree {
using namespace dev;

namespace {

// Engines: ID, L (= N/2), E, frames per block, name, radices. IDs below kLinesFirst are
// tile-engine shapes (FPW frames per tile); the others are line-aligned engines (tiles of
// FPW frames: K = FPW / P iterations of P = 256/(L/E) frames, a whole number of lines).
#define FFTFREE_RAGGED_TILES(X)               \
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
// Line-aligned rows also carry the body's OPT: 3 = tiles in launch order, one plane per
// pass; 263 = the same with the twiddle tables in registers and the window taps in LDS
// (N = 2048, whose LDS allows two workgroups per CU: registers are what is left); 515 = the
// window taps in registers, the twiddles in LDS (N = 4096, two waves per frame, whose LDS
// has room for the twiddles but not for 16 KB of taps). Walking even tiles last frames
// first (bit 7) took the batch's sample reads from 1.07x to 1.04x but not its time
// (753-757 us against 747-748 us), so the rows leave it out.
#define FFTFREE_RAGGED_LINES(X)                             \
  X(10, 128, 4, 64, "lines 4x4x4x2", 3, 4, 4, 4, 2)        \
  X(11, 256, 4, 32, "lines 4x4x4x4", 3, 4, 4, 4, 4)        \
  X(12, 512, 8, 32, "lines 8x8x8", 3, 8, 8, 8)             \
  X(13, 1024, 16, 32, "lines 16x16x4", 263, 16, 16, 4)      \
  X(14, 2048, 16, 32, "lines 16x16x8", 515, 16, 16, 8)
constexpr int kLinesFirst = 10;
// iterations per line-aligned tile of FPW frames (P = 256 / (L / E) frames per iteration)
template <int L, int E, int FPW>
constexpr int lines_k() {
  return FPW / (256 / (L / E));
}

#define FFTFREE_RAGGED_INFO(ID, L, E, FPW, NAME, ...) {L, E, FPW, NAME},
const VariantInfo kRaggedCfgs[] = {FFTFREE_RAGGED_TILES(FFTFREE_RAGGED_INFO) FFTFREE_RAGGED_LINES(FFTFREE_RAGGED_INFO)};
#undef FFTFREE_RAGGED_INFO
constexpr int kRaggedCfgCount = sizeof(kRaggedCfgs) / sizeof(kRaggedCfgs[0]);

// Items of XCD x: workgroups with blockIdx % 8 == x stride through [chunks[x], chunks[x+1]).
#define FFTFREE_RAGGED_ITEMS(BODY)                                                        \
  const int x = blockIdx.x & 7;                                                           \
  const unsigned per_xcd = (gridDim.x - x + 7) >> 3; /* workgroups dealt to this XCD */   \
  for (unsigned i = chunks[x] + (blockIdx.x >> 3); i < chunks[x + 1]; i += per_xcd) {     \
    const RaggedItem w = items[i];                                                        \
    for (unsigned k = 0; k < w.ntiles; ++k) {                                             \
      BODY;                                                                               \
      __syncthreads(); /* the next block reuses the LDS */                                \
    }                                                                                     \
  }

// One block is a call that is not inlined: register allocation then sees the block's
// routine alone, not the routine interleaved with the item loop (inlined, the compiler
// keeps values live across blocks and the kernels lose most of their occupancy). The
// request's arguments are passed as their global address (a struct in registers would be
// spilled to scratch for every call) and read with uniform loads.
template <int MODE, int L, int E, int FPW, int... Rs>
__device__ __noinline__ void ragged_tile_block(const StftArgs* __restrict__ ap, size_t tile, float2* lds) {
  const StftArgs a = *ap;
  StftTile<L, E, FPW, 1, MODE, 1, -1, Rs...>::run(a, tile, lds, nullptr);
}


// Tile engine, one shape per kernel.
template <int MODE, int L, int E, int FPW, int... Rs>
__global__ void __launch_bounds__(256) stft_ragged_tile_kernel(const StftArgs* __restrict__ reqs,
                                                               const RaggedItem* __restrict__ items,
                                                               const unsigned* __restrict__ chunks) {
  __shared__ float2 lds[StftTile<L, E, FPW, 1, MODE, 1, -1, Rs...>::kLds];
  FFTFREE_RAGGED_ITEMS((ragged_tile_block<MODE, L, E, FPW, Rs...>(reqs + w.req, (size_t)w.tile + k, lds)))
}

// Line-aligned engine, one block per workgroup: the workgroups of XCD x take the items of
// chunk x in order, one each (every item is one block, block 0 of a request being the
// frames after its last tile when it has them; the host set each request's `lead` and
// sizes the grid to 8 times the longest chunk). With no loop around the body its register
// allocation is the standalone kernel's (117 registers at N = 1024; the 256 of two waves
// per SIMD at N >= 2048), where a persistent loop over items needed 244 at N = 1024 and
// spilled at N = 2048 (as a call or inlined). The graph runs the launches of different N
// side by side, so each launch's last partial wave of workgroups overlaps other work.
template <int MODE, int L, int E, int K, int OPT, int... Rs>
__global__ void __launch_bounds__(256, (lines_waves_per_eu<L, OPT>()))
    stft_ragged_lines_kernel(const StftArgs* __restrict__ reqs, const RaggedItem* __restrict__ items,
                             const unsigned* __restrict__ chunks) {
  const int x = blockIdx.x & 7;
  const unsigned i = chunks[x] + (blockIdx.x >> 3);
  if (i >= chunks[x + 1]) return;  // uniform over the workgroup
  const RaggedItem w = items[i];
  // by reference: the body reads the request's fields from global memory (uniform loads);
  // a local copy would have its address taken and live in scratch
  const StftArgs& a = reqs[w.req];
  stft_lines_body<L, E, K, OPT, MODE, -1, Rs...>(a, (int)w.tile, 0);
}

template <int MODE>
const void* ragged_kernel_mode(int cfg) {
  switch (cfg) {
#define FFTFREE_RAGGED_TILE_CASE(ID, L, E, FPW, NAME, ...) \
  case ID: return reinterpret_cast<const void*>(&stft_ragged_tile_kernel<MODE, L, E, FPW, __VA_ARGS__>);
    FFTFREE_RAGGED_TILES(FFTFREE_RAGGED_TILE_CASE)
#undef FFTFREE_RAGGED_TILE_CASE
#define FFTFREE_RAGGED_LINES_CASE(ID, L, E, FPW, NAME, OPT, ...) \
  case ID:                                                                  \
    return reinterpret_cast<const void*>(&stft_ragged_lines_kernel<MODE, L, E, lines_k<L, E, FPW>(), OPT, __VA_ARGS__>);
    FFTFREE_RAGGED_LINES(FFTFREE_RAGGED_LINES_CASE)
#undef FFTFREE_RAGGED_LINES_CASE
    default: return nullptr;
  }
}

const void* ragged_kernel(int mode, int cfg) {
  switch (mode) {
    case kOutComplex: return ragged_kernel_mode<kOutComplex>(cfg);
    case kOutMagnitude: return ragged_kernel_mode<kOutMagnitude>(cfg);
    case kOutPolar: return ragged_kernel_mode<kOutPolar>(cfg);
    default: return nullptr;
  }
}

}  // namespace

int ragged_cfg_count() { return kRaggedCfgCount; }
const VariantInfo* ragged_cfg(int id) { return id >= 0 && id < kRaggedCfgCount ? &kRaggedCfgs[id] : nullptr; }
bool ragged_cfg_lines(int id) { return id >= kLinesFirst && id < kRaggedCfgCount; }
bool ragged_cfg_flat(int id) { return ragged_cfg_lines(id); }

bool ragged_lines_layout(int cfg, const StftArgs& a, size_t* lead, size_t* blocks, size_t* take) {
  switch (cfg) {
#define FFTFREE_RAGGED_LAYOUT_CASE(ID, L, E, FPW, NAME, ...) \
  case ID: return lines_layout<L, E, lines_k<L, E, FPW>()>(a, lead, blocks, take);
    FFTFREE_RAGGED_LINES(FFTFREE_RAGGED_LAYOUT_CASE)
#undef FFTFREE_RAGGED_LAYOUT_CASE
    default: return false;
  }
}

hipError_t ragged_grid(int mode, int cfg, int* grid) {
  const void* k = ragged_kernel(mode, cfg);
  if (!k) return hipErrorInvalidValue;
  int occ = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, 256, 0);
  if (e != hipSuccess) return e;
  *grid = device_cus() * (occ > 0 ? occ : 1);
  return hipSuccess;
}

hipError_t launch_stft_ragged(const StftArgs* reqs, const RaggedItem* items, const unsigned* chunks, int mode,
                              int cfg, int grid, hipStream_t s) {
  const void* k = ragged_kernel(mode, cfg);
  if (!k || grid <= 0) return hipErrorInvalidValue;
  void* args[] = {&reqs, &items, &chunks};
  return hipLaunchKernel(k, dim3((unsigned)grid), dim3(256), args, 0, s);
}

}  // namespace fftfree

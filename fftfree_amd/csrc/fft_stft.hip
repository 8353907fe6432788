// fft_stft.hip -- the real-input STFT variant table, the default-variant choice and
// launch_stft; the small-N and generic kernels:
// stft_tiny_kernel       direct DFT for N < 16.
// frame_complex_kernel, planes_kernel  the generic path around fft_apply_axis.
// The fused kernels live in fft_stft_impl.hpp and are instantiated here (stft_r2c_kernel
// rows), in fft_stft_lines.hip and in fft_stft_mf.hip.
#include "fft_stft_impl.hpp"

#include <algorithm>
#include <iterator>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace fftfree {

// Compute units of the current device, cached per device.
int device_cus() {
  static std::mutex mu;
  static std::map<int, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lk(mu);
  const auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  cache[dev] = cus;
  return cus;
}

// Workgroups of kernel k that the current device holds at once (CUs x occupancy at 256
// threads), cached per device.
size_t resident_slots(const void* k) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  {
    std::lock_guard<std::mutex> lk(mu);
    const auto it = cache.find({k, dev});
    if (it != cache.end()) return it->second;
  }
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, 256, 0) != hipSuccess || occ <= 0) occ = 1;
  const size_t slots = (size_t)device_cus() * (size_t)occ;
  std::lock_guard<std::mutex> lk(mu);
  cache[{k, dev}] = slots;
  return slots;
}

// Dynamic LDS bytes that cap kernel k at `wpc` workgroups per CU (0: no cap): the
// padding makes wpc + 1 workgroups overflow the CU's LDS. Cached per device.
size_t lds_cap_bytes(const void* k, int wpc) {
  if (wpc <= 0) return 0;
  static std::mutex mu;
  static std::map<std::tuple<const void*, int, int>, size_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_tuple(k, wpc, dev);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cu_lds = 0;
  hipFuncAttributes attr{};
  size_t pad = 0;
  if (hipDeviceGetAttribute(&cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
      hipFuncGetAttributes(&attr, k) == hipSuccess && cu_lds > 0) {
    const size_t per = (size_t)cu_lds / (size_t)(wpc + 1) + 512;  // just over a (wpc+1)-th
    pad = per > attr.sharedSizeBytes ? per - attr.sharedSizeBytes : 0;
    if ((attr.sharedSizeBytes + pad) * (size_t)wpc > (size_t)cu_lds) pad = 0;  // cannot cap this way
  }
  cache[key] = pad;
  return pad;
}

namespace {

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

// This file's rows of the variant table (`order` places them among the rows of the
// other translation units).
const StftEntry kLocal[] = {
    FFTFREE_STFT(0, 8, 8, 64, 0, "8", 8),
    FFTFREE_STFT(0, 16, 16, 64, 0, "16", 16),
    FFTFREE_STFT(0, 32, 8, 64, 1, "8x4 inplace", 8, 4),
    FFTFREE_STFT(0, 64, 8, 32, 1, "8x8 inplace", 8, 8),
    FFTFREE_STFT(0, 128, 16, 32, 1, "16x8 inplace", 16, 8),
    FFTFREE_STFT(1, 256, 16, 16, 1, "16x16 inplace", 16, 16),
    FFTFREE_STFT(4, 512, 8, 4, 1, "8x8x8 inplace", 8, 8, 8),
    FFTFREE_STFT(5, 512, 16, 8, 1, "16x16x2 inplace", 16, 16, 2),
    FFTFREE_STFT(2, 1024, 16, 4, 1, "16x16x4 inplace", 16, 16, 4),
    FFTFREE_STFT(3, 1024, 8, 2, 1, "8x8x8x2 inplace", 8, 8, 8, 2),
    FFTFREE_STFT(1, 2048, 16, 2, 1, "16x16x8 inplace", 16, 16, 8),
    FFTFREE_STFT(2, 2048, 8, 1, 1, "8x8x8x4 inplace", 8, 8, 8, 4),
    FFTFREE_STFT(0, 4096, 16, 1, 1, "16x16x16 inplace", 16, 16, 16),
};

#undef FFTFREE_STFT
#undef FFTFREE_STFT_MF
#undef FFTFREE_STFT_PL
#undef FFTFREE_STFT_LN

// The variant table: this file's rows and those of the kernel translation units, in
// order of N and, within N, of each row's `order`.
const std::vector<StftEntry>& stft_table() {
  static const std::vector<StftEntry> tab = [] {
    std::vector<StftEntry> t(std::begin(kLocal), std::end(kLocal));
    for (auto rows : {stft_lines_entries, stft_lines_exch_entries, stft_lines_order_entries, stft_lines_2048_entries,
                      stft_lines_4096_entries,
                      stft_mf_entries}) {
      int c = 0;
      const StftEntry* e = rows(&c);
      t.insert(t.end(), e, e + c);
    }
    std::stable_sort(t.begin(), t.end(), [](const StftEntry& x, const StftEntry& y) {
      return x.n != y.n ? x.n < y.n : x.order < y.order;
    });
    return t;
  }();
  return tab;
}

const StftEntry* find_variant(int n, int idx) {
  int seen = 0;
  for (const StftEntry& e : stft_table()) {
    if (e.n != n) continue;
    if (seen == idx) return &e;
    ++seen;
  }
  return nullptr;
}

int count_variants(int n) {
  int c = 0;
  for (const StftEntry& e : stft_table()) c += (e.n == n);
  return c;
}

constexpr int kTinyBlock = 256;

// Default kernel variant for an STFT. A persistent first entry of N takes every batch
// that gives each of its workgroups at least 4 iterations. Otherwise: among the leading
// multi-frame entries that share the kind of the first non-persistent one (same
// description up to "k="), the tile size K whose whole tiles fill the resident
// workgroup slots best: fewest rounds x K, ties to the larger K. Tiles are
// P = 256/(L/E) frames by K; a variant's resident slots are CUs x its kernel's
// occupancy. When that choice leaves more than half of the slots empty (a small batch),
// the "small-batch" entries of N (the same kernel with K = 2, 1) compete on the same
// cost, counting partial tiles.
int choose_stft_variant(const StftArgs& a) {
  const int count = count_variants(a.n);
  if (count < 2) return 0;
  auto frames_per_iter = [](const StftEntry* e) { return (double)(256 / (e->info.l / e->info.e)); };
  auto frames_per_tile = [&](const StftEntry* e) { return frames_per_iter(e) * e->info.fpw; };
  int base = 0;
  const StftEntry* first = find_variant(a.n, 0);
  if (first->slots && std::strstr(first->info.radices, "persistent")) {
    if (std::floor((double)a.frames / frames_per_iter(first)) >= 4.0 * (double)first->slots()) return 0;
    base = 1;
    first = find_variant(a.n, base);
  }
  if (!first || !first->slots) return 0;
  const char* k0 = std::strstr(first->info.radices, "k=");
  if (!k0) return base;
  const size_t stem = (size_t)(k0 - first->info.radices);
  int best = base;
  double best_cost = 1e300;
  for (int i = base; i < count; ++i) {
    const StftEntry* e = find_variant(a.n, i);
    if (std::strncmp(e->info.radices, first->info.radices, stem + 2) != 0 || !e->slots) break;
    const double tiles = std::floor((double)a.frames / frames_per_tile(e));
    const double cost = std::ceil(tiles / (double)e->slots()) * e->info.fpw;
    if (cost < best_cost || (cost == best_cost && e->info.fpw > find_variant(a.n, best)->info.fpw)) {
      best_cost = cost;
      best = i;
    }
  }
  // small batches: partial tiles count too (the launcher runs a remainder in workgroup 0)
  auto small_cost = [&](const StftEntry* e) {
    const double tiles = std::ceil((double)a.frames / frames_per_tile(e));
    return std::ceil(tiles / (double)e->slots()) * e->info.fpw;
  };
  const StftEntry* b = find_variant(a.n, best);
  if (std::floor((double)a.frames / frames_per_tile(b)) * 2.0 >= (double)b->slots()) return best;
  const int chosen = best;
  best_cost = small_cost(b);
  for (int i = 0; i < count; ++i) {
    const StftEntry* e = find_variant(a.n, i);
    if (!std::strstr(e->info.radices, "small-batch") || !e->slots) continue;
    const double cost = small_cost(e);
    if (cost < best_cost || (cost == best_cost && best != chosen && e->info.fpw > find_variant(a.n, best)->info.fpw)) {
      best_cost = cost;
      best = i;
    }
  }
  return best;
}

}  // namespace

int stft_max_lds_n() { return 8192; }
int stft_variant_count(int n) { return count_variants(n); }
const VariantInfo* stft_variant(int n, int idx) {
  const StftEntry* e = find_variant(n, idx);
  return e ? &e->info : nullptr;
}

// The choice of choose_stft_variant, kept per (device, N, frames): repeated launches of
// one shape skip the table scan and the occupancy queries.
int default_stft_variant(const StftArgs& a) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  static std::mutex mu;
  static std::map<std::tuple<int, int, size_t>, int> memo;
  const auto key = std::make_tuple(dev, a.n, a.frames);
  {
    std::lock_guard<std::mutex> lk(mu);
    const auto it = memo.find(key);
    if (it != memo.end()) return it->second;
  }
  const int v = choose_stft_variant(a);
  std::lock_guard<std::mutex> lk(mu);
  if (memo.size() > 4096) memo.clear();  // callers with ever-changing frame counts
  memo[key] = v;
  return v;
}

// For planes that the next launch reads straight back (Griffin-Lim's forward planes, read
// by its next inverse): the in-place tile entries of N, whose ordinary stores leave the
// planes in the caches where the default's streaming stores send them to HBM (at N = 1024,
// 4093 frames: 28.4 against 33.9 us per Griffin-Lim iteration). The one with the most
// frames per workgroup that still gives every CU a workgroup, else the fewest; -1 when N
// has none.
int readback_stft_variant(const StftArgs& a) {
  const int count = count_variants(a.n);
  const size_t cus = (size_t)device_cus();
  int best = -1, small = -1;
  for (int i = 0; i < count; ++i) {
    const StftEntry* e = find_variant(a.n, i);
    if (!std::strstr(e->info.radices, "inplace")) continue;
    const int fpw = e->info.fpw;
    if (small < 0 || fpw < find_variant(a.n, small)->info.fpw) small = i;
    if ((a.frames + fpw - 1) / (size_t)fpw >= cus && (best < 0 || fpw > find_variant(a.n, best)->info.fpw)) best = i;
  }
  return best >= 0 ? best : small;
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
  const StftEntry* e = find_variant(a.n, variant);
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


}  // namespace fftfree

// plan.hpp -- host-side plan objects of the GPU path, in the reference's shapes.
//
// The reference's C++ plan layer keeps these shapes, and so does this one:
//   KernelKind / KernelAccuracy      eigen_fft.hpp:1892-1894
//   AxisLayout                       eigen_fft.hpp:1503-1525
//   JobDispatcher::parallel_for      eigen_fft.hpp:317-347
//   Plan::KernelDescriptor           eigen_fft.hpp:2086-2097
//   Plan::use_kernel / set_dispatcher eigen_fft.hpp:2372 and the select_kernel family
//   fft_apply_axis                   eigen_fft.hpp:3104-3124
//   PlanRuntimeConfig                plan_support.hpp:24-46
//   PlanCache::get_plan              plan_support.hpp:409-426
// What changes is where the work runs: a Plan owns a device twiddle table and its
// descriptors launch HIP kernels on the dispatcher's stream. The reference's CPU
// WorkerPool fan-out over frames becomes the launch grid (one workgroup per group of
// frames), so the GPU dispatcher hands the whole frame range to one launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace fftfree {

enum class KernelKind { CooleyTukey, Stockham, External };
enum class KernelAccuracy { Default, HighPrecision, Reference };
enum class TransformMode { C2C = 0, R2C = 1, C2R = 2, R2R = 3 };

// A batch of transforms in device memory: element (i, b) at base[i*axis_stride +
// b*batch_stride]. The GPU kernels need axis_stride == 1 and batch_stride == axis_size
// (the layout fft_inplace_batched builds, eigen_fft.hpp:3159-3166).
struct AxisLayout {
  float2* base = nullptr;
  int64_t axis_size = 0;
  int64_t batch_size = 0;
  int64_t axis_stride = 1;
  int64_t batch_stride = 0;
  // Out-of-place input with the same layout (a GPU addition; nullptr = in place).
  const float2* source = nullptr;
  bool valid() const { return base != nullptr && axis_size > 0; }
};

class JobDispatcher {
 public:
  using Fn = std::function<void(size_t, size_t, int)>;
  virtual ~JobDispatcher() = default;
  virtual void parallel_for(size_t total, size_t chunk, const Fn& fn) = 0;
};

// The GPU grid is the worker pool: the whole range goes to one launch, worker 0.
class StreamDispatcher final : public JobDispatcher {
 public:
  explicit StreamDispatcher(hipStream_t s = nullptr) : stream_(s) {}
  void parallel_for(size_t total, size_t /*chunk*/, const Fn& fn) override {
    if (total) fn(0, total, 0);
  }
  hipStream_t stream() const { return stream_; }
  void set_stream(hipStream_t s) { stream_ = s; }

 private:
  hipStream_t stream_;
};

struct KernelContext {
  virtual ~KernelContext() = default;
};

struct PlanRuntimeConfig {
  int threads = 4;
  int lanes = 2;
  bool inverse = false;
  int radix = 0;
  std::vector<int> radix_pattern;
  int transform = 0;
  bool reduce_magnitude = false;
  bool store_polar = false;
  bool half_spectrum = false;
  bool allow_outer_parallel = true;
  bool allow_inner_parallel = false;
  int inner_threads = 0;
};

struct Plan;

struct KernelDescriptor {
  KernelKind kind = KernelKind::CooleyTukey;
  const char* name = nullptr;
  KernelAccuracy accuracy = KernelAccuracy::Default;
  bool realtime_safe = true;
  using CreateStateFn = std::unique_ptr<KernelContext> (*)(Plan&);
  using ExecuteFn = void (*)(const Plan&, const AxisLayout&, KernelContext*);
  using DestroyStateFn = void (*)(Plan&, KernelContext*);
  CreateStateFn create_state = nullptr;
  ExecuteFn execute_axis = nullptr;
  DestroyStateFn destroy_state = nullptr;
};

// Thrown by plan construction and execution; the C ABI catches it at the boundary.
struct HipFailure : std::runtime_error {
  using std::runtime_error::runtime_error;
};
void hip_check(hipError_t e, const char* what);

struct Plan {
  int N = 0;
  int lgN = 0;
  bool inverse = false;
  TransformMode transform_mode = TransformMode::C2C;
  bool reduce_magnitude = false;
  bool store_polar = false;
  bool half_spectrum = false;
  int radix = 0;                   // copied, a hint only (the reference never reads it)
  std::vector<int> radix_pattern;  // copied, a hint only
  int device = 0;
  float2* W = nullptr;             // device table exp(-2 pi i m / N), m < N, FP64-built
  float2* scratch = nullptr;       // device scratch for N > 4096 (pass-A output), grown on use
  size_t scratch_elems = 0;
  hipStream_t aux_stream = nullptr;  // second stream of the large-N chunk pipeline (default priority)
  hipEvent_t aux_fork = nullptr, aux_join = nullptr;
  int stft_variant = -1;           // kernel variant (autotuner), -1 = default
  int c2c_variant = -1;

  Plan(int n, bool inv, const PlanRuntimeConfig& cfg);
  ~Plan();
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;

  // Uploads the twiddle table on first use (lazy, so host-only queries need no GPU).
  void ensure_device();
  // The second stream and its fork/join events, created on first use on the plan's device.
  hipStream_t ensure_aux_stream() const;
  // Scratch of at least `elems` complex values (large-N two-pass transforms).
  float2* ensure_scratch(size_t elems) const;

  void use_kernel(KernelKind kind);
  const KernelDescriptor& kernel() const { return *kernel_desc_; }
  KernelContext* kernel_state() const { return kernel_state_.get(); }
  void set_dispatcher(JobDispatcher* d) { dispatcher_ = d ? d : &inline_dispatcher(); }
  JobDispatcher* dispatcher() const { return dispatcher_; }
  hipStream_t stream() const;

  static const KernelDescriptor& cooleytukey_kernel();
  static const KernelDescriptor& stockham_kernel();
  static const KernelDescriptor& external_kernel();
  static StreamDispatcher& inline_dispatcher();

 private:
  const KernelDescriptor* kernel_desc_ = nullptr;
  std::unique_ptr<KernelContext> kernel_state_;
  JobDispatcher* dispatcher_ = nullptr;
};

// Runs the plan's kernel over a device layout (eigen_fft.hpp:3104-3124).
void fft_apply_axis(const Plan& P, const AxisLayout& layout);

// Plans keyed by (N, direction, transform flags); one per key, created on demand.
class PlanCache {
 public:
  std::shared_ptr<Plan> get_plan(int N, bool inverse, const PlanRuntimeConfig& cfg);
  void clear();

 private:
  std::mutex mu_;
  std::map<std::string, std::shared_ptr<Plan>> plans_;
};

// FP64-built forward twiddles exp(-2 pi i m / N) rounded to float, m < N, with exact
// values at multiples of a quarter turn.
std::vector<float2> make_twiddles(int N);

}  // namespace fftfree

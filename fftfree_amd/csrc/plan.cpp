// plan.cpp -- GPU Plan, kernel descriptors, fft_apply_axis and the plan cache.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "fft_kernels.hpp"

namespace fftfree {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw HipFailure(std::string(what) + ": " + hipGetErrorString(e));
  }
}

std::vector<float2> make_twiddles(int N) {
  // exp(-2 pi i m / N): reduce m to a quadrant q and a remainder r (4m = qN + r) in
  // exact integer arithmetic, evaluate the small angle pi/2 * r/N in double (using the
  // complementary angle past an octant), then rotate by q quarter turns exactly.
  std::vector<float2> tw(static_cast<size_t>(N));
  const double half_pi = 1.57079632679489661923;
  for (int64_t m = 0; m < N; ++m) {
    const int64_t q = (4 * m) / N;
    const int64_t r = 4 * m - q * static_cast<int64_t>(N);
    double c, s;  // cos/sin of +2 pi m / N within the quadrant
    if (r == 0) {
      c = 1.0;
      s = 0.0;
    } else if (2 * r <= N) {
      const double th = half_pi * static_cast<double>(r) / static_cast<double>(N);
      c = std::cos(th);
      s = std::sin(th);
    } else {
      const double th = half_pi * static_cast<double>(N - r) / static_cast<double>(N);
      c = std::sin(th);
      s = std::cos(th);
    }
    double C, S;
    switch (q & 3) {
      case 0: C = c; S = s; break;
      case 1: C = -s; S = c; break;
      case 2: C = -c; S = -s; break;
      default: C = s; S = -c; break;
    }
    tw[static_cast<size_t>(m)] = make_float2(static_cast<float>(C), static_cast<float>(-S));
  }
  return tw;
}

Plan::Plan(int n, bool inv, const PlanRuntimeConfig& cfg) : N(n), inverse(inv) {
  // power-of-two check (eigen_fft.hpp:2228-2232); N == 0 is refused here too
  if (N <= 0) throw std::invalid_argument("N must be power of two");
  int t = N;
  while ((t & 1) == 0) {
    ++lgN;
    t >>= 1;
  }
  if ((1 << lgN) != N) throw std::invalid_argument("N must be power of two");
  switch (cfg.transform) {
    case 1: transform_mode = TransformMode::R2C; break;
    case 2: transform_mode = TransformMode::C2R; break;
    case 3: transform_mode = TransformMode::R2R; break;
    default: transform_mode = TransformMode::C2C; break;
  }
  reduce_magnitude = cfg.reduce_magnitude;
  store_polar = cfg.store_polar;
  half_spectrum = cfg.half_spectrum;
  radix = cfg.radix;
  radix_pattern = cfg.radix_pattern;
  dispatcher_ = &inline_dispatcher();
  kernel_desc_ = &cooleytukey_kernel();  // reference default (eigen_fft.hpp:2296)
}

Plan::~Plan() {
  if (W || scratch || aux_stream) (void)hipSetDevice(device);
  if (aux_stream) (void)hipStreamDestroy(aux_stream);
  if (aux_fork) (void)hipEventDestroy(aux_fork);
  if (aux_join) (void)hipEventDestroy(aux_join);
  if (W) (void)hipFree(W);
  if (scratch) (void)hipFree(scratch);
}

hipStream_t Plan::ensure_aux_stream() const {
  auto* self = const_cast<Plan*>(this);
  if (!aux_stream) {
    hip_check(hipEventCreateWithFlags(&self->aux_fork, hipEventDisableTiming), "hipEventCreate(fork)");
    hip_check(hipEventCreateWithFlags(&self->aux_join, hipEventDisableTiming), "hipEventCreate(join)");
    hip_check(hipStreamCreateWithFlags(&self->aux_stream, hipStreamNonBlocking), "hipStreamCreate(aux)");
  }
  return aux_stream;
}

float2* Plan::ensure_scratch(size_t elems) const {
  auto* self = const_cast<Plan*>(this);
  if (elems > scratch_elems) {
    if (scratch) (void)hipFree(self->scratch);
    self->scratch = nullptr;
    self->scratch_elems = 0;
    hip_check(hipMalloc(reinterpret_cast<void**>(&self->scratch), elems * sizeof(float2)), "hipMalloc(scratch)");
    self->scratch_elems = elems;
  }
  return scratch;
}

void Plan::ensure_device() {
  if (W) return;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw HipFailure("no HIP device available");
  hip_check(hipGetDevice(&device), "hipGetDevice");
  const std::vector<float2> host = make_twiddles(N);
  float2* d = nullptr;
  hip_check(hipMalloc(reinterpret_cast<void**>(&d), sizeof(float2) * host.size()), "hipMalloc(twiddles)");
  const hipError_t e = hipMemcpy(d, host.data(), sizeof(float2) * host.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    hip_check(e, "hipMemcpy(twiddles)");
  }
  W = d;
}

hipStream_t Plan::stream() const {
  auto* sd = dynamic_cast<StreamDispatcher*>(dispatcher_);
  return sd ? sd->stream() : nullptr;
}

StreamDispatcher& Plan::inline_dispatcher() {
  static StreamDispatcher d(nullptr);
  return d;
}

void Plan::use_kernel(KernelKind kind) {
  const KernelDescriptor* d = nullptr;
  switch (kind) {
    case KernelKind::CooleyTukey: d = &cooleytukey_kernel(); break;
    case KernelKind::Stockham: d = &stockham_kernel(); break;
    case KernelKind::External: d = &external_kernel(); break;
  }
  if (!d || !d->execute_axis) throw std::invalid_argument("Unsupported KernelKind.");
  if (kernel_desc_ == d) return;
  if (kernel_desc_ && kernel_desc_->destroy_state) kernel_desc_->destroy_state(*this, kernel_state_.get());
  kernel_state_.reset();
  kernel_desc_ = d;
  if (d->create_state) kernel_state_ = d->create_state(*this);
}

namespace {

// Both built-in descriptors run the same gfx950 Stockham engine; they differ in the
// hooks the reference runs around them: Cooley-Tukey applies the pre-transform hook
// (eigen_fft.hpp:2671), Stockham does not (eigen_fft.hpp:3227-3623). Both apply the
// 1/N inverse scale and the post-transform hook.
C2CArgs make_c2c_args(const Plan& P, bool cooley_tukey) {
  C2CArgs a;
  a.tw = P.W;
  a.n = P.N;
  if (cooley_tukey) {
    if (P.transform_mode == TransformMode::R2C) a.pre_mode = 1;
    if (P.transform_mode == TransformMode::C2R && P.half_spectrum) {
      a.pre_mode = 2;
      a.pre_polar = P.store_polar ? 1 : 0;
    }
  }
  const bool r2c_half = P.transform_mode == TransformMode::R2C && P.half_spectrum;
  a.post_half = r2c_half ? 1 : 0;
  a.post_mode = P.store_polar ? 3 : P.reduce_magnitude ? 2 : (P.transform_mode == TransformMode::C2R ? 1 : 0);
  a.post_upper = r2c_half ? P.N / 2 + 1 : P.N;
  a.scale = P.inverse ? 1.0f / static_cast<float>(P.N) : 1.0f;
  return a;
}

void gpu_execute_axis(const Plan& P, const AxisLayout& layout, bool cooley_tukey) {
  if (layout.axis_stride != 1 || layout.batch_stride != layout.axis_size)
    throw std::invalid_argument("GPU kernels need contiguous transforms (axis_stride 1, batch_stride N)");
  if (!P.W) throw std::logic_error("plan twiddles not on the device");
  C2CArgs a = make_c2c_args(P, cooley_tukey);
  const bool large = P.N > c2c_max_lds_n();
  if (large && !large_supported(static_cast<size_t>(P.N))) throw std::invalid_argument("transform size not supported");
  P.dispatcher()->parallel_for(static_cast<size_t>(layout.batch_size), 1, [&](size_t b, size_t e, int) {
    const float2* src = layout.source ? layout.source : layout.base;
    if (!large) {
      a.in = src + b * static_cast<size_t>(P.N);
      a.out = layout.base + b * static_cast<size_t>(P.N);
      a.batch = e - b;
      hip_check(launch_c2c(a, P.inverse ? 1 : 0, P.c2c_variant, P.stream()), "launch_c2c");
      return;
    }
    // Two-pass large transform in chunks of transforms. A chunk's pass-A output (the
    // scratch) is read back by its pass B straight away, so a chunk's scratch is kept
    // small enough to stay in the caches (about 32 MiB); chunks alternate between the
    // plan's stream and a second one, each with its own scratch half, so one chunk's
    // pass A overlaps the other's pass B and neither waits for a launch's tail.
    const size_t chunk_elems = size_t(4) << 20;
    const size_t n = static_cast<size_t>(P.N);
    const size_t per = std::max<size_t>(1, std::min<size_t>(65535, chunk_elems / n));
    const size_t chunks = (e - b + per - 1) / per;
    const bool two = chunks > 1;
    float2* tmp = P.ensure_scratch((two ? 2 : 1) * per * n);
    hipStream_t s0 = P.stream();
    hipStream_t s1 = two ? P.ensure_aux_stream() : s0;
    if (two) {
      hip_check(hipEventRecord(P.aux_fork, s0), "hipEventRecord(fork)");
      hip_check(hipStreamWaitEvent(s1, P.aux_fork, 0), "hipStreamWaitEvent(fork)");
    }
    for (size_t i = 0; i < chunks; ++i) {
      const size_t c = b + i * per;
      const size_t cnt = std::min(per, e - c);
      a.in = src + c * n;
      a.out = layout.base + c * n;
      a.batch = cnt;
      hip_check(launch_c2c_large(a, P.inverse ? 1 : 0, P.c2c_variant, tmp + (i & 1) * per * n, (i & 1) ? s1 : s0),
                "launch_c2c_large");
    }
    if (two) {
      hip_check(hipEventRecord(P.aux_join, s1), "hipEventRecord(join)");
      hip_check(hipStreamWaitEvent(s0, P.aux_join, 0), "hipStreamWaitEvent(join)");
    }
  });
}

void ct_execute(const Plan& P, const AxisLayout& l, KernelContext*) { gpu_execute_axis(P, l, true); }
void stockham_execute(const Plan& P, const AxisLayout& l, KernelContext*) { gpu_execute_axis(P, l, false); }
void external_execute(const Plan&, const AxisLayout&, KernelContext*) {
  throw std::logic_error("External FFT kernel requested but no implementation supplied.");
}

}  // namespace

const KernelDescriptor& Plan::cooleytukey_kernel() {
  static const KernelDescriptor d{KernelKind::CooleyTukey, "cooleytukey-gfx950", KernelAccuracy::Default, true,
                                  nullptr, &ct_execute, nullptr};
  return d;
}
const KernelDescriptor& Plan::stockham_kernel() {
  static const KernelDescriptor d{KernelKind::Stockham, "stockham-autosort-gfx950", KernelAccuracy::Default, true,
                                  nullptr, &stockham_execute, nullptr};
  return d;
}
const KernelDescriptor& Plan::external_kernel() {
  static const KernelDescriptor d{KernelKind::External, "external-provider", KernelAccuracy::HighPrecision, false,
                                  nullptr, &external_execute, nullptr};
  return d;
}

void fft_apply_axis(const Plan& P, const AxisLayout& layout) {
  if (!layout.valid()) throw std::invalid_argument("AxisLayout must reference valid data.");
  if (layout.axis_size != P.N) throw std::invalid_argument("AxisLayout axis_size must match Plan::N.");
  P.kernel().execute_axis(P, layout, P.kernel_state());
}

std::shared_ptr<Plan> PlanCache::get_plan(int N, bool inverse, const PlanRuntimeConfig& cfg) {
  char key[128];
  std::snprintf(key, sizeof(key), "%d/%d/%d/%d/%d/%d", N, inverse ? 1 : 0, cfg.transform, cfg.reduce_magnitude ? 1 : 0,
                cfg.store_polar ? 1 : 0, cfg.half_spectrum ? 1 : 0);
  std::lock_guard<std::mutex> lk(mu_);
  auto it = plans_.find(key);
  if (it != plans_.end()) return it->second;
  auto p = std::make_shared<Plan>(N, inverse, cfg);
  plans_.emplace(key, p);
  return p;
}

void PlanCache::clear() {
  std::lock_guard<std::mutex> lk(mu_);
  plans_.clear();
}

}  // namespace fftfree

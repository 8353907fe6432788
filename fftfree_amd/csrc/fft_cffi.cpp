// fft_cffi.cpp -- the C ABI (include/fft_cffi.h) over the gfx950 kernels.
//
// FftContext mirrors the reference context (fft_cffi.cpp:40-82): the init arguments,
// windows built to N taps, a plan cache. It adds a HIP stream, device copies of the
// windows and reusable device staging buffers. Host-pointer entry points stage through
// device memory; the *_device entry points launch straight on caller buffers.
//
// Errors never cross the C boundary (reference: catch(...) at fft_cffi.cpp:1031-1033,
// 1955-1973): they are reported on stderr and turned into NULL / 0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fft_cffi.h"
#include "fft_kernels.hpp"
#include "host_pipeline.hpp"
#include "mt19937.hpp"
#include "plan.hpp"

namespace fftfree {
namespace {

// Reusable device allocation that only grows.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  int device = 0;
  void reserve(size_t want) {
    if (want <= bytes) return;
    release();
    if (want == 0) return;
    hip_check(hipMalloc(&ptr, want), "hipMalloc(staging)");
    bytes = want;
    (void)hipGetDevice(&device);
  }
  void release() {
    if (ptr) {
      (void)hipSetDevice(device);
      (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
  ~DeviceBuffer() { release(); }
};

struct FftContext {
  int N = 0;
  bool inverse = false;
  int kernel = 0;  // C-level code: 2 = Stockham, anything else = Cooley-Tukey
  int radix = 0;
  int pad_mode = 0;
  std::vector<int> radix_pattern;
  int window = 0;
  int hop = 0;
  int transform = 0;
  int reduce_magnitude = 0;
  int store_polar = 0;
  int half_spectrum = 0;
  int allow_outer_parallel = 1;
  int allow_inner_parallel = 0;
  int inner_threads = 0;
  bool has_pool = false;  // the reference creates a WorkerPool here (fft_cffi.cpp:1928-1931)
  int analysis_win_kind = 0;
  int synthesis_win_kind = 0;
  float win_param1 = 0.0f, win_param2 = 0.0f;
  int win_norm_policy = 0;
  int cola_mode = 0;
  std::vector<float> analysis_win;
  std::vector<float> synthesis_win;
  bool windows_enabled = false;
  int apply_ola = 0;
  // streaming state (fft_cffi.cpp:1342-1742): PCM not yet consumed, and the inverse
  // direction's overlap-add accumulator and queue of finished samples
  std::vector<float> stream_pcm;
  size_t stream_offset = 0;
  std::vector<float> stream_queue;
  size_t stream_queue_offset = 0;
  std::vector<float> stream_accum, stream_norm;
  PlanRuntimeConfig cfg;
  PlanCache cache;
  std::shared_ptr<Plan> plan;

  // device side (created on first use)
  bool gpu_ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  StreamDispatcher dispatcher;
  DeviceBuffer d_awin, d_swin, d_pcm, d_planes, d_cplx;
  DeviceBuffer d_sink;  // N floats that kernels write skipped output planes to
  DeviceBuffer d_halo;  // adjoint run-boundary partial sums
  DeviceBuffer d_gl;    // Griffin-Lim working set (on the inverse context)
  uint32_t* h_words = nullptr;       // Griffin-Lim start words (pinned host memory)
  // the start words now in d_gl: their seed and count (0: none); they depend on nothing else
  unsigned gl_words_seed = 0;
  size_t gl_words_count = 0;
  size_t h_words_cap = 0;
  hipEvent_t words_copied = nullptr;  // the last copy out of h_words
  bool windows_dirty = true;
  // host-pointer pipeline (created on first use): upload and download streams, pinned
  // bounce slots, one event per slot and stage, and the host copy threads
  hipStream_t up_stream = nullptr, down_stream = nullptr;
  PinnedBuffer h_in, h_out;
  hipEvent_t ev_in[2] = {nullptr, nullptr};
  hipEvent_t ev_kernel[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_out[3] = {nullptr, nullptr, nullptr};
  std::unique_ptr<CopyPool> copy_pool;

  ~FftContext() {
    if (up_stream || down_stream) {
      (void)hipSetDevice(device);
      if (up_stream) (void)hipStreamSynchronize(up_stream);
      if (down_stream) (void)hipStreamSynchronize(down_stream);
      if (stream) (void)hipStreamSynchronize(stream);
      for (hipEvent_t e : ev_in)
        if (e) (void)hipEventDestroy(e);
      for (hipEvent_t e : ev_kernel)
        if (e) (void)hipEventDestroy(e);
      for (hipEvent_t e : ev_out)
        if (e) (void)hipEventDestroy(e);
      if (up_stream) (void)hipStreamDestroy(up_stream);
      if (down_stream) (void)hipStreamDestroy(down_stream);
    }
    h_in.release();
    h_out.release();
    copy_pool.reset();
    d_awin.release();
    d_swin.release();
    d_pcm.release();
    d_planes.release();
    d_cplx.release();
    d_sink.release();
    d_halo.release();
    d_gl.release();
    if (words_copied) {
      (void)hipEventSynchronize(words_copied);
      (void)hipEventDestroy(words_copied);
    }
    if (h_words) (void)hipHostFree(h_words);
    plan.reset();
    cache.clear();
    if (stream) {
      (void)hipSetDevice(device);
      (void)hipStreamDestroy(stream);
    }
  }
};

bool verbose() {
  static const bool v = [] {
    const char* e = std::getenv("FFTFREE_VERBOSE");
    return e && *e && *e != '0';
  }();
  return v;
}

void report(const char* where, const char* what) { std::fprintf(stderr, "[fftfree-amd] %s: %s\n", where, what); }

bool is_pow2(size_t v) { return v != 0 && (v & (v - 1)) == 0; }
size_t next_pow2(size_t v) {
  size_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// build_window / normalize_window (fft_cffi.cpp:163-208), same double formulas.
std::vector<float> build_window(int kind, int W, float p1, float /*p2*/) {
  const double PI = 3.141592653589793238462643383279502884;
  std::vector<float> w(static_cast<size_t>(std::max(W, 0)), 1.0f);
  switch (kind) {
    case FFT_WINDOW_HANN:
    case FFT_WINDOW_KAISER:  // the reference uses Hann as its Kaiser placeholder
      for (int n = 0; n < W; ++n) w[n] = float(0.5 * (1.0 - std::cos(2.0 * PI * n / (W - 1))));
      break;
    case FFT_WINDOW_HAMMING:
      for (int n = 0; n < W; ++n) w[n] = float(0.54 - 0.46 * std::cos(2.0 * PI * n / (W - 1)));
      break;
    case FFT_WINDOW_BLACKMAN:
      for (int n = 0; n < W; ++n)
        w[n] = float(0.42 - 0.5 * std::cos(2.0 * PI * n / (W - 1)) + 0.08 * std::cos(4.0 * PI * n / (W - 1)));
      break;
    case FFT_WINDOW_TUKEY: {
      const double alpha = (p1 == 0.0f ? 0.5 : p1);
      const int L = std::max(1, W - 1);
      for (int n = 0; n < W; ++n) {
        const double x = double(n) / double(L);
        if (x < alpha / 2) w[n] = float(0.5 * (1 + std::cos(PI * (2 * x / alpha - 1))));
        else if (x <= 1 - alpha / 2) w[n] = 1.0f;
        else w[n] = float(0.5 * (1 + std::cos(PI * (2 * x / alpha - 2 / alpha + 1))));
      }
      break;
    }
    default:
      break;
  }
  return w;
}

void normalize_window(std::vector<float>& w, int policy) {
  if (w.empty()) return;
  if (policy == FFT_WINDOW_NORM_L2) {
    double s2 = 0.0;
    for (float v : w) s2 += double(v) * double(v);
    s2 = std::sqrt(std::max(1e-30, s2));
    for (auto& v : w) v = float(v / s2);
  } else if (policy == FFT_WINDOW_NORM_AREA) {
    double s = 0.0;
    for (float v : w) s += double(v);
    if (std::abs(s) > 1e-30)
      for (auto& v : w) v = float(v / s);
  }
}

int parse_kernel_token(std::string t) {
  auto b = t.begin();
  while (b != t.end() && std::isspace((unsigned char)*b)) ++b;
  auto e = t.end();
  while (e != b && std::isspace((unsigned char)*(e - 1))) --e;
  t = std::string(b, e);
  for (auto& c : t) c = (char)std::tolower((unsigned char)c);
  if (t.empty() || t == "0" || t == "auto" || t == "default") return 0;
  if (t == "cooleytukey" || t == "cooley" || t == "cooley-tukey" || t == "cooley_tukey" || t == "ct")
    return FFT_KERNEL_COOLEYTUKEY;
  if (t == "stockham" || t == "stock" || t == "sh") return FFT_KERNEL_STOCKHAM;
  if (t == "external" || t == "ext" || t == "custom") return FFT_KERNEL_EXTERNAL;
  char* end = nullptr;
  const long v = std::strtol(t.c_str(), &end, 10);
  if (end && *end == '\0' && (v == 0 || v == 1 || v == 2)) return (int)v;
  return 0;
}

// FFTFREE_DEFAULT_KERNEL (fft_cffi.cpp:127-159), consulted by fft_init_ex only.
int env_default_kernel() {
  static const int k = [] {
    const char* e = std::getenv("FFTFREE_DEFAULT_KERNEL");
    return e ? parse_kernel_token(e) : 0;
  }();
  return k;
}

bool use_stockham(const FftContext* c) { return c->kernel == 2; }

void bind_plan(FftContext* c) {
  c->plan = c->cache.get_plan(c->N, c->inverse, c->cfg);
  c->plan->use_kernel(use_stockham(c) ? KernelKind::Stockham : KernelKind::CooleyTukey);
}

void ensure_gpu(FftContext* c) {
  if (!c->gpu_ready) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw HipFailure("no HIP device available");
    hip_check(hipGetDevice(&c->device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    c->dispatcher.set_stream(c->stream);
    c->gpu_ready = true;
  }
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  c->plan->ensure_device();
  c->plan->set_dispatcher(&c->dispatcher);
  c->d_sink.reserve(std::max<size_t>(64, (size_t)c->N) * sizeof(float));
  if (c->windows_dirty) {
    c->d_awin.reserve(std::max<size_t>(1, c->analysis_win.size()) * sizeof(float));
    c->d_swin.reserve(std::max<size_t>(1, c->synthesis_win.size()) * sizeof(float));
    if (!c->analysis_win.empty())
      hip_check(hipMemcpy(c->d_awin.ptr, c->analysis_win.data(), c->analysis_win.size() * sizeof(float),
                          hipMemcpyHostToDevice),
                "upload analysis window");
    if (!c->synthesis_win.empty())
      hip_check(hipMemcpy(c->d_swin.ptr, c->synthesis_win.data(), c->synthesis_win.size() * sizeof(float),
                          hipMemcpyHostToDevice),
                "upload synthesis window");
    c->windows_dirty = false;
  }
}

// Frame count of fft_execute_batched (fft_cffi.cpp:713-750); 0 = refused.
size_t frame_count(const FftContext* c, size_t pcm_len, int pad_mode, size_t max_frames) {
  const int N = c->N;
  if (N <= 0 || pcm_len == 0) return 0;
  const int W = c->window ? c->window : N;
  const int H = c->hop ? c->hop : W;
  if (W <= 0 || H <= 0 || W > N) return 0;
  const int pad = (pad_mode == 0) ? c->pad_mode : pad_mode;
  size_t frames;
  if (pcm_len < (size_t)W) {
    if (pad != 1) return 0;
    frames = 1;
  } else {
    frames = 1 + (pcm_len - (size_t)W) / (size_t)H;
    const size_t last = (frames - 1) * (size_t)H;
    const size_t remaining = pcm_len > last ? pcm_len - last : 0;
    if (remaining < (size_t)W) {
      if (pad == 1) {
      } else if (pad == 2) {
        frames -= 1;
      } else {
        return 0;
      }
    }
  }
  if (frames == 0) return 0;
  if (max_frames != 0 && frames > max_frames) frames = max_frames;
  return frames;
}

size_t bins_of(const FftContext* c) { return c->half_spectrum ? (size_t)(c->N / 2 + 1) : (size_t)c->N; }

// The fused kernel assumes real input. Only the Cooley-Tukey C2R half-spectrum hook
// (Hermitian rebuild of the framed input, eigen_fft.hpp:2501-2527) breaks that.
bool fused_ok(const FftContext* c) {
  if (c->N > stft_max_lds_n()) return false;
  return !(!use_stockham(c) && c->transform == FFT_TRANSFORM_C2R && c->half_spectrum);
}

StftArgs stft_args(FftContext* c, const float* d_pcm, size_t pcm_len) {
  StftArgs a;
  a.pcm = d_pcm;
  a.pcm_len = pcm_len;
  a.window = c->window ? c->window : c->N;
  a.hop = c->hop ? c->hop : a.window;
  a.win = (c->windows_enabled && !c->analysis_win.empty()) ? c->d_awin.as<float>() : nullptr;
  a.tw = c->plan->W;
  a.n = c->N;
  a.bins = (int)bins_of(c);
  a.out_mode = c->store_polar ? 3 : c->reduce_magnitude ? 2 : (c->transform == FFT_TRANSFORM_C2R ? 1 : 0);
  a.scale = c->inverse ? 1.0f / (float)c->N : 1.0f;
  a.sink = c->d_sink.as<float>();
  return a;
}

// Frames [f0, f0+nf) of the STFT into planes (already offset to frame f0).
// `variant`: the kernel variant when the context has none set (-1: the default choice).
void run_stft_frames(FftContext* c, const StftArgs& base, size_t f0, size_t nf, float* re, float* im, float* mag,
                     hipStream_t s, int variant = -1) {
  StftArgs a = base;
  a.frames = nf;
  a.frame_offset = f0;
  a.o_re = re;
  a.o_im = im;
  a.o_mag = mag;
  if (fused_ok(c)) {
    hip_check(launch_stft(a, c->inverse ? 1 : 0, c->plan->stft_variant >= 0 ? c->plan->stft_variant : variant, s),
              "launch_stft");
    return;
  }
  // generic path: frame -> plan kernel (hooks) -> planes, in chunks of <= 256 MiB
  const size_t per_frame = (size_t)c->N * sizeof(float2);
  const size_t chunk = std::max<size_t>(1, std::min<size_t>(nf, (size_t(256) << 20) / per_frame));
  c->d_cplx.reserve(chunk * per_frame);
  for (size_t b = 0; b < nf; b += chunk) {
    const size_t cnt = std::min(chunk, nf - b);
    StftArgs fa = a;
    fa.frames = cnt;
    fa.frame_offset = f0 + b;
    hip_check(launch_frame_complex(fa, c->d_cplx.as<float2>(), s), "launch_frame_complex");
    AxisLayout l;
    l.base = c->d_cplx.as<float2>();
    l.axis_size = c->N;
    l.batch_size = (int64_t)cnt;
    l.axis_stride = 1;
    l.batch_stride = c->N;
    c->dispatcher.set_stream(s);
    fft_apply_axis(*c->plan, l);
    c->dispatcher.set_stream(c->stream);
    const size_t bins = bins_of(c);
    hip_check(launch_planes(c->d_cplx.as<float2>(), cnt, c->N, (int)bins, c->store_polar, re ? re + b * bins : nullptr,
                            im ? im + b * bins : nullptr, mag ? mag + b * bins : nullptr, s),
              "launch_planes");
  }
}

// After an error part-way: wait for every stage so no copy is left writing into the
// caller's or the context's buffers.
void drain_pipeline(FftContext* c) {
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->up_stream) (void)hipStreamSynchronize(c->up_stream);
  if (c->down_stream) (void)hipStreamSynchronize(c->down_stream);
}

// Streams, events and copy threads of the host-pointer pipeline.
void ensure_pipeline(FftContext* c) {
  if (c->up_stream) return;
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  for (hipEvent_t& e : c->ev_in) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  for (hipEvent_t& e : c->ev_kernel) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  for (hipEvent_t& e : c->ev_out) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipStreamCreateWithFlags(&c->down_stream, hipStreamNonBlocking), "hipStreamCreate(down)");
  hip_check(hipStreamCreateWithFlags(&c->up_stream, hipStreamNonBlocking), "hipStreamCreate(up)");
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  c->copy_pool = std::make_unique<CopyPool>((int)std::min(7u, hw > 1 ? hw - 1 : 0u));
}

// fft_execute_batched on host pointers (fft_cffi.cpp:713-804): `frames` frames of the STFT
// of pcm[0, pcm_len) into the three host planes, in chunks of about 24 MiB of planes.
// Chunk i: its new samples are copied into pinned input slot i % 2 and uploaded; its
// kernel writes device plane slot i % 3; the planes are downloaded into pinned slot i % 3;
// two chunks later they are copied into the caller's planes. Returns the frames whose
// planes reached the caller (all of them unless an error is thrown part-way).
size_t stft_host_pipelined(FftContext* c, const float* pcm, size_t pcm_len, size_t frames, float* o_re, float* o_im,
                           float* o_mag, size_t* produced) {
  constexpr size_t K = 3;
  ensure_pipeline(c);
  const size_t bins = bins_of(c);
  const size_t W = (size_t)(c->window ? c->window : c->N);
  const size_t H = (size_t)(c->hop ? c->hop : (int)W);
  const size_t plane_bytes = bins * sizeof(float);
  // Small calls (at most 2 MiB of planes: a streaming push of a few hops, say): one
  // upload, the launch over all frames (the pipeline's single chunk) and one download on
  // the context's stream, the host copies on this thread. No copy threads and no events
  // between streams, which cost more than such a call's copies.
  if (frames * 3 * plane_bytes <= (size_t(2) << 20)) {
    const size_t need = std::min(pcm_len, (frames - 1) * H + W);
    const size_t stride = (frames * bins + 63) / 64 * 64;
    c->d_pcm.reserve(pcm_len * sizeof(float));
    c->d_planes.reserve(3 * stride * sizeof(float));
    c->h_in.reserve(need * sizeof(float));
    c->h_out.reserve(3 * stride * sizeof(float));
    std::memcpy(c->h_in.as<float>(), pcm, need * sizeof(float));
    hip_check(hipMemcpyAsync(c->d_pcm.as<float>(), c->h_in.as<float>(), need * sizeof(float), hipMemcpyHostToDevice,
                             c->stream),
              "H2D pcm");
    float* d = c->d_planes.as<float>();
    run_stft_frames(c, stft_args(c, c->d_pcm.as<float>(), pcm_len), 0, frames, d, d + stride, d + 2 * stride,
                    c->stream);
    hip_check(hipMemcpyAsync(c->h_out.as<float>(), d, 3 * stride * sizeof(float), hipMemcpyDeviceToHost, c->stream),
              "D2H planes");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    const float* h = c->h_out.as<float>();
    std::memcpy(o_re, h, frames * plane_bytes);
    std::memcpy(o_im, h + stride, frames * plane_bytes);
    std::memcpy(o_mag, h + 2 * stride, frames * plane_bytes);
    *produced = frames;
    return frames;
  }
  // A multiple of every kernel variant's tile (4 to 32 frames), so chunk edges fall on
  // tile edges of the single launch over all frames and the planes are bit-identical.
  constexpr size_t kTileMultiple = 192;
  size_t chunk = (size_t(24) << 20) / (3 * plane_bytes);
  chunk = std::max(kTileMultiple, chunk / kTileMultiple * kTileMultiple);
  chunk = std::min(chunk, frames);
  const size_t chunks = (frames + chunk - 1) / chunk;
  // samples [0, end(i)) are needed by chunks 0..i
  auto sample_end = [&](size_t i) {
    const size_t last = std::min(frames, (i + 1) * chunk) - 1;
    return std::min(pcm_len, last * H + W);
  };
  size_t max_piece = 0;
  for (size_t i = 0, prev = 0; i < chunks; ++i) {
    const size_t e = std::max(prev, sample_end(i));
    max_piece = std::max(max_piece, e - prev);
    prev = e;
  }
  // each plane of a slot starts on a 256-byte boundary, as the kernels' vector stores want
  const size_t plane_stride = (chunk * bins + 63) / 64 * 64;
  const size_t slot_floats = 3 * plane_stride;
  c->d_pcm.reserve(pcm_len * sizeof(float));
  c->d_planes.reserve(K * slot_floats * sizeof(float));
  c->h_in.reserve(std::max<size_t>(1, 2 * max_piece) * sizeof(float));
  c->h_out.reserve(K * slot_floats * sizeof(float));
  const StftArgs base = stft_args(c, c->d_pcm.as<float>(), pcm_len);
  float* const d_slots = c->d_planes.as<float>();
  float* const h_slots = c->h_out.as<float>();

  auto copy_out = [&](size_t j) {
    hip_check(hipEventSynchronize(c->ev_out[j % K]), "download");
    const size_t f0 = j * chunk, nf = std::min(chunk, frames - f0);
    const float* src = h_slots + (j % K) * slot_floats;
    const CopyPool::Copy copies[3] = {{o_re + f0 * bins, src, nf * plane_bytes},
                                      {o_im + f0 * bins, src + plane_stride, nf * plane_bytes},
                                      {o_mag + f0 * bins, src + 2 * plane_stride, nf * plane_bytes}};
    c->copy_pool->run(copies, 3);
    *produced = f0 + nf;
  };
  size_t staged = 0;  // samples uploaded so far
  bool in_used[2] = {false, false};
  for (size_t i = 0; i < chunks; ++i) {
    const size_t f0 = i * chunk, nf = std::min(chunk, frames - f0);
    const size_t end = std::max(staged, sample_end(i));
    if (end > staged) {
      const size_t in_slot = i % 2;
      // the upload that last read this pinned slot (chunk i - 2) has finished
      if (in_used[in_slot]) hip_check(hipEventSynchronize(c->ev_in[in_slot]), "upload");
      in_used[in_slot] = true;
      float* stage = c->h_in.as<float>() + in_slot * max_piece;
      const CopyPool::Copy in{stage, pcm + staged, (end - staged) * sizeof(float)};
      c->copy_pool->run(&in, 1);
      hip_check(hipMemcpyAsync(c->d_pcm.as<float>() + staged, stage, (end - staged) * sizeof(float),
                               hipMemcpyHostToDevice, c->up_stream),
                "H2D pcm");
      hip_check(hipEventRecord(c->ev_in[in_slot], c->up_stream), "hipEventRecord");
      hip_check(hipStreamWaitEvent(c->stream, c->ev_in[in_slot], 0), "hipStreamWaitEvent");
      staged = end;
    }
    const size_t k = i % K;
    float* d_re = d_slots + k * slot_floats;
    // the download that last read this device slot (chunk i - K) has finished
    if (i >= K) hip_check(hipStreamWaitEvent(c->stream, c->ev_out[k], 0), "hipStreamWaitEvent");
    run_stft_frames(c, base, f0, nf, d_re, d_re + plane_stride, d_re + 2 * plane_stride, c->stream);
    hip_check(hipEventRecord(c->ev_kernel[k], c->stream), "hipEventRecord");
    // pinned slot k was copied out (chunk i - K) before this iteration began
    hip_check(hipStreamWaitEvent(c->down_stream, c->ev_kernel[k], 0), "hipStreamWaitEvent");
    float* h_re = h_slots + k * slot_floats;
    if (nf == chunk) {
      hip_check(hipMemcpyAsync(h_re, d_re, slot_floats * sizeof(float), hipMemcpyDeviceToHost, c->down_stream),
                "D2H planes");
    } else {
      for (int p = 0; p < 3; ++p)
        hip_check(hipMemcpyAsync(h_re + p * plane_stride, d_re + p * plane_stride, nf * plane_bytes,
                                 hipMemcpyDeviceToHost, c->down_stream),
                  "D2H planes");
    }
    hip_check(hipEventRecord(c->ev_out[k], c->down_stream), "hipEventRecord");
    if (i + 1 >= K) copy_out(i + 1 - K);
  }
  for (size_t j = chunks >= K - 1 ? chunks - (K - 1) : 0; j < chunks; ++j) copy_out(j);
  return *produced;
}

FftContext* make_context(size_t n, int threads, int lanes, int inverse, int kernel, int radix,
                         const int* radix_pattern, size_t radix_pattern_len, int pad_mode, int window, int hop) {
  auto c = std::make_unique<FftContext>();
  c->N = (int)n;
  c->inverse = inverse != 0;
  c->kernel = kernel;
  c->radix = radix;
  c->pad_mode = pad_mode;
  c->window = window;
  c->hop = hop;
  if (radix_pattern && radix_pattern_len > 0) c->radix_pattern.assign(radix_pattern, radix_pattern + radix_pattern_len);
  c->cfg.threads = threads > 0 ? threads : 1;
  c->cfg.lanes = lanes > 0 ? lanes : 1;
  c->cfg.inverse = c->inverse;
  c->cfg.radix = radix;
  c->cfg.radix_pattern = c->radix_pattern;
  return c.release();
}

// Which kernel table fft_set_variant / fft_variant_* address: the fused STFT tiles for
// real-input and real-output contexts, the complex-transform kernels for C2C contexts
// (whose batched complex entries run those) and for contexts the fused STFT cannot take.
bool stft_variants(const FftContext* c) { return fused_ok(c) && c->transform != FFT_TRANSFORM_C2C; }

int variant_count(const FftContext* c) {
  if (stft_variants(c)) return stft_variant_count(c->N);
  if (c->N > c2c_max_lds_n()) return large_variant_count(c->N);
  return c2c_variant_count(c->N);
}

const char* variant_desc(const FftContext* c, int idx, char* buf, size_t len) {
  if (!stft_variants(c) && c->N > c2c_max_lds_n()) return large_variant_desc(c->N, idx, buf, len) ? buf : nullptr;
  const VariantInfo* v = stft_variants(c) ? stft_variant(c->N, idx) : c2c_variant(c->N, idx);
  if (!v) return nullptr;
  std::snprintf(buf, len, "%s fpw=%d", v->radices, v->fpw);
  return buf;
}

}  // namespace
}  // namespace fftfree

using fftfree::FftContext;
using fftfree::HipFailure;

namespace fftfree {
namespace {

// A ragged batch: one context per distinct (window, hop) supplies windows, twiddles and
// the plane sink; every request becomes a StftArgs block in device memory. Per window size
// N the plan holds an engine (a tile shape, or the line-aligned engine) and the blocks per
// work item; the work items of each N are sorted by cost and split into 8 cost-balanced
// XCD chunks, and the graph holds one persistent launch per N. The autotuner re-times the
// choices on the batch itself.
struct RaggedPlan {
  std::vector<fft_ragged_request> reqs;
  std::vector<void*> ctxs;           // FftContext handles
  std::vector<int> req_ctx;
  std::vector<StftArgs> args;        // as the per-request launches take them
  int mode = 0;
  std::map<int, int> cfg_of_n;       // N -> engine (ragged_cfg id)
  std::map<int, int> tiles_of_n;     // N -> blocks per work item
  std::map<int, std::vector<size_t>> reqs_of_n;
  struct Launch {
    int cfg;
    size_t chunk_off;  // its 9 chunk bounds (absolute item indices)
    int grid;
  };
  std::vector<Launch> launches;      // of the staged work lists
  std::map<int, int> grid_of_cfg;
  DeviceBuffer d_args, d_items, d_chunks;
  size_t n_items = 0;
  int device = 0;
  hipStream_t stream = nullptr;      // private stream for capture and tuning
  // capture-only branches: the graph runs the launches of different N side by side
  static constexpr int kBranches = 4;
  hipStream_t branch[kBranches] = {};
  std::vector<hipEvent_t> marks;     // fork and joins of the capture
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~RaggedPlan() {
    (void)hipSetDevice(device);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : marks) (void)hipEventDestroy(e);
    for (hipStream_t b : branch)
      if (b) (void)hipStreamDestroy(b);
    if (stream) (void)hipStreamDestroy(stream);
    d_args.release();
    d_items.release();
    d_chunks.release();
    for (void* c : ctxs) fft_free(c);
  }

  int grid(int cfg) {
    auto it = grid_of_cfg.find(cfg);
    if (it != grid_of_cfg.end()) return it->second;
    int g = 0;
    hip_check(ragged_grid(mode, cfg, &g), "ragged_grid");
    g = std::max(g, 8);
    grid_of_cfg[cfg] = g;
    return g;
  }

  // The line-aligned engine `cfg` takes every request of N whole (each qualifies and its
  // frames after the last tile fit the rest block).
  bool lines_ok(int n, int cfg) const {
    for (size_t r : reqs_of_n.at(n)) {
      size_t lead = 0, blocks = 0, take = 0;
      if (!ragged_lines_layout(cfg, args[r], &lead, &blocks, &take) || take != args[r].frames) return false;
    }
    return true;
  }

  // Blocks of request r under engine cfg, and the frames (cost) of one block.
  size_t blocks_of(size_t r, int cfg, size_t* lead) const {
    *lead = 0;
    if (ragged_cfg_lines(cfg)) {
      size_t blocks = 0, take = 0;
      ragged_lines_layout(cfg, args[r], lead, &blocks, &take);
      return blocks;
    }
    const int fpw = ragged_cfg(cfg)->fpw;
    return (args[r].frames + fpw - 1) / fpw;
  }

  // Work items of the requests `rs` (all of one N) under that N's current choices.
  void build_n(int n, const std::vector<size_t>& rs, std::vector<RaggedItem>& items,
               std::vector<unsigned>& chunks) const {
    items.clear();
    std::vector<double> cost;
    const int cfg = cfg_of_n.at(n);
    const unsigned per = ragged_cfg_flat(cfg) ? 1u : (unsigned)tiles_of_n.at(n);
    for (size_t r : rs) {
      size_t lead = 0;
      const size_t blocks = blocks_of(r, cfg, &lead);
      const double block_cost = (double)ragged_cfg(cfg)->fpw * (4.0 * reqs[r].hop + 12.0 * (n / 2 + 1));
      for (size_t t0 = 0; t0 < blocks; t0 += per) {
        const unsigned cnt = (unsigned)std::min<size_t>(per, blocks - t0);
        items.push_back(RaggedItem{(unsigned)r, (unsigned)cfg, (unsigned)t0, cnt});
        cost.push_back(block_cost * cnt);
      }
    }
    // costliest first (stable, so the blocks of a request stay in order)
    std::vector<size_t> order(items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<RaggedItem> sorted(items.size());
    std::vector<double> sc(items.size());
    for (size_t i = 0; i < order.size(); ++i) {
      sorted[i] = items[order[i]];
      sc[i] = cost[order[i]];
    }
    items.swap(sorted);
    double total = 0;
    for (double c : sc) total += c;
    chunks.assign(9, 0);
    double run = 0;
    size_t i = 0;
    for (int x = 1; x < 8; ++x) {
      while (i < items.size() && run + sc[i] * 0.5 < total * x / 8.0) run += sc[i++];
      chunks[x] = (unsigned)i;
    }
    chunks[8] = (unsigned)items.size();
  }

  // Upload the requests (with each one's `lead` for its N's engine) and the work lists of
  // the sizes `ns` (all of their requests); the launches of run() follow.
  void stage(const std::vector<int>& ns) {
    std::vector<StftArgs> dev_args = args;
    for (auto& kv : reqs_of_n) {
      const int cfg = cfg_of_n.at(kv.first);
      if (!ragged_cfg_lines(cfg)) continue;
      for (size_t r : kv.second) blocks_of(r, cfg, &dev_args[r].lead);
    }
    d_args.reserve(dev_args.size() * sizeof(StftArgs));
    hip_check(hipMemcpy(d_args.ptr, dev_args.data(), dev_args.size() * sizeof(StftArgs), hipMemcpyHostToDevice),
              "upload requests");
    std::vector<RaggedItem> all;
    std::vector<unsigned> all_chunks;
    launches.clear();
    for (int n : ns) {
      std::vector<RaggedItem> items;
      std::vector<unsigned> chunks;
      build_n(n, reqs_of_n.at(n), items, chunks);
      if (items.empty()) continue;
      const int cfg = cfg_of_n.at(n);
      int g = 0;
      if (ragged_cfg_flat(cfg)) {  // one workgroup per item of the longest chunk, per XCD
        unsigned longest = 0;
        for (int x = 0; x < 8; ++x) longest = std::max(longest, chunks[x + 1] - chunks[x]);
        g = 8 * (int)longest;
      } else {
        g = grid(cfg);
      }
      launches.push_back(Launch{cfg, all_chunks.size(), g});
      for (unsigned& c : chunks) c += (unsigned)all.size();
      all.insert(all.end(), items.begin(), items.end());
      all_chunks.insert(all_chunks.end(), chunks.begin(), chunks.end());
    }
    d_items.reserve(std::max<size_t>(1, all.size()) * sizeof(RaggedItem));
    if (!all.empty())
      hip_check(hipMemcpy(d_items.ptr, all.data(), all.size() * sizeof(RaggedItem), hipMemcpyHostToDevice),
                "upload items");
    d_chunks.reserve(std::max<size_t>(1, all_chunks.size()) * sizeof(unsigned));
    if (!all_chunks.empty())
      hip_check(hipMemcpy(d_chunks.ptr, all_chunks.data(), all_chunks.size() * sizeof(unsigned),
                          hipMemcpyHostToDevice),
                "upload chunks");
    n_items = all.size();
  }

  hipError_t run(hipStream_t s) const {
    for (const Launch& l : launches) {
      // chunk bounds are absolute item indices, so every launch takes the whole item list
      const hipError_t e = launch_stft_ragged(d_args.as<StftArgs>(), d_items.as<RaggedItem>(),
                                              d_chunks.as<unsigned>() + l.chunk_off, mode, l.cfg, l.grid, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }

  std::vector<int> sizes() const {
    std::vector<int> ns;
    for (auto& kv : reqs_of_n) ns.push_back(kv.first);
    std::sort(ns.rbegin(), ns.rend());  // the costliest transforms first
    return ns;
  }

  // The staged launches as branches of the capture: launch i on branch i % kBranches,
  // forked from and joined back into `stream`, so that the launches of different N run
  // side by side and each one's tail (its last partial wave of workgroups) overlaps the
  // others' work instead of leaving the GPU idle. Launches of one N share no outputs with
  // another N's, so any order is correct.
  hipError_t run_branched() {
    if (launches.size() < 2) return run(stream);
    if (!branch[0])
      for (hipStream_t& b : branch) hip_check(hipStreamCreateWithFlags(&b, hipStreamNonBlocking), "hipStreamCreate");
    while (marks.size() < launches.size() + 1) {
      hipEvent_t e = nullptr;
      hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
      marks.push_back(e);
    }
    hipError_t err = hipEventRecord(marks[0], stream);
    for (size_t i = 0; i < launches.size() && err == hipSuccess; ++i) {
      const Launch& l = launches[i];
      hipStream_t b = branch[i % kBranches];
      if (i < kBranches) err = hipStreamWaitEvent(b, marks[0], 0);
      if (err == hipSuccess)
        err = launch_stft_ragged(d_args.as<StftArgs>(), d_items.as<RaggedItem>(),
                                 d_chunks.as<unsigned>() + l.chunk_off, mode, l.cfg, l.grid, b);
      if (err == hipSuccess && i + kBranches >= launches.size()) {  // the branch's last launch
        err = hipEventRecord(marks[1 + i], b);
        if (err == hipSuccess) err = hipStreamWaitEvent(stream, marks[1 + i], 0);
      }
    }
    return err;
  }

  // Stage every size and capture the launches in a graph.
  void commit() {
    stage(sizes());
    if (exec) hip_check(hipGraphExecDestroy(exec), "hipGraphExecDestroy");
    if (graph) hip_check(hipGraphDestroy(graph), "hipGraphDestroy");
    exec = nullptr;
    graph = nullptr;
    hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const hipError_t le = run_branched();
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(stream, &g);
    hip_check(le, "launch_stft_ragged (capture)");
    hip_check(ee, "hipStreamEndCapture");
    graph = g;
    hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }

  // Median time (ms) of `repeats` timed runs of the staged launches, after one warm-up.
  float median_ms(int repeats, hipEvent_t e0, hipEvent_t e1) {
    hip_check(run(stream), "launch_stft_ragged");
    std::vector<float> t;
    for (int i = 0; i < repeats; ++i) {
      hip_check(hipEventRecord(e0, stream), "hipEventRecord");
      hip_check(run(stream), "launch_stft_ragged");
      hip_check(hipEventRecord(e1, stream), "hipEventRecord");
      hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
  }

  // Per N: time every engine of N (its tile shapes; the line-aligned engine when the
  // requests qualify) at 1, 2, 4 and 8 blocks per item on that N's requests, by the median
  // of `repeats` runs; a candidate replaces the current choice only when it is more than
  // 3% faster (timings of near-equal choices otherwise flip with noise).
  void tune(int repeats) {
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    for (int n : sizes()) {
      std::vector<int> engines;
      for (int cfg = 0; cfg < ragged_cfg_count(); ++cfg)
        if (2 * ragged_cfg(cfg)->l == n && (!ragged_cfg_lines(cfg) || lines_ok(n, cfg))) engines.push_back(cfg);
      stage({n});
      float best = median_ms(repeats, e0, e1);
      int best_cfg = cfg_of_n[n], best_per = tiles_of_n[n];
      for (int cfg : engines)
        for (int per : {1, 2, 4, 8}) {
          if (cfg == best_cfg && per == best_per) continue;
          if (ragged_cfg_flat(cfg) && per > 1) continue;  // one block per item
          cfg_of_n[n] = cfg;
          tiles_of_n[n] = per;
          stage({n});
          const float ms = median_ms(repeats, e0, e1);
          if (ms < best * 0.97f) {
            best = ms;
            best_cfg = cfg;
            best_per = per;
          }
        }
      cfg_of_n[n] = best_cfg;
      tiles_of_n[n] = best_per;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
};

}  // namespace
}  // namespace fftfree


extern "C" {

void fft_pcm_to_channels(const float* in_pcm, float* out_real, float* out_imag, float* out_mag, size_t n,
                         int threads) {
  if (!in_pcm || !out_real || !out_imag || !out_mag || n == 0) return;
  void* h = nullptr;
  try {
    const size_t plan_n = fftfree::next_pow2(n);
    h = fft_init_full(plan_n, threads, 1, 0, 0, 0, nullptr, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1);
    if (!h) throw std::runtime_error("context");
    std::vector<float> re(plan_n), im(plan_n), mag(plan_n);
    if (!fft_execute(h, in_pcm, re.data(), im.data(), mag.data(), n)) throw std::runtime_error("execute");
    std::memcpy(out_real, re.data(), n * sizeof(float));
    std::memcpy(out_imag, im.data(), n * sizeof(float));
    std::memcpy(out_mag, mag.data(), n * sizeof(float));
  } catch (...) {
    // reference behaviour: zero the outputs (fft_cffi.cpp:489-510)
    std::memset(out_real, 0, n * sizeof(float));
    std::memset(out_imag, 0, n * sizeof(float));
    std::memset(out_mag, 0, n * sizeof(float));
  }
  fft_free(h);
}

void* fft_init(size_t n, int threads, int lanes, int inverse, int kernel, int radix, const int* radix_pattern,
               size_t radix_pattern_len, int pad_mode) {
  return fft_init_ex(n, threads, lanes, inverse, kernel, radix, radix_pattern, radix_pattern_len, pad_mode, 0, 0);
}

void* fft_init_ex(size_t n, int threads, int lanes, int inverse, int kernel, int radix, const int* radix_pattern,
                  size_t radix_pattern_len, int pad_mode, int window, int hop) {
  // fft_cffi.cpp:540-658: N must be a power of two unless pad_mode == 1 (pad up).
  if (n == 0) return nullptr;
  FftContext* c = nullptr;
  try {
    if (kernel == 0) kernel = fftfree::env_default_kernel();
    size_t plan_n = n;
    if (!fftfree::is_pow2(plan_n)) {
      if (pad_mode != 1) {
        std::fprintf(stderr,
                     "fft_init: refusing to auto-pad N=%zu -> %zu because pad_mode!=ALWAYS; set pad_mode=1 to "
                     "permit padding or supply power-of-two N\n",
                     n, fftfree::next_pow2(n));
        return nullptr;
      }
      plan_n = fftfree::next_pow2(plan_n);
    }
    if (plan_n > (size_t)(1u << 30)) return nullptr;
    c = fftfree::make_context(plan_n, threads, lanes, inverse, kernel, radix, radix_pattern, radix_pattern_len,
                              pad_mode, window, hop);
    c->cfg.threads = threads > 0 ? threads : 4;  // Plan default runtime threads
    c->cfg.lanes = lanes > 0 ? lanes : 0;
    fftfree::bind_plan(c);
    return c;
  } catch (const std::exception& e) {
    fftfree::report("fft_init", e.what());
  } catch (...) {
    fftfree::report("fft_init", "unknown exception");
  }
  delete c;
  return nullptr;
}

void* fft_init_full(size_t n, int threads, int lanes, int inverse, int kernel, int radix, const int* radix_pattern,
                    size_t radix_pattern_len, int pad_mode, int window, int hop, int transform,
                    int reduce_magnitude, int store_polar, int half_spectrum, int allow_outer_parallel,
                    int allow_inner_parallel, int inner_threads, int save_crash_logs, int silent_crash_reports) {
  return fft_init_full_v2(n, threads, lanes, inverse, kernel, radix, radix_pattern, radix_pattern_len, pad_mode,
                          window, hop, transform, reduce_magnitude, store_polar, half_spectrum, allow_outer_parallel,
                          allow_inner_parallel, inner_threads, save_crash_logs, silent_crash_reports, 0, 0, 0, 0.0f,
                          0.0f, 0, 0.0f, 0.0f, 0, 0);
}

void* fft_init_full_v2(size_t n, int threads, int lanes, int inverse, int kernel, int radix,
                       const int* radix_pattern, size_t radix_pattern_len, int pad_mode, int window, int hop,
                       int transform, int reduce_magnitude, int store_polar, int half_spectrum,
                       int allow_outer_parallel, int allow_inner_parallel, int inner_threads, int save_crash_logs,
                       int silent_crash_reports, int apply_windows, int apply_ola, int analysis_window_kind,
                       float analysis_param1, float analysis_param2, int synthesis_window_kind,
                       float synthesis_param1, float synthesis_param2, int window_norm_policy, int cola_mode) {
  (void)save_crash_logs;
  (void)silent_crash_reports;
  FftContext* c = nullptr;
  try {
    // The reference's Plan constructor loops forever on N == 0; refuse it instead.
    if (n == 0 || n > (size_t)(1u << 30)) return nullptr;
    c = fftfree::make_context(n, threads, lanes, inverse, kernel, radix, radix_pattern, radix_pattern_len, pad_mode,
                              window, hop);
    c->transform = transform;
    c->reduce_magnitude = reduce_magnitude;
    c->store_polar = store_polar;
    c->half_spectrum = half_spectrum;
    c->allow_outer_parallel = allow_outer_parallel;
    c->allow_inner_parallel = allow_inner_parallel;
    c->inner_threads = inner_threads;
    c->analysis_win_kind = analysis_window_kind;
    c->synthesis_win_kind = synthesis_window_kind;
    c->win_param1 = analysis_param1;
    c->win_param2 = analysis_param2;
    c->win_norm_policy = window_norm_policy;
    c->cola_mode = cola_mode;
    c->windows_enabled = apply_windows != 0;
    c->apply_ola = apply_ola != 0;
    // Windows are N taps: the first W come from the builder, the rest are 1.0
    // (fft_cffi.cpp:1894-1906).
    const int useW = window > 0 ? window : c->N;
    {
      auto aw = fftfree::build_window(analysis_window_kind, useW, analysis_param1, analysis_param2);
      fftfree::normalize_window(aw, window_norm_policy);
      c->analysis_win.assign((size_t)c->N, 1.0f);
      for (int i = 0; i < useW && i < c->N; ++i) c->analysis_win[i] = aw[i];
    }
    {
      auto sw = fftfree::build_window(synthesis_window_kind, useW, synthesis_param1, synthesis_param2);
      fftfree::normalize_window(sw, window_norm_policy);
      c->synthesis_win.assign((size_t)c->N, 1.0f);
      for (int i = 0; i < useW && i < c->N; ++i) c->synthesis_win[i] = sw[i];
    }
    c->cfg.transform = transform;
    c->cfg.reduce_magnitude = reduce_magnitude != 0;
    c->cfg.store_polar = store_polar != 0;
    c->cfg.half_spectrum = half_spectrum != 0;
    c->cfg.allow_outer_parallel = allow_outer_parallel != 0;
    c->cfg.allow_inner_parallel = allow_inner_parallel != 0;
    c->cfg.inner_threads = inner_threads;
    c->has_pool = allow_outer_parallel != 0;
    if (fftfree::verbose()) {
      std::fprintf(stderr, "[diag] fft_init_full: N=%d threads=%d lanes=%d allow_outer=%d allow_inner=%d\n", c->N,
                   c->cfg.threads, c->cfg.lanes, allow_outer_parallel ? 1 : 0, allow_inner_parallel ? 1 : 0);
    }
    fftfree::bind_plan(c);  // throws for a non-power-of-two N, as the Plan ctor does
    return c;
  } catch (const std::exception& e) {
    fftfree::report("fft_init_full", e.what());
  } catch (...) {
    fftfree::report("fft_init_full", "unknown exception");
  }
  delete c;
  return nullptr;
}

int fft_execute(void* handle, const float* in_pcm, float* out_real, float* out_imag, float* out_mag, size_t n) {
  if (!handle || !in_pcm || !out_real || !out_imag || !out_mag || n == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (c->pad_mode == 2 && (int)n != c->N) return 0;
  if (n > (size_t)c->N) return 0;
  try {
    fftfree::ensure_gpu(c);
    const size_t N = (size_t)c->N;
    std::vector<std::complex<float>> col(N);
    for (size_t i = 0; i < n; ++i) col[i] = std::complex<float>(in_pcm[i], 0.0f);
    c->d_cplx.reserve(N * sizeof(float2));
    fftfree::hip_check(hipMemcpyAsync(c->d_cplx.ptr, col.data(), N * sizeof(float2), hipMemcpyHostToDevice, c->stream),
                       "H2D");
    fftfree::AxisLayout l;
    l.base = c->d_cplx.as<float2>();
    l.axis_size = c->N;
    l.batch_size = 1;
    l.batch_stride = c->N;
    fftfree::fft_apply_axis(*c->plan, l);
    fftfree::hip_check(hipMemcpyAsync(col.data(), c->d_cplx.ptr, N * sizeof(float2), hipMemcpyDeviceToHost, c->stream),
                       "D2H");
    fftfree::hip_check(hipStreamSynchronize(c->stream), "sync");
    // fft_cffi.cpp:690-695: re, im, |v| of the post-transformed value
    for (size_t j = 0; j < N; ++j) {
      out_real[j] = col[j].real();
      out_imag[j] = col[j].imag();
      out_mag[j] = std::abs(col[j]);
    }
    return 1;
  } catch (const std::exception& e) {
    fftfree::report("fft_execute", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_frame_count(void* handle, size_t pcm_len, int pad_mode, size_t max_frames) {
  if (!handle) return 0;
  return fftfree::frame_count(static_cast<FftContext*>(handle), pcm_len, pad_mode, max_frames);
}

size_t fft_execute_batched(void* handle, const float* pcm, size_t pcm_len, float* out_real, float* out_imag,
                           float* out_mag, int pad_mode, int enable_backup, size_t max_frames) {
  (void)enable_backup;
  if (!handle || !pcm || pcm_len == 0 || !out_real || !out_imag || !out_mag) return 0;
  auto* c = static_cast<FftContext*>(handle);
  const size_t frames = fftfree::frame_count(c, pcm_len, pad_mode, max_frames);
  if (frames == 0) return 0;
  size_t produced = 0;
  try {
    fftfree::ensure_gpu(c);
    return fftfree::stft_host_pipelined(c, pcm, pcm_len, frames, out_real, out_imag, out_mag, &produced);
  } catch (const std::exception& e) {
    fftfree::drain_pipeline(c);
    fftfree::report("fft_execute_batched", e.what());
    return produced;
  } catch (...) {
    fftfree::drain_pipeline(c);
    return produced;
  }
}

size_t fft_stft_execute(void* handle, const float* pcm, size_t pcm_len, float* out_real, float* out_imag,
                        float* out_mag, size_t max_frames) {
  return fft_execute_batched(handle, pcm, pcm_len, out_real, out_imag, out_mag, 1, 0, max_frames);
}

size_t fft_execute_batched_device(void* handle, const float* d_pcm, size_t pcm_len, float* d_out_real,
                                  float* d_out_imag, float* d_out_mag, int pad_mode, size_t max_frames,
                                  void* stream) {
  if (!handle || !d_pcm || pcm_len == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  const size_t frames = fftfree::frame_count(c, pcm_len, pad_mode, max_frames);
  if (frames == 0) return 0;
  try {
    fftfree::ensure_gpu(c);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const fftfree::StftArgs base = fftfree::stft_args(c, d_pcm, pcm_len);
    fftfree::run_stft_frames(c, base, 0, frames, d_out_real, d_out_imag, d_out_mag, s);
    return frames;
  } catch (const std::exception& e) {
    fftfree::report("fft_execute_batched_device", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

// The signals entries: a forward R2C half-spectrum context; N and the window/hop must
// be ones the signal kernels take (signals_supported).
bool signals_ok(const FftContext* c, bool adjoint) {
  if (c->inverse || c->transform != FFT_TRANSFORM_R2C || !c->half_spectrum) return false;
  const int W = c->window ? c->window : c->N;
  const int H = c->hop ? c->hop : W;
  return fftfree::signals_supported(c->N, W, H, adjoint);
}

fftfree::SignalArgs signal_args(FftContext* c, const float* d_pcm, size_t signals, size_t pcm_len, size_t frames) {
  const fftfree::StftArgs s = fftfree::stft_args(c, d_pcm, pcm_len);
  fftfree::SignalArgs a;
  a.pcm = d_pcm;
  a.len = pcm_len;
  a.signals = signals;
  a.frames = frames;
  a.win = s.win;
  a.window = s.window;
  a.hop = s.hop;
  a.n = c->N;
  a.tw = s.tw;
  return a;
}

size_t fft_stft_signals_device(void* handle, const float* d_pcm, size_t signals, size_t pcm_len, float* d_out_real,
                               float* d_out_imag, float* d_out_mag, int pad_mode, void* stream) {
  if (!handle || !d_pcm || signals == 0 || pcm_len == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (!signals_ok(c, false)) return 0;
  const size_t frames = fftfree::frame_count(c, pcm_len, pad_mode, 0);
  if (frames == 0) return 0;
  try {
    fftfree::ensure_gpu(c);
    fftfree::SignalArgs a = signal_args(c, d_pcm, signals, pcm_len, frames);
    a.o_re = d_out_real;
    a.o_im = d_out_imag;
    a.o_mag = d_out_mag;
    const int mode = c->store_polar ? 3 : c->reduce_magnitude ? 2 : 0;  // polar, magnitude, complex
    fftfree::hip_check(fftfree::launch_stft_signals(a, mode, static_cast<hipStream_t>(stream)), "launch_stft_signals");
    return frames;
  } catch (const std::exception& e) {
    fftfree::report("fft_stft_signals_device", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_stft_adjoint_device(void* handle, const float* d_pcm, size_t signals, size_t pcm_len,
                               const float* d_grad_real, const float* d_grad_imag, float* d_grad_pcm, int pad_mode,
                               void* stream) {
  if (!handle || !d_pcm || !d_grad_pcm || signals == 0 || pcm_len == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (!signals_ok(c, true)) return 0;
  const size_t frames = fftfree::frame_count(c, pcm_len, pad_mode, 0);
  if (frames == 0) return 0;
  try {
    fftfree::ensure_gpu(c);
    fftfree::SignalArgs a = signal_args(c, d_pcm, signals, pcm_len, frames);
    a.g_mag = d_grad_real;
    a.g_phase = d_grad_imag;
    a.grad = d_grad_pcm;
    const size_t halo_floats = fftfree::adjoint_scratch(c->N, a.window, a.hop, signals, pcm_len);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (halo_floats > 0) {
      c->d_halo.reserve(halo_floats * sizeof(float));
      a.halo = c->d_halo.as<float>();
      a.halo_capacity = c->d_halo.bytes / sizeof(float);
    }
    const int mode = c->store_polar ? 3 : c->reduce_magnitude ? 2 : 0;  // polar, magnitude, complex
    fftfree::hip_check(fftfree::launch_stft_adjoint(a, mode, s), "launch_stft_adjoint");
    return frames;
  } catch (const std::exception& e) {
    fftfree::report("fft_stft_adjoint_device", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_stft_adjoint_saved_device(void* handle, const float* d_plane_real, const float* d_plane_imag,
                                     size_t signals, size_t pcm_len, const float* d_grad_real,
                                     const float* d_grad_imag, float* d_grad_pcm, int pad_mode, void* stream) {
  if (!handle || !d_grad_pcm || signals == 0 || pcm_len == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (!signals_ok(c, true)) return 0;
  if (c->reduce_magnitude && !c->store_polar) {
    fftfree::report("fft_stft_adjoint_saved_device",
                    "magnitude planes do not hold the direction of X: use fft_stft_adjoint_device");
    return 0;
  }
  const bool polar = c->store_polar;
  if (polar && (!d_plane_real || !d_plane_imag)) {
    fftfree::report("fft_stft_adjoint_saved_device", "polar planes |X| and arg X are required");
    return 0;
  }
  const size_t frames = fftfree::frame_count(c, pcm_len, pad_mode, 0);
  if (frames == 0) return 0;
  try {
    fftfree::ensure_gpu(c);
    fftfree::SignalArgs a = signal_args(c, nullptr, signals, pcm_len, frames);
    a.g_mag = d_grad_real;
    a.g_phase = d_grad_imag;
    a.p_mag = polar ? d_plane_real : nullptr;
    a.p_phase = polar ? d_plane_imag : nullptr;
    a.grad = d_grad_pcm;
    const size_t halo_floats = fftfree::adjoint_scratch(c->N, a.window, a.hop, signals, pcm_len);
    if (halo_floats > 0) {
      c->d_halo.reserve(halo_floats * sizeof(float));
      a.halo = c->d_halo.as<float>();
      a.halo_capacity = c->d_halo.bytes / sizeof(float);
    }
    fftfree::hip_check(fftfree::launch_stft_adjoint(a, polar ? 3 : 0, static_cast<hipStream_t>(stream), nullptr, true),
                       "launch_stft_adjoint(saved)");
    return frames;
  } catch (const std::exception& e) {
    fftfree::report("fft_stft_adjoint_saved_device", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_stft_adjoint_device_noalloc(void* handle, const float* d_pcm, size_t signals, size_t pcm_len,
                                       const float* d_grad_real, const float* d_grad_imag, float* d_grad_pcm,
                                       int pad_mode, void* stream) {
  if (!handle || !d_pcm || !d_grad_pcm || signals == 0 || pcm_len == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (!signals_ok(c, true)) return 0;
  const size_t frames = fftfree::frame_count(c, pcm_len, pad_mode, 0);
  if (frames == 0) return 0;
  try {
    // no allocation here: the context's device state (stream, twiddles, window) must
    // exist already, and the halo slots come from the library's static scratch
    if (!c->gpu_ready || c->windows_dirty || !c->plan->W) {
      fftfree::report("fft_stft_adjoint_device_noalloc",
                      "the context has no device state yet: make one allocating device call on it first "
                      "(e.g. fft_stft_signals_device)");
      return 0;
    }
    fftfree::hip_check(hipSetDevice(c->device), "hipSetDevice");
    fftfree::SignalArgs a = signal_args(c, d_pcm, signals, pcm_len, frames);
    a.g_mag = d_grad_real;
    a.g_phase = d_grad_imag;
    a.grad = d_grad_pcm;
    const int mode = c->store_polar ? 3 : c->reduce_magnitude ? 2 : 0;  // polar, magnitude, complex
    hipStream_t s = static_cast<hipStream_t>(stream);
    size_t need = 0, capacity = 0;
    fftfree::hip_check(fftfree::launch_stft_adjoint(a, mode, s, &need), "plan adjoint");
    if (need > 0) {
      float* scr = fftfree::noalloc_halo_scratch(&capacity);
      if (!scr) throw fftfree::HipFailure("hipGetSymbolAddress(scratch) failed");
      if (need > capacity) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "this shape needs %zu floats of halo scratch; the allocation-free entry holds %zu",
                      need, capacity);
        fftfree::report("fft_stft_adjoint_device_noalloc", msg);
        return 0;
      }
      a.halo = scr;
      a.halo_capacity = capacity;
    }
    fftfree::hip_check(fftfree::launch_stft_adjoint(a, mode, s), "launch_stft_adjoint");
    return frames;
  } catch (const std::exception& e) {
    fftfree::report("fft_stft_adjoint_device_noalloc", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_execute_c2c_batched_device(void* handle, const float* d_in, float* d_out, size_t batch, void* stream) {
  if (!handle || !d_in || !d_out || batch == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  try {
    fftfree::ensure_gpu(c);
    hipStream_t s = static_cast<hipStream_t>(stream);
    fftfree::AxisLayout l;
    l.base = reinterpret_cast<float2*>(d_out);
    if (d_in != d_out) l.source = reinterpret_cast<const float2*>(d_in);
    l.axis_size = c->N;
    l.batch_size = (int64_t)batch;
    l.batch_stride = c->N;
    c->dispatcher.set_stream(s);
    fftfree::fft_apply_axis(*c->plan, l);
    c->dispatcher.set_stream(c->stream);
    return batch;
  } catch (const std::exception& e) {
    c->dispatcher.set_stream(c->stream);
    fftfree::report("fft_execute_c2c_batched_device", e.what());
    return 0;
  } catch (...) {
    c->dispatcher.set_stream(c->stream);
    return 0;
  }
}

size_t fft_execute_c2c_batched(void* handle, const float* in, float* out, size_t batch) {
  if (!handle || !in || !out || batch == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  try {
    fftfree::ensure_gpu(c);
    const size_t bytes = batch * (size_t)c->N * sizeof(float2);
    c->d_cplx.reserve(bytes);
    fftfree::hip_check(hipMemcpyAsync(c->d_cplx.ptr, in, bytes, hipMemcpyHostToDevice, c->stream), "H2D");
    const size_t got = fft_execute_c2c_batched_device(handle, c->d_cplx.as<float>(), c->d_cplx.as<float>(), batch,
                                                      c->stream);
    if (got != batch) return 0;
    fftfree::hip_check(hipMemcpyAsync(out, c->d_cplx.ptr, bytes, hipMemcpyDeviceToHost, c->stream), "D2H");
    fftfree::hip_check(hipStreamSynchronize(c->stream), "sync");
    return batch;
  } catch (const std::exception& e) {
    fftfree::report("fft_execute_c2c_batched", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

namespace fftfree {
namespace {

// The device form of fft_execute_complex_batched (fft_cffi.cpp:1042-1157).
size_t complex_batched(FftContext* c, const float* d_re, const float* d_im, size_t frames, float* d_out,
                       hipStream_t s) {
  IstftArgs a;
  a.in_re = d_re;
  a.in_im = d_im;
  a.frames = frames;
  a.n = c->N;
  a.bins = (int)bins_of(c);
  a.hop = c->hop > 0 ? c->hop : (c->window > 0 ? c->window : c->N);  // fft_cffi.cpp:1120
  a.tw = c->plan->W;
  a.swin = (c->windows_enabled && !c->synthesis_win.empty()) ? c->d_swin.as<float>() : nullptr;
  a.awin = (c->windows_enabled && !c->analysis_win.empty()) ? c->d_awin.as<float>() : nullptr;
  a.ola = c->apply_ola ? 1 : 0;
  a.cola = c->cola_mode == FFT_COLA_NORMALIZE ? 1 : 0;
  a.out = d_out;
  // An inverse C2R half-spectrum plan without a polar/magnitude post hook computes
  // x = Re IDFT of the Hermitian rebuild on either kernel: the fused kernel.
  const bool fused = c->inverse && c->transform == FFT_TRANSFORM_C2R && c->half_spectrum && !c->store_polar &&
                     !c->reduce_magnitude && istft_fused_supported(c->N, a.hop, a.ola);
  if (fused) {
    hip_check(launch_istft(a, s), "launch_istft");
    return frames;
  }
  const size_t bytes = frames * (size_t)c->N * sizeof(float2);
  c->d_cplx.reserve(bytes);
  const int hermitian = (c->half_spectrum && c->transform == FFT_TRANSFORM_C2R) ? 1 : 0;
  hip_check(launch_complex_columns(a, hermitian, c->d_cplx.as<float2>(), s), "launch_complex_columns");
  AxisLayout l;
  l.base = c->d_cplx.as<float2>();
  l.axis_size = c->N;
  l.batch_size = (int64_t)frames;
  l.axis_stride = 1;
  l.batch_stride = c->N;
  c->dispatcher.set_stream(s);
  try {
    fft_apply_axis(*c->plan, l);
  } catch (...) {
    c->dispatcher.set_stream(c->stream);
    throw;
  }
  c->dispatcher.set_stream(c->stream);
  hip_check(launch_overlap(a, c->d_cplx.as<float2>(), s), "launch_overlap");
  return frames;
}

// Griffin-Lim (phase_infer.hpp:69-204, called from fft_cffi.cpp:2176-2248) with every
// buffer on the device. Per iteration: the inverse context's transform of the current
// spectrum, overlap-add averaged over the frames per sample, the forward context's STFT,
// and the phase update X = |X| (cos, sin)(atan2(Im, Re)) of the frames the forward
// produced. When the inverse context is what the fused inverse kernel runs, that kernel
// does the transform and the averaged overlap-add, with the phase update applied as it
// loads the bins, and the update is written out once at the end.
bool griffin_lim(FftContext* fw, FftContext* iv, const float* d_mag, size_t frames, int hop, int half,
                 int iterations, int pad_mode, unsigned seed, float* d_re, float* d_im, hipStream_t s) {
  const int N = iv->N;
  if (N <= 0 || fw->N != N) return false;  // the reference would write past its buffers
  const size_t bins = half ? (size_t)(N / 2 + 1) : (size_t)N;
  if (bins != bins_of(iv) || bins != bins_of(fw)) return false;
  const size_t total = frames * bins;
  const size_t ola_len = (frames - 1) * (size_t)hop + (size_t)N;
  const int iters = iterations > 0 ? iterations : 32;
  const size_t fwd_frames = frame_count(fw, ola_len, pad_mode, frames);
  if (fwd_frames == 0) return false;
  const size_t uf = std::min(fwd_frames, frames);
  ensure_gpu(fw);
  const bool fused = iv->inverse && iv->transform == FFT_TRANSFORM_C2R && iv->half_spectrum && !iv->store_polar &&
                     !iv->reduce_magnitude && !iv->apply_ola && istft_fused_supported(N, hop, 1);
  // working set: start words, the overlap-added signal, the forward planes, and (general
  // path) the frame-major inverse output
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t o_words = 0, o_ola = up(total * 4), o_tre = o_ola + up(ola_len * 4), o_tim = o_tre + up(total * 4),
               o_fr = o_tim + up(total * 4), bytes = o_fr + (fused ? 0 : up(frames * (size_t)N * 4));
  const size_t had = iv->d_gl.bytes;
  iv->d_gl.reserve(bytes);
  if (iv->d_gl.bytes != had) iv->gl_words_count = 0;  // a new allocation holds no words
  char* base = iv->d_gl.as<char>();
  uint32_t* d_words = reinterpret_cast<uint32_t*>(base + o_words);
  float* d_ola = reinterpret_cast<float*>(base + o_ola);
  float* d_tre = reinterpret_cast<float*>(base + o_tre);
  float* d_tim = reinterpret_cast<float*>(base + o_tim);
  float* d_fr = reinterpret_cast<float*>(base + o_fr);

  // the start: mt19937 is sequential, so its words come from the host (phase_infer.hpp:124-134),
  // into a pinned buffer of the context, copied in chunks while the next chunk is drawn. The
  // words depend only on the seed and their count: a call that repeats both on this context
  // finds them on the device already.
  const unsigned seed_used = seed ? seed : 0x47d7bfadu;
  if (iv->gl_words_count != total || iv->gl_words_seed != seed_used) {
    iv->gl_words_count = 0;
    if (iv->words_copied) hip_check(hipEventSynchronize(iv->words_copied), "wait for the last start copy");
    else hip_check(hipEventCreateWithFlags(&iv->words_copied, hipEventDisableTiming), "hipEventCreate");
    if (iv->h_words_cap < total) {
      if (iv->h_words) hip_check(hipHostFree(iv->h_words), "hipHostFree");
      iv->h_words = nullptr;
      iv->h_words_cap = 0;
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&iv->h_words), total * 4, hipHostMallocDefault),
                "hipHostMalloc");
      iv->h_words_cap = total;
    }
    Mt19937Block rng(seed_used);  // the values of std::mt19937
    constexpr size_t kChunk = size_t(1) << 18;
    for (size_t b = 0; b < total; b += kChunk) {
      const size_t e = std::min(total, b + kChunk);
      rng.fill(iv->h_words + b, e - b);
      hip_check(hipMemcpyAsync(d_words + b, iv->h_words + b, (e - b) * 4, hipMemcpyHostToDevice, s), "H2D words");
    }
    hip_check(hipEventRecord(iv->words_copied, s), "hipEventRecord");
    iv->gl_words_seed = seed_used;
    iv->gl_words_count = total;
  } else {
    // the copy that put them there may have been enqueued on another stream
    hip_check(hipStreamWaitEvent(s, iv->words_copied, 0), "hipStreamWaitEvent(start words)");
  }
  hip_check(launch_gl_start(d_mag, d_words, total, d_re, d_im, s), "launch_gl_start");
  const StftArgs fa = stft_args(fw, d_ola, ola_len);
  // the forward planes are read back by the next iteration's inverse straight away
  StftArgs fq = fa;
  fq.frames = uf;
  const int fwd_variant = readback_stft_variant(fq);
  for (int it = 0; it < iters; ++it) {
    if (fused) {
      IstftArgs a;
      a.in_re = d_re;
      a.in_im = d_im;
      a.frames = frames;
      a.n = N;
      a.bins = (int)bins;
      a.hop = hop;
      a.tw = iv->plan->W;
      a.swin = (iv->windows_enabled && !iv->synthesis_win.empty()) ? iv->d_swin.as<float>() : nullptr;
      a.ola = 1;
      a.cola = 1;
      a.unit_cola = 1;
      a.out = d_ola;
      a.out_len = ola_len;
      a.gl_mag = d_mag;
      a.gl_re = d_tre;
      a.gl_im = d_tim;
      a.gl_frames = it > 0 ? uf : 0;
      hip_check(launch_istft(a, s), "launch_istft (Griffin-Lim)");
    } else {
      if (it > 0)
        hip_check(launch_gl_update(d_mag, d_tre, d_tim, uf, (int)bins, half, (N & 1) == 0, d_re, d_im, s),
                  "launch_gl_update");
      complex_batched(iv, d_re, d_im, frames, d_fr, s);
      hip_check(launch_gl_overlap(d_fr, frames, N, hop, d_ola, ola_len, s), "launch_gl_overlap");
    }
    run_stft_frames(fw, fa, 0, uf, d_tre, d_tim, nullptr, s, fwd_variant);
  }
  hip_check(launch_gl_update(d_mag, d_tre, d_tim, uf, (int)bins, half, (N & 1) == 0, d_re, d_im, s),
            "launch_gl_update");
  return true;
}

// Argument checks and hop choice of fft_griffin_lim (fft_cffi.cpp:2176-2248).
int griffin_lim_args(void* ctx_forward, void* ctx_inverse, const void* in_mag, size_t frames, int* hop,
                     int half_spectrum, const void* out_real, const void* out_imag) {
  if (!ctx_forward || !ctx_inverse || !in_mag || !out_real || !out_imag || frames == 0) return 0;
  if (in_mag == out_real || in_mag == out_imag || out_real == out_imag) return 0;
  auto* fw = static_cast<FftContext*>(ctx_forward);
  auto* iv = static_cast<FftContext*>(ctx_inverse);
  const int N = iv->N > 0 ? iv->N : fw->N;
  if (N <= 0) return 0;
  const bool half = half_spectrum != 0;
  if (half != (iv->half_spectrum != 0) || half != (fw->half_spectrum != 0)) return 0;
  if (*hop <= 0) *hop = iv->hop > 0 ? iv->hop : (fw->hop > 0 ? fw->hop : N);
  return 1;
}

}  // namespace
}  // namespace fftfree

int fft_griffin_lim_device(void* ctx_forward, void* ctx_inverse, const float* d_in_mag, size_t frames, int hop,
                           int half_spectrum, int iterations, int pad_mode, unsigned int seed, float* d_out_real,
                           float* d_out_imag, void* stream) {
  if (!fftfree::griffin_lim_args(ctx_forward, ctx_inverse, d_in_mag, frames, &hop, half_spectrum, d_out_real,
                                 d_out_imag))
    return 0;
  try {
    auto* iv = static_cast<FftContext*>(ctx_inverse);
    fftfree::ensure_gpu(iv);
    return fftfree::griffin_lim(static_cast<FftContext*>(ctx_forward), iv, d_in_mag, frames, hop, half_spectrum != 0,
                                iterations, pad_mode, seed, d_out_real, d_out_imag, static_cast<hipStream_t>(stream))
               ? 1
               : 0;
  } catch (const std::exception& e) {
    fftfree::report("fft_griffin_lim_device", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

int fft_griffin_lim(void* ctx_forward, void* ctx_inverse, const float* in_mag, size_t frames, int hop,
                    int half_spectrum, int iterations, int pad_mode, unsigned int seed, float* out_real,
                    float* out_imag) {
  if (!fftfree::griffin_lim_args(ctx_forward, ctx_inverse, in_mag, frames, &hop, half_spectrum, out_real, out_imag))
    return 0;
  try {
    auto* iv = static_cast<FftContext*>(ctx_inverse);
    fftfree::ensure_gpu(iv);
    const size_t bins = fftfree::bins_of(iv);
    const size_t bytes = frames * bins * sizeof(float);
    const size_t off = (bytes + 255) & ~size_t(255);
    iv->d_planes.reserve(3 * off);
    float* d_mag = iv->d_planes.as<float>();
    float* d_re = reinterpret_cast<float*>(iv->d_planes.as<char>() + off);
    float* d_im = reinterpret_cast<float*>(iv->d_planes.as<char>() + 2 * off);
    fftfree::hip_check(hipMemcpyAsync(d_mag, in_mag, bytes, hipMemcpyHostToDevice, iv->stream), "H2D");
    if (!fftfree::griffin_lim(static_cast<FftContext*>(ctx_forward), iv, d_mag, frames, hop, half_spectrum != 0,
                              iterations, pad_mode, seed, d_re, d_im, iv->stream))
      return 0;
    fftfree::hip_check(hipMemcpyAsync(out_real, d_re, bytes, hipMemcpyDeviceToHost, iv->stream), "D2H");
    fftfree::hip_check(hipMemcpyAsync(out_imag, d_im, bytes, hipMemcpyDeviceToHost, iv->stream), "D2H");
    fftfree::hip_check(hipStreamSynchronize(iv->stream), "sync");
    return 1;
  } catch (const std::exception& e) {
    fftfree::report("fft_griffin_lim", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_execute_complex_batched_device(void* handle, const float* d_in_real, const float* d_in_imag,
                                          size_t frames, float* d_out_pcm, size_t max_frames, void* stream) {
  if (!handle || !d_in_real || !d_in_imag || !d_out_pcm || frames == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (c->N <= 0) return 0;
  if (max_frames != 0 && frames > max_frames) frames = max_frames;
  try {
    fftfree::ensure_gpu(c);
    return fftfree::complex_batched(c, d_in_real, d_in_imag, frames, d_out_pcm, static_cast<hipStream_t>(stream));
  } catch (const std::exception& e) {
    fftfree::report("fft_execute_complex_batched_device", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_execute_complex_batched(void* handle, const float* in_real, const float* in_imag, size_t frames,
                                   float* out_pcm, int pad_mode, int enable_backup, size_t max_frames) {
  (void)pad_mode;       // not read by the reference either
  (void)enable_backup;  // selects the reference's retry path, same output
  if (!handle || !in_real || !in_imag || !out_pcm || frames == 0) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (c->N <= 0) return 0;
  if (max_frames != 0 && frames > max_frames) frames = max_frames;
  try {
    fftfree::ensure_gpu(c);
    const size_t N = (size_t)c->N;
    const size_t bins = fftfree::bins_of(c);
    const size_t in_bytes = frames * bins * sizeof(float);
    const size_t out_bytes = frames * N * sizeof(float);
    // staging: the two input planes, then the output
    const size_t in_off = (2 * in_bytes + 255) & ~size_t(255);
    c->d_planes.reserve(in_off + out_bytes);
    float* d_re = c->d_planes.as<float>();
    float* d_im = d_re + frames * bins;
    float* d_out = reinterpret_cast<float*>(c->d_planes.as<char>() + in_off);
    fftfree::hip_check(hipMemcpyAsync(d_re, in_real, in_bytes, hipMemcpyHostToDevice, c->stream), "H2D");
    fftfree::hip_check(hipMemcpyAsync(d_im, in_imag, in_bytes, hipMemcpyHostToDevice, c->stream), "H2D");
    fftfree::complex_batched(c, d_re, d_im, frames, d_out, c->stream);
    fftfree::hip_check(hipMemcpyAsync(out_pcm, d_out, out_bytes, hipMemcpyDeviceToHost, c->stream), "D2H");
    fftfree::hip_check(hipStreamSynchronize(c->stream), "sync");
    return frames;
  } catch (const std::exception& e) {
    fftfree::report("fft_execute_complex_batched", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

// ---- phase inference (fft_cffi.cpp:2058-2174) -------------------------------------------
namespace fftfree {
namespace {

// phase_init's context (fft_cffi.cpp:2060-2080): the configuration, and for mode 1 two
// complex contexts of size N (inverse and forward) whose transforms the cepstrum runs on.
struct PhaseCtx {
  int n = 0, hop = 0, half = 1, mode = 0, iterations = 0;
  FftContext* inv = nullptr;
  FftContext* fwd = nullptr;
  DeviceBuffer scratch;  // mode 0: phase checkpoints; mode 1: frames x N complex
  DeviceBuffer staging;  // host entry: magnitude and output planes
  ~PhaseCtx() {
    scratch.release();
    staging.release();
    if (inv) fft_free(inv);
    if (fwd) fft_free(fwd);
  }
};

size_t phase_execute(PhaseCtx* c, const float* d_mag, size_t frames, float* d_re, float* d_im, hipStream_t s) {
  const int N = c->n;
  const int bins = c->half ? N / 2 + 1 : N;
  if (c->mode == 0) {
    hip_check(launch_phase_linear(d_mag, frames, bins, N, c->hop, d_re, d_im, s), "launch_phase_linear");
    return frames;
  }
  if (c->mode != 1 || !c->inv || !c->fwd) return 0;  // unknown mode, or N with no plan
  if (phase_minimum_fused(N)) {  // every step in one launch, one wave per frame
    ensure_gpu(c->fwd);
    hip_check(launch_phase_minimum(d_mag, frames, bins, N, c->half, c->fwd->plan->W, d_re, d_im, s),
              "launch_phase_minimum");
    return frames;
  }
  c->scratch.reserve(frames * (size_t)N * sizeof(float2));
  float2* cols = c->scratch.as<float2>();
  hip_check(launch_phase_log(d_mag, frames, bins, N, c->half, cols, s), "launch_phase_log");
  if (fft_execute_c2c_batched_device(c->inv, reinterpret_cast<float*>(cols), reinterpret_cast<float*>(cols), frames,
                                     s) != frames)
    return 0;
  hip_check(launch_phase_fold(frames, N, cols, s), "launch_phase_fold");
  if (fft_execute_c2c_batched_device(c->fwd, reinterpret_cast<float*>(cols), reinterpret_cast<float*>(cols), frames,
                                     s) != frames)
    return 0;
  hip_check(launch_phase_exp(cols, frames, bins, N, d_re, d_im, s), "launch_phase_exp");
  return frames;
}

}  // namespace
}  // namespace fftfree

void* phase_init(int N, int hop, int half_spectrum, int mode, int iterations) {
  if (N <= 0) return nullptr;
  try {
    auto* c = new fftfree::PhaseCtx();
    c->n = N;
    c->hop = hop > 0 ? hop : N;
    c->half = half_spectrum != 0;
    c->mode = mode;
    c->iterations = iterations;
    if (mode == 1 && fftfree::is_pow2((size_t)N)) {  // the reference's plans need a power of two
      c->inv = static_cast<FftContext*>(fft_init_full_v2((size_t)N, 1, 1, 1, 0, 0, nullptr, 0, 0, 0, 0,
                                                         FFT_TRANSFORM_C2C, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0.f,
                                                         0.f, 0, 0.f, 0.f, 0, 0));
      c->fwd = static_cast<FftContext*>(fft_init_full_v2((size_t)N, 1, 1, 0, 0, 0, nullptr, 0, 0, 0, 0,
                                                         FFT_TRANSFORM_C2C, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0.f,
                                                         0.f, 0, 0.f, 0.f, 0, 0));
    }
    return c;
  } catch (...) {
    return nullptr;
  }
}

size_t phase_infer_execute_device(void* handle, const float* d_in_mag, size_t frames, float* d_out_real,
                                  float* d_out_imag, void* stream) {
  if (!handle || !d_in_mag || !d_out_real || !d_out_imag || frames == 0) return 0;
  try {
    return fftfree::phase_execute(static_cast<fftfree::PhaseCtx*>(handle), d_in_mag, frames, d_out_real, d_out_imag,
                                  static_cast<hipStream_t>(stream));
  } catch (const std::exception& e) {
    fftfree::report("phase_infer_execute_device", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t phase_infer_execute(void* handle, const float* in_mag, size_t frames, float* out_real, float* out_imag) {
  if (!handle || !in_mag || !out_real || !out_imag || frames == 0) return 0;
  auto* c = static_cast<fftfree::PhaseCtx*>(handle);
  try {
    const size_t bins = c->half ? (size_t)(c->n / 2 + 1) : (size_t)c->n;
    const size_t bytes = frames * bins * sizeof(float);
    const size_t off = (bytes + 255) & ~size_t(255);
    c->staging.reserve(3 * off);
    float* d_mag = c->staging.as<float>();
    float* d_re = reinterpret_cast<float*>(c->staging.as<char>() + off);
    float* d_im = reinterpret_cast<float*>(c->staging.as<char>() + 2 * off);
    fftfree::hip_check(hipMemcpy(d_mag, in_mag, bytes, hipMemcpyHostToDevice), "H2D");
    const size_t got = fftfree::phase_execute(c, d_mag, frames, d_re, d_im, nullptr);
    if (got == 0) return 0;
    fftfree::hip_check(hipMemcpy(out_real, d_re, bytes, hipMemcpyDeviceToHost), "D2H");
    fftfree::hip_check(hipMemcpy(out_imag, d_im, bytes, hipMemcpyDeviceToHost), "D2H");
    return got;
  } catch (const std::exception& e) {
    fftfree::report("phase_infer_execute", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

void phase_free(void* handle) { delete static_cast<fftfree::PhaseCtx*>(handle); }

// ---- streaming (fft_cffi.cpp:1342-1742) -----------------------------------------------
// The sliding buffers and the queue live on the host, as the reference keeps them. A PCM
// push runs through fft_execute_batched (one launch; small pushes take its single-stream
// path); a spectrum push runs the inverse transform and the overlap-add onto the carried
// accumulator on the device and brings back only the overlap-added samples.
namespace fftfree {
namespace {

int stream_window(const FftContext* c) { return c->window > 0 ? c->window : c->N; }
int stream_hop(const FftContext* c) { return c->hop > 0 ? c->hop : stream_window(c); }

size_t stream_backlog(const FftContext* c) {
  return c->stream_offset >= c->stream_pcm.size() ? 0 : c->stream_pcm.size() - c->stream_offset;
}

size_t stream_frames(const FftContext* c) {
  const size_t backlog = stream_backlog(c);
  const int W = stream_window(c), H = stream_hop(c);
  if (W <= 0 || H <= 0 || backlog < (size_t)W) return 0;
  return 1 + (backlog - (size_t)W) / (size_t)H;
}

void stream_compact(FftContext* c) {
  if (c->stream_offset == 0) return;
  if (c->stream_offset >= c->stream_pcm.size()) {
    c->stream_pcm.clear();
  } else {
    c->stream_pcm.erase(c->stream_pcm.begin(), c->stream_pcm.begin() + (std::ptrdiff_t)c->stream_offset);
  }
  c->stream_offset = 0;
}

size_t stream_pending(const FftContext* c) {
  return c->stream_queue_offset >= c->stream_queue.size() ? 0 : c->stream_queue.size() - c->stream_queue_offset;
}

// fft_cffi.cpp:1420-1442: up to max_samples (0 = all) finished samples to out_pcm.
size_t stream_emit(FftContext* c, float* out_pcm, size_t max_samples) {
  size_t take = stream_pending(c);
  if (max_samples != 0 && take > max_samples) take = max_samples;
  if (!out_pcm || take == 0) return 0;
  std::copy_n(c->stream_queue.data() + c->stream_queue_offset, take, out_pcm);
  c->stream_queue_offset += take;
  if (c->stream_queue_offset >= c->stream_queue.size()) {
    c->stream_queue.clear();
  } else {
    c->stream_queue.erase(c->stream_queue.begin(), c->stream_queue.begin() + (std::ptrdiff_t)c->stream_queue_offset);
  }
  c->stream_queue_offset = 0;
  return take;
}

// cola_window_weight (fft_cffi.cpp:210-226).
float stream_cola_weight(const FftContext* c, size_t i) {
  if (!c->windows_enabled) return 1.0f;
  const float a = i < c->analysis_win.size() ? c->analysis_win[i] : 1.0f;
  const float b = i < c->synthesis_win.size() ? c->synthesis_win[i] : 1.0f;
  return a * b;
}

}  // namespace
}  // namespace fftfree

void fft_stream_reset(void* handle) {
  if (!handle) return;
  auto* c = static_cast<FftContext*>(handle);
  c->stream_pcm.clear();
  c->stream_offset = 0;
  c->stream_queue.clear();
  c->stream_queue_offset = 0;
  c->stream_accum.clear();
  c->stream_norm.clear();
}

size_t fft_stream_pending_frames(void* handle) {
  return handle ? fftfree::stream_frames(static_cast<FftContext*>(handle)) : 0;
}

size_t fft_stream_backlog_samples(void* handle) {
  return handle ? fftfree::stream_backlog(static_cast<FftContext*>(handle)) : 0;
}

size_t fft_stream_pending_pcm(void* handle) {
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  return c->inverse ? fftfree::stream_pending(c) : 0;
}

size_t fft_stream_push_pcm(void* handle, const float* pcm, size_t samples, float* out_real, float* out_imag,
                           float* out_mag, size_t max_frames, int flush_mode) {
  (void)flush_mode;  // not read by the reference either
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  try {
    if (samples > 0 && pcm) c->stream_pcm.insert(c->stream_pcm.end(), pcm, pcm + samples);
    const int W = fftfree::stream_window(c), H = fftfree::stream_hop(c);
    if (fftfree::verbose())
      std::fprintf(stderr, "[diag] fft_stream_push_pcm samples=%zu W=%d H=%d offset=%zu buffered=%zu\n", samples, W,
                   H, c->stream_offset, c->stream_pcm.size());
    if (W <= 0 || H <= 0) return 0;
    size_t frames = fftfree::stream_frames(c);
    if (frames == 0) return 0;
    if (max_frames != 0 && frames > max_frames) frames = max_frames;
    const size_t backlog = fftfree::stream_backlog(c);
    size_t len = (size_t)W + (frames - 1) * (size_t)H;
    if (len > backlog) len = backlog;
    // the frames of the next len samples (fft_cffi.cpp:1553-1590: the chunk the reference
    // copies is exactly this stretch of the buffer), partial frames dropped (pad_mode 2)
    const size_t bins = fftfree::bins_of(c);
    std::vector<float> local;
    if (!out_real || !out_imag || !out_mag) local.assign(3 * frames * bins, 0.0f);
    float* re = out_real ? out_real : local.data();
    float* im = out_imag ? out_imag : local.data() + frames * bins;
    float* mg = out_mag ? out_mag : local.data() + 2 * frames * bins;
    const size_t produced =
        fft_execute_batched(handle, c->stream_pcm.data() + c->stream_offset, len, re, im, mg, 2, 0, frames);
    if (produced == 0) return 0;
    c->stream_offset += produced * (size_t)H;
    fftfree::stream_compact(c);
    return produced;
  } catch (const std::exception& e) {
    fftfree::report("fft_stream_push_pcm", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_stream_push_spectrum(void* handle, const float* in_real, const float* in_imag, size_t frames,
                                float* out_pcm, size_t max_samples, int flush_mode) {
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (!c->inverse) return 0;
  try {
    const size_t n = (size_t)std::max(c->N, 1);
    if (c->stream_accum.size() != n) c->stream_accum.assign(n, 0.0f);
    const bool use_norm = c->cola_mode == FFT_COLA_NORMALIZE;
    if (use_norm) {
      if (c->stream_norm.size() != n) c->stream_norm.assign(n, 0.0f);
    } else {
      c->stream_norm.clear();
    }
    if (frames > 0) {
      // the reference's complex entry refuses a NULL plane: nothing is added
      if (!in_real || !in_imag) return fftfree::stream_emit(c, out_pcm, max_samples);
      // Per frame the reference adds the frame's N samples to the accumulator, queues the
      // first e = min(hop, N) of them (all N after the last frame of a final flush) and
      // shifts the accumulator by that much. On the device: the frames' inverse transform,
      // then the overlap-add of frame f at offset f*e onto the carried accumulator (the
      // same additions in the same order); the host divides by the COLA sums, queues the
      // first (F-1)*e + (e or N) samples and carries the rest.
      fftfree::ensure_gpu(c);
      const size_t bins = fftfree::bins_of(c);
      const size_t e = std::min((size_t)std::max(fftfree::stream_hop(c), 1), n);
      const size_t len = (frames - 1) * e + n;
      auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
      const size_t in_bytes = frames * bins * sizeof(float), n_bytes = n * sizeof(float);
      // device staging: what goes up (planes, carried accumulator and norm, weights), the
      // frames, then what comes down (overlap-add and norm sums), each run contiguous
      const size_t o_im = up(in_bytes), o_carry = o_im + up(in_bytes), o_cnorm = o_carry + up(n_bytes),
                   o_w = o_cnorm + up(n_bytes), o_fr = o_w + up(n_bytes),
                   o_ext = o_fr + up(frames * n * sizeof(float)), o_norm = o_ext + up(len * sizeof(float)),
                   o_end = use_norm ? o_norm + up(len * sizeof(float)) : o_norm;
      c->d_planes.reserve(o_end);  // (complex_batched's own scratch is d_cplx)
      char* base = c->d_planes.as<char>();
      auto at = [&](size_t off) { return reinterpret_cast<float*>(base + off); };
      hipStream_t st = c->stream;
      // small pushes go up in one copy from pinned memory (a copy's fixed cost is most of
      // their time); large ones copy the planes straight from the caller
      const bool small = o_fr <= (size_t(1) << 20);
      const size_t h0 = small ? 0 : o_carry;
      c->h_in.reserve(o_fr - h0);
      auto hin = [&](size_t off) { return c->h_in.as<char>() + (off - h0); };  // by device offset
      if (small) {
        std::memcpy(hin(0), in_real, in_bytes);
        std::memcpy(hin(o_im), in_imag, in_bytes);
      }
      std::memcpy(hin(o_carry), c->stream_accum.data(), n_bytes);
      if (use_norm) {
        std::memcpy(hin(o_cnorm), c->stream_norm.data(), n_bytes);
        float* w = reinterpret_cast<float*>(hin(o_w));
        for (size_t i = 0; i < n; ++i) w[i] = fftfree::stream_cola_weight(c, i);
      }
      fftfree::hip_check(hipMemcpyAsync(base + h0, hin(h0), o_fr - h0, hipMemcpyHostToDevice, st), "H2D stream");
      if (!small) {
        fftfree::hip_check(hipMemcpyAsync(at(0), in_real, in_bytes, hipMemcpyHostToDevice, st), "H2D re");
        fftfree::hip_check(hipMemcpyAsync(at(o_im), in_imag, in_bytes, hipMemcpyHostToDevice, st), "H2D im");
      }
      fftfree::complex_batched(c, at(0), at(o_im), frames, at(o_fr), st);
      fftfree::hip_check(fftfree::launch_stream_ola(at(o_fr), frames, (int)n, (int)e, at(o_carry), at(o_cnorm),
                                                    use_norm ? at(o_w) : nullptr, at(o_ext), at(o_norm), len, st),
                         "launch_stream_ola");
      c->h_out.reserve(o_end - o_ext);
      fftfree::hip_check(hipMemcpyAsync(c->h_out.ptr, base + o_ext, o_end - o_ext, hipMemcpyDeviceToHost, st),
                         "D2H overlap-add");
      fftfree::hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
      const float* ext = c->h_out.as<float>();
      const float* en = use_norm ? c->h_out.as<float>() + (o_norm - o_ext) / sizeof(float) : nullptr;
      const size_t out_n = (frames - 1) * e + (flush_mode == FFT_STREAM_FLUSH_FINAL ? n : e);
      const size_t q0 = c->stream_queue.size();
      c->stream_queue.resize(q0 + out_n);
      float* q = c->stream_queue.data() + q0;
      for (size_t t = 0; t < out_n; ++t) {
        float v = ext[t];
        if (use_norm && std::fabs(en[t]) > 1e-12f) v /= en[t];
        q[t] = v;
      }
      for (size_t i = 0; i < n; ++i) {  // what is left: len - out_n <= n - e samples
        const size_t t = out_n + i;
        c->stream_accum[i] = t < len ? ext[t] : 0.0f;
        if (use_norm) c->stream_norm[i] = t < len ? en[t] : 0.0f;
      }
    }
    if (flush_mode == FFT_STREAM_FLUSH_FINAL && frames == 0 && !c->stream_accum.empty()) {
      const bool norm_local = !c->stream_norm.empty();
      for (size_t i = 0; i < c->stream_accum.size(); ++i) {
        float v = c->stream_accum[i];
        if (norm_local && std::fabs(c->stream_norm[i]) > 1e-12f) v /= c->stream_norm[i];
        c->stream_queue.push_back(v);
      }
      std::fill(c->stream_accum.begin(), c->stream_accum.end(), 0.0f);
      if (norm_local) std::fill(c->stream_norm.begin(), c->stream_norm.end(), 0.0f);
    }
    return fftfree::stream_emit(c, out_pcm, max_samples);
  } catch (const std::exception& e) {
    fftfree::report("fft_stream_push_spectrum", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

size_t fft_ctx_size(void* handle) { return handle ? (size_t) static_cast<FftContext*>(handle)->N : 0; }

size_t fft_ctx_worker_threads(void* handle) {
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  int t = c->cfg.threads > 0 ? c->cfg.threads : 1;
  if (c->has_pool) {
    const int hw = (int)std::thread::hardware_concurrency();
    if (hw > 0) t = std::min(t, hw);
  }
  return (size_t)t;
}

size_t fft_ctx_effective_threads(void* handle) {
  // eigen_fft.hpp:2332-2342: min(requested, hardware, 16)
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  int lim = c->cfg.threads > 0 ? std::min(c->cfg.threads, 16) : 4;
  const int hw = (int)std::thread::hardware_concurrency();
  if (hw > 0) lim = std::min(lim, hw);
  return (size_t)std::max(1, lim);
}

void fft_free(void* handle) {
  if (!handle) return;
  delete static_cast<FftContext*>(handle);
}

int fft_config_window(void* handle, int analysis_kind, int synthesis_kind, float param1, float param2,
                      int norm_policy, int cola_mode) {
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  const int W = c->window > 0 ? c->window : c->N;
  if (W <= 0) return 0;
  try {
    c->analysis_win_kind = analysis_kind;
    c->synthesis_win_kind = synthesis_kind;
    c->win_param1 = param1;
    c->win_param2 = param2;
    c->win_norm_policy = norm_policy;
    c->cola_mode = cola_mode;
    c->analysis_win = fftfree::build_window(analysis_kind, W, param1, param2);
    c->synthesis_win = fftfree::build_window(synthesis_kind, W, param1, param2);
    fftfree::normalize_window(c->analysis_win, norm_policy);
    fftfree::normalize_window(c->synthesis_win, norm_policy);
    c->windows_enabled = false;  // fft_cffi.cpp:2005-2006
    c->windows_dirty = true;
    return 1;
  } catch (...) {
    return 0;
  }
}

int fft_config_window_custom(void* handle, const float* analysis_window, size_t analysis_len,
                             const float* synthesis_window, size_t synthesis_len, int norm_policy, int cola_mode) {
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  const int W = c->window > 0 ? c->window : c->N;
  if (W <= 0) return 0;
  try {
    if (analysis_window && analysis_len == (size_t)W) c->analysis_win.assign(analysis_window, analysis_window + W);
    if (synthesis_window && synthesis_len == (size_t)W) c->synthesis_win.assign(synthesis_window, synthesis_window + W);
    c->win_norm_policy = norm_policy;
    c->cola_mode = cola_mode;
    fftfree::normalize_window(c->analysis_win, norm_policy);
    fftfree::normalize_window(c->synthesis_win, norm_policy);
    c->windows_enabled = false;
    c->windows_dirty = true;
    return 1;
  } catch (...) {
    return 0;
  }
}

void fft_set_workerpool_debug_dropout(int) {}
void fft_set_workerpool_dropout_pattern(int, int, int) {}
void fft_set_workerpool_dropout_single(long long) {}
void fft_set_workerpool_dropout_persistent(int) {}
void fft_clear_workerpool_dropout_history(void) {}
void fft_register_restore_observer(fft_restore_observer_t) {}

size_t fft_plan_metadata(void* handle, int kind, float* out, size_t capacity) {
  if (!handle || !out) return 0;
  auto* c = static_cast<FftContext*>(handle);
  try {
    fftfree::ensure_gpu(c);
    const bool r2c_half = c->transform == FFT_TRANSFORM_R2C && c->half_spectrum;
    // capacity bound for the device scratch: N*lgN entries at most
    size_t need = 0;
    fftfree::hip_check(fftfree::launch_metadata(c->N, fftfree::use_stockham(c) ? 2 : 0, kind, r2c_half ? 1 : 0,
                                                c->plan->W, c->inverse ? 1 : 0, nullptr, 0, &need, c->stream),
                       "metadata size");
    if (need == 0 || need > capacity) return 0;
    c->d_cplx.reserve(need * sizeof(float2));
    size_t got = 0;
    fftfree::hip_check(fftfree::launch_metadata(c->N, fftfree::use_stockham(c) ? 2 : 0, kind, r2c_half ? 1 : 0,
                                                c->plan->W, c->inverse ? 1 : 0, c->d_cplx.as<float2>(), need, &got,
                                                c->stream),
                       "launch_metadata");
    fftfree::hip_check(hipMemcpyAsync(out, c->d_cplx.ptr, got * sizeof(float2), hipMemcpyDeviceToHost, c->stream),
                       "D2H");
    fftfree::hip_check(hipStreamSynchronize(c->stream), "sync");
    return got;
  } catch (const std::exception& e) {
    fftfree::report("fft_plan_metadata", e.what());
    return 0;
  } catch (...) {
    return 0;
  }
}

int fft_set_variant(void* handle, int variant) {
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  if (variant < 0) {
    c->plan->stft_variant = -1;
    c->plan->c2c_variant = -1;
    return 1;
  }
  if (variant >= fftfree::variant_count(c)) return 0;
  if (fftfree::stft_variants(c))
    c->plan->stft_variant = variant;
  else
    c->plan->c2c_variant = variant;
  return 1;
}

int fft_variant_count(void* handle) {
  if (!handle) return 0;
  auto* c = static_cast<FftContext*>(handle);
  return fftfree::variant_count(c);
}

int fft_variant_name(void* handle, int variant, char* buf, size_t buflen) {
  if (!handle || !buf || buflen == 0) return 0;
  return fftfree::variant_desc(static_cast<FftContext*>(handle), variant, buf, buflen) ? 1 : 0;
}

int fft_variant_default(void* handle, size_t frames) {
  if (!handle) return -1;
  auto* c = static_cast<FftContext*>(handle);
  if (fftfree::stft_variants(c)) {
    if (c->plan->stft_variant >= 0) return c->plan->stft_variant;
    if (c->N < 16) return 0;  // the direct-DFT kernel
    fftfree::StftArgs a;
    a.n = c->N;
    a.frames = frames;
    return fftfree::default_stft_variant(a);
  }
  return c->plan->c2c_variant >= 0 ? c->plan->c2c_variant : 0;
}

void* fft_ragged_create(fft_ragged_request* requests, size_t count, int pad_mode, int reduce_magnitude,
                        int store_polar, int apply_windows, int window_kind) {
  if (!requests || count == 0) return nullptr;
  std::unique_ptr<fftfree::RaggedPlan> p(new fftfree::RaggedPlan());
  try {
    std::map<std::pair<int, int>, int> ctx_of;
    for (size_t i = 0; i < count; ++i) {
      fft_ragged_request& q = requests[i];
      const int n = q.window;
      if (!q.pcm || q.pcm_len == 0 || q.hop <= 0 || (n != 256 && n != 512 && n != 1024 && n != 2048 && n != 4096))
        return nullptr;
      const auto key = std::make_pair(n, q.hop);
      if (!ctx_of.count(key)) {
        void* h = fft_init_full_v2((size_t)n, 1, 1, 0, 2, 0, nullptr, 0, pad_mode, n, q.hop, FFT_TRANSFORM_R2C,
                                   reduce_magnitude, store_polar, 1, 0, 0, 0, 0, 1, apply_windows, 0, window_kind,
                                   0.0f, 0.0f, window_kind, 0.0f, 0.0f, FFT_WINDOW_NORM_NONE, FFT_COLA_OFF);
        if (!h) return nullptr;
        ctx_of[key] = (int)p->ctxs.size();
        p->ctxs.push_back(h);
      }
      p->req_ctx.push_back(ctx_of[key]);
    }
    p->reqs.assign(requests, requests + count);
    for (size_t i = 0; i < count; ++i) {
      auto* c = static_cast<FftContext*>(p->ctxs[p->req_ctx[i]]);
      fftfree::ensure_gpu(c);
      fft_ragged_request& q = p->reqs[i];
      q.frames = fftfree::frame_count(c, q.pcm_len, 0, 0);
      requests[i].frames = q.frames;
      fftfree::StftArgs a = fftfree::stft_args(c, q.pcm, q.pcm_len);
      a.frames = q.frames;
      a.o_re = q.out_real;
      a.o_im = q.out_imag;
      a.o_mag = q.out_mag;
      p->args.push_back(a);
      p->reqs_of_n[q.window].push_back(i);
    }
    p->mode = store_polar ? 3 : reduce_magnitude ? 2 : 0;
    fftfree::hip_check(hipGetDevice(&p->device), "hipGetDevice");
    fftfree::hip_check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
    // defaults before tuning: the line-aligned engine where it takes every request of N,
    // else the E=8 tile shape; 2 blocks per item
    for (auto& kv : p->reqs_of_n) {
      const int n = kv.first;
      int pick = -1;
      for (int cfg = 0; cfg < fftfree::ragged_cfg_count(); ++cfg) {
        if (2 * fftfree::ragged_cfg(cfg)->l != n) continue;
        if (fftfree::ragged_cfg_lines(cfg) ? p->lines_ok(n, cfg) : fftfree::ragged_cfg(cfg)->e == 8) pick = cfg;
        if (pick >= 0 && fftfree::ragged_cfg_lines(pick)) break;
      }
      p->cfg_of_n[n] = pick;
      p->tiles_of_n[n] = 2;
    }
    p->commit();
    return p.release();
  } catch (const std::exception& e) {
    fftfree::report("fft_ragged_create", e.what());
    return nullptr;
  } catch (...) {
    return nullptr;
  }
}

int fft_ragged_autotune(void* ragged, int repeats) {
  if (!ragged) return 0;
  auto* p = static_cast<fftfree::RaggedPlan*>(ragged);
  try {
    fftfree::hip_check(hipSetDevice(p->device), "hipSetDevice");
    p->tune(repeats > 0 ? repeats : 3);
    p->commit();
    return 1;
  } catch (const std::exception& e) {
    fftfree::report("fft_ragged_autotune", e.what());
    return 0;
  }
}

int fft_ragged_set_engine(void* ragged, int n, int engine, int blocks_per_item) {
  if (!ragged || blocks_per_item < 1) return 0;
  auto* p = static_cast<fftfree::RaggedPlan*>(ragged);
  const fftfree::VariantInfo* v = fftfree::ragged_cfg(engine);
  if (!v || 2 * v->l != n || !p->reqs_of_n.count(n)) return 0;
  try {
    fftfree::hip_check(hipSetDevice(p->device), "hipSetDevice");
    if (fftfree::ragged_cfg_lines(engine) && !p->lines_ok(n, engine)) return 0;
    p->cfg_of_n[n] = engine;
    p->tiles_of_n[n] = blocks_per_item;
    p->commit();
    return 1;
  } catch (const std::exception& e) {
    fftfree::report("fft_ragged_set_engine", e.what());
    return 0;
  }
}

int fft_ragged_engine_count(void) { return fftfree::ragged_cfg_count(); }

int fft_ragged_execute(void* ragged, void* stream) {
  if (!ragged) return 0;
  auto* p = static_cast<fftfree::RaggedPlan*>(ragged);
  return hipGraphLaunch(p->exec, static_cast<hipStream_t>(stream)) == hipSuccess ? 1 : 0;
}

int fft_ragged_execute_plain(void* ragged, void* stream) {
  if (!ragged) return 0;
  auto* p = static_cast<fftfree::RaggedPlan*>(ragged);
  try {
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (size_t i = 0; i < p->args.size(); ++i) {
      auto* c = static_cast<FftContext*>(p->ctxs[p->req_ctx[i]]);
      fftfree::hip_check(fftfree::launch_stft(p->args[i], 0, c->plan->stft_variant, s), "launch_stft");
    }
    return 1;
  } catch (const std::exception& e) {
    fftfree::report("fft_ragged_execute_plain", e.what());
    return 0;
  }
}

int fft_ragged_describe(void* ragged, char* buf, size_t buflen) {
  if (!ragged || !buf || buflen == 0) return 0;
  auto* p = static_cast<fftfree::RaggedPlan*>(ragged);
  std::string out = "launches=" + std::to_string(p->launches.size()) + " items=" + std::to_string(p->n_items);
  for (auto& kv : p->cfg_of_n) {
    const fftfree::VariantInfo* v = fftfree::ragged_cfg(kv.second);
    const bool flat = fftfree::ragged_cfg_flat(kv.second);
    int g = 0;  // the staged launch's grid
    for (const auto& l : p->launches)
      if (l.cfg == kv.second) g = l.grid;
    out += " N" + std::to_string(kv.first) + ":" + v->radices + "/fpb" + std::to_string(v->fpw) + "/x" +
           std::to_string(flat ? 1 : p->tiles_of_n[kv.first]) + "/grid" + std::to_string(g);
  }
  std::snprintf(buf, buflen, "%s", out.c_str());
  return 1;
}

void fft_ragged_free(void* ragged) { delete static_cast<fftfree::RaggedPlan*>(ragged); }

}  // extern "C"

// ft_cffi.cpp -- ft_apply_complex_pipeline / ft_apply_polar_pipeline (ft_tools_cffi.cpp:76-170)
// and their device-pointer forms: the configuration mapping and the order of the steps as
// the reference runs them (run_pipeline, ft_tools.hpp:463-477) over the launches of
// ft_tools.hip.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/ft_tools_cffi.h"
#include "ft_kernels.hpp"

namespace {

struct Step {
  bool ok = true;
  void check(hipError_t e) {
    if (e != hipSuccess) ok = false;
  }
};

// The settings as the reference maps them (to_grid_spec / to_plan / configure_*,
// ft_tools_cffi.cpp:18-58), with the shapes through the pipeline.
struct Settings {
  double f_origin = 0.0, f_step = 1.0;
  int ut = 1, uf = 1, dt = 1, df = 1, preserve = 1;
  bool use_q = false, use_c = false;
  fftfree::FtQuantizer q;
  fftfree::FtContrast c;
  long long t1 = 0, f1 = 0;  // after upsampling
  bool up() const { return ut > 1 || uf > 1; }
  bool down() const { return dt > 1 || df > 1; }
};

// False when the grid is empty or too large, or the result would have another shape (the
// reference computes that result and then returns 0 without writing,
// ft_tools_cffi.cpp:111-113).
bool make_settings(const ft_grid_spec* spec, const ft_resample_plan_c* plan, const ft_quantizer_config* quant,
                   const ft_contrast_config* contrast, int tb, int fb, Settings& st) {
  if (tb <= 0 || fb <= 0 || (long long)tb * (long long)fb >= (long long)INT_MAX) return false;
  st.f_origin = spec ? spec->freq_origin : 0.0;
  st.f_step = (spec && spec->freq_step != 0.0) ? spec->freq_step : 1.0;
  st.ut = (plan && plan->upsample_time > 0) ? plan->upsample_time : 1;
  st.uf = (plan && plan->upsample_freq > 0) ? plan->upsample_freq : 1;
  st.dt = (plan && plan->downsample_time > 0) ? plan->downsample_time : 1;
  st.df = (plan && plan->downsample_freq > 0) ? plan->downsample_freq : 1;
  st.preserve = plan ? (plan->preserve_energy != 0) : 1;
  st.use_q = quant && quant->enabled != 0;
  st.use_c = contrast && contrast->enabled != 0;
  if (st.use_q) {
    st.q.base_frequency = quant->base_frequency;
    st.q.magnitude_step = quant->magnitude_step;
    st.q.mix = quant->mix;
    st.q.distance_decay = quant->distance_decay;
  }
  st.q.inv_base = 1.0f / st.q.base_frequency;  // used only where base > 0 / step > 0
  st.q.inv_step = 1.0f / st.q.magnitude_step;
  if (st.use_c) {
    st.c.kernel_time = contrast->kernel_time;
    st.c.kernel_freq = contrast->kernel_freq;
    st.c.brightness = contrast->brightness;
    st.c.contrast = contrast->contrast;
    st.c.gamma = contrast->gamma;
    st.c.clamp_zero = contrast->clamp_zero != 0;
  }
  st.t1 = st.up() ? (long long)tb * st.ut : tb;
  st.f1 = st.up() ? (long long)fb * st.uf : fb;
  const long long t2 = st.down() ? (st.t1 + st.dt - 1) / st.dt : st.t1;
  const long long f2 = st.down() ? (st.f1 + st.df - 1) / st.df : st.f1;
  if (t2 != tb || f2 != fb) return false;
  return st.t1 * st.f1 < (long long)INT_MAX;
}

// The contrast box reads neighbouring cells of the magnitude (polar) or of both planes
// (complex): the one-pass kernel cannot then write over them.
bool box_reads_output(bool polar, const Settings& st, const float* in_a, const float* in_b, const float* out_a,
                      const float* out_b) {
  return st.use_c && (out_a == in_a || (!polar && out_b == in_b));
}

// The pipeline from device planes (in_a, in_b) to device planes (out_a, out_b) on stream
// s. Without resampling and with a local (or no) contrast box it is one launch; otherwise
// a stream-ordered workspace carries the complex grid through one launch per step.
bool run_device(bool polar, const Settings& st, const float* in_a, const float* in_b, float* out_a, float* out_b,
                int tb, int fb, hipStream_t s) {
  const size_t n0 = (size_t)tb * fb, n1 = (size_t)(st.t1 * st.f1);
  const bool q_on = st.use_q && st.q.mix > 0.0f;
  if (!st.up() && !st.down() && fftfree::ft_fused_supported(st.use_c, st.c) && !box_reads_output(polar, st, in_a, in_b,
                                                                                                   out_a, out_b))
    return fftfree::ft_fused(polar, in_a, in_b, out_a, out_b, tb, fb, st.q, q_on ? 1 : 0, st.f_origin, st.f_step,
                             st.c, st.use_c ? 1 : 0, s) == hipSuccess;
  // workspace: two complex grids, magnitude and phase planes, contrast scratch
  const size_t cz = n1 * sizeof(float2), cf = n1 * sizeof(float);
  auto round_up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t bytes = 2 * round_up(cz) + 2 * round_up(cf) + round_up((2 * n1 + 1) * sizeof(float));
  void* base = nullptr;
  if (hipMallocAsync(&base, bytes, s) != hipSuccess) return false;
  char* p = static_cast<char*>(base);
  auto take = [&](size_t b) {
    char* r = p;
    p += round_up(b);
    return r;
  };
  float2* za = reinterpret_cast<float2*>(take(cz));
  float2* zb = reinterpret_cast<float2*>(take(cz));
  float* ma = reinterpret_cast<float*>(take(cf));
  float* pa = reinterpret_cast<float*>(take(cf));
  float* scratch = reinterpret_cast<float*>(take((2 * n1 + 1) * sizeof(float)));
  Step ok;
  const double f_step = st.up() ? st.f_step / (double)st.uf : st.f_step;
  const int t1 = (int)st.t1, f1 = (int)st.f1;
  if (!polar) {
    ok.check(fftfree::ft_planes_to_complex(in_a, in_b, n0, za, s));
    float2* cur = za;
    float2* other = zb;
    if (st.up()) {
      ok.check(fftfree::ft_upsample(cur, tb, fb, st.ut, st.uf, other, s));
      std::swap(cur, other);
    }
    if (q_on) ok.check(fftfree::ft_quantize(cur, nullptr, t1, f1, st.q, st.f_origin, f_step, s));
    if (st.use_c) ok.check(fftfree::ft_contrast(cur, nullptr, t1, f1, st.c, scratch, s));
    if (st.down()) {
      ok.check(fftfree::ft_downsample(cur, t1, f1, st.dt, st.df, st.preserve, other, s));
      std::swap(cur, other);
    }
    ok.check(fftfree::ft_complex_to_planes(cur, n0, out_a, out_b, s));
  } else {
    // polar grid: resampling goes through complex values (upsample_polar /
    // downsample_polar, ft_tools.hpp:229-245)
    if (st.up()) {
      ok.check(fftfree::ft_polar_to_complex(in_a, in_b, n0, za, s));
      ok.check(fftfree::ft_upsample(za, tb, fb, st.ut, st.uf, zb, s));
      ok.check(fftfree::ft_complex_to_polar(zb, n1, ma, pa, s));
    } else {
      ok.check(hipMemcpyAsync(ma, in_a, n0 * sizeof(float), hipMemcpyDeviceToDevice, s));
      ok.check(hipMemcpyAsync(pa, in_b, n0 * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    if (q_on) ok.check(fftfree::ft_quantize(nullptr, ma, t1, f1, st.q, st.f_origin, f_step, s));
    if (st.use_c) ok.check(fftfree::ft_contrast(nullptr, ma, t1, f1, st.c, scratch, s));
    if (st.down()) {
      ok.check(fftfree::ft_polar_to_complex(ma, pa, n1, za, s));
      ok.check(fftfree::ft_downsample(za, t1, f1, st.dt, st.df, st.preserve, zb, s));
      ok.check(fftfree::ft_complex_to_polar(zb, n0, out_a, out_b, s));
    } else {
      ok.check(hipMemcpyAsync(out_a, ma, n0 * sizeof(float), hipMemcpyDeviceToDevice, s));
      ok.check(hipMemcpyAsync(out_b, pa, n0 * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
  }
  ok.check(hipFreeAsync(base, s));
  return ok.ok;
}

// Host planes, changed in place: one device allocation holds the input and output planes.
int pipeline_host(bool polar, const ft_grid_spec* spec, const ft_resample_plan_c* plan,
                  const ft_quantizer_config* quant, const ft_contrast_config* contrast, float* a, float* b, int tb,
                  int fb) {
  Settings st;
  if (!a || !b || !make_settings(spec, plan, quant, contrast, tb, fb, st)) return 0;
  const size_t n0 = (size_t)tb * fb, plane = ((n0 * sizeof(float)) + 255) & ~size_t(255);
  void* base = nullptr;
  if (hipMalloc(&base, 4 * plane) != hipSuccess) return 0;
  char* p = static_cast<char*>(base);
  float* d_in_a = reinterpret_cast<float*>(p);
  float* d_in_b = reinterpret_cast<float*>(p + plane);
  float* d_out_a = reinterpret_cast<float*>(p + 2 * plane);
  float* d_out_b = reinterpret_cast<float*>(p + 3 * plane);
  hipStream_t s = nullptr;
  Step ok;
  ok.check(hipMemcpy(d_in_a, a, n0 * sizeof(float), hipMemcpyHostToDevice));
  ok.check(hipMemcpy(d_in_b, b, n0 * sizeof(float), hipMemcpyHostToDevice));
  if (ok.ok && !run_device(polar, st, d_in_a, d_in_b, d_out_a, d_out_b, tb, fb, s)) ok.ok = false;
  if (ok.ok) ok.check(hipStreamSynchronize(s));
  // the planes are written only once the whole result is there
  std::vector<float> ha, hb;
  if (ok.ok) {
    ha.resize(n0);
    hb.resize(n0);
    ok.check(hipMemcpy(ha.data(), d_out_a, n0 * sizeof(float), hipMemcpyDeviceToHost));
    ok.check(hipMemcpy(hb.data(), d_out_b, n0 * sizeof(float), hipMemcpyDeviceToHost));
  }
  (void)hipFree(base);
  if (!ok.ok) return 0;
  std::memcpy(a, ha.data(), n0 * sizeof(float));
  std::memcpy(b, hb.data(), n0 * sizeof(float));
  return 1;
}

int pipeline_device(bool polar, const ft_grid_spec* spec, const ft_resample_plan_c* plan,
                    const ft_quantizer_config* quant, const ft_contrast_config* contrast, const float* in_a,
                    const float* in_b, float* out_a, float* out_b, int tb, int fb, void* stream) {
  Settings st;
  if (!in_a || !in_b || !out_a || !out_b || !make_settings(spec, plan, quant, contrast, tb, fb, st)) return 0;
  if (box_reads_output(polar, st, in_a, in_b, out_a, out_b) && !st.up() && !st.down() &&
      fftfree::ft_fused_supported(1, st.c)) {
    // the contrast box reads neighbours: run the one-pass kernel into a stream-ordered copy
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t n0 = (size_t)tb * fb, plane = ((n0 * sizeof(float)) + 255) & ~size_t(255);
    void* base = nullptr;
    if (hipMallocAsync(&base, 2 * plane, s) != hipSuccess) return 0;
    float* ta = static_cast<float*>(base);
    float* tbuf = reinterpret_cast<float*>(static_cast<char*>(base) + plane);
    Step ok;
    ok.check(hipMemcpyAsync(ta, in_a, n0 * sizeof(float), hipMemcpyDeviceToDevice, s));
    ok.check(hipMemcpyAsync(tbuf, in_b, n0 * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (ok.ok && !run_device(polar, st, ta, tbuf, out_a, out_b, tb, fb, s)) ok.ok = false;
    ok.check(hipFreeAsync(base, s));
    return ok.ok ? 1 : 0;
  }
  return run_device(polar, st, in_a, in_b, out_a, out_b, tb, fb, static_cast<hipStream_t>(stream)) ? 1 : 0;
}

}  // namespace

extern "C" {

int ft_apply_complex_pipeline(const ft_grid_spec* spec, const ft_resample_plan_c* plan, const ft_quantizer_config* quant,
                              const ft_contrast_config* contrast, float* real, float* imag, int time_bins,
                              int freq_bins) {
  try {
    return pipeline_host(false, spec, plan, quant, contrast, real, imag, time_bins, freq_bins);
  } catch (...) {
    return 0;
  }
}

int ft_apply_polar_pipeline(const ft_grid_spec* spec, const ft_resample_plan_c* plan, const ft_quantizer_config* quant,
                            const ft_contrast_config* contrast, float* magnitude, float* phase, int time_bins,
                            int freq_bins) {
  try {
    return pipeline_host(true, spec, plan, quant, contrast, magnitude, phase, time_bins, freq_bins);
  } catch (...) {
    return 0;
  }
}

int ft_apply_complex_pipeline_device(const ft_grid_spec* spec, const ft_resample_plan_c* plan,
                                     const ft_quantizer_config* quant, const ft_contrast_config* contrast,
                                     const float* d_real, const float* d_imag, float* d_out_real, float* d_out_imag,
                                     int time_bins, int freq_bins, void* stream) {
  try {
    return pipeline_device(false, spec, plan, quant, contrast, d_real, d_imag, d_out_real, d_out_imag, time_bins,
                           freq_bins, stream);
  } catch (...) {
    return 0;
  }
}

int ft_apply_polar_pipeline_device(const ft_grid_spec* spec, const ft_resample_plan_c* plan,
                                   const ft_quantizer_config* quant, const ft_contrast_config* contrast,
                                   const float* d_magnitude, const float* d_phase, float* d_out_magnitude,
                                   float* d_out_phase, int time_bins, int freq_bins, void* stream) {
  try {
    return pipeline_device(true, spec, plan, quant, contrast, d_magnitude, d_phase, d_out_magnitude, d_out_phase,
                           time_bins, freq_bins, stream);
  } catch (...) {
    return 0;
  }
}

}  // extern "C"

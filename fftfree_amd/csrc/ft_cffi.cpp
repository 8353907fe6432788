// ft_cffi.cpp -- ft_apply_complex_pipeline / ft_apply_polar_pipeline (ft_tools_cffi.cpp:76-170):
// the configuration mapping and the order of the steps as the reference runs them
// (run_pipeline, ft_tools.hpp:463-477), each step a launch of ft_tools.hip.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdio>
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

// Everything the pipeline needs, in one device allocation.
struct Buffers {
  void* base = nullptr;
  float2 *za = nullptr, *zb = nullptr;
  float *ma = nullptr, *pa = nullptr, *scratch = nullptr;
  bool alloc(size_t n_in, size_t n_big) {
    const size_t c = n_big * sizeof(float2), f = n_big * sizeof(float);
    const size_t bytes = 2 * c + 2 * f + (2 * n_big + 1) * sizeof(float) + 5 * 256;
    if (hipMalloc(&base, bytes) != hipSuccess) return false;
    char* p = static_cast<char*>(base);
    auto take = [&](size_t b) {
      char* r = p;
      p += (b + 255) & ~size_t(255);
      return r;
    };
    za = reinterpret_cast<float2*>(take(c));
    zb = reinterpret_cast<float2*>(take(c));
    ma = reinterpret_cast<float*>(take(f));
    pa = reinterpret_cast<float*>(take(f));
    scratch = reinterpret_cast<float*>(take((2 * n_big + 1) * sizeof(float)));
    (void)n_in;
    return true;
  }
  ~Buffers() {
    if (base) (void)hipFree(base);
  }
};

int pipeline(bool polar, const ft_grid_spec* spec, const ft_resample_plan_c* plan, const ft_quantizer_config* quant,
             const ft_contrast_config* contrast, float* a, float* b, int tb, int fb) {
  if (!a || !b || tb <= 0 || fb <= 0 || (long long)tb * (long long)fb >= (long long)INT_MAX) return 0;
  // to_grid_spec / to_plan / configure_* (ft_tools_cffi.cpp:18-58)
  const double f_origin = spec ? spec->freq_origin : 0.0;
  double f_step = (spec && spec->freq_step != 0.0) ? spec->freq_step : 1.0;
  const int ut = (plan && plan->upsample_time > 0) ? plan->upsample_time : 1;
  const int uf = (plan && plan->upsample_freq > 0) ? plan->upsample_freq : 1;
  const int dt = (plan && plan->downsample_time > 0) ? plan->downsample_time : 1;
  const int df = (plan && plan->downsample_freq > 0) ? plan->downsample_freq : 1;
  const int preserve = plan ? (plan->preserve_energy != 0) : 1;
  const bool use_q = quant && quant->enabled != 0;
  const bool use_c = contrast && contrast->enabled != 0;
  fftfree::FtQuantizer q;
  if (use_q) {
    q.base_frequency = quant->base_frequency;
    q.magnitude_step = quant->magnitude_step;
    q.mix = quant->mix;
    q.distance_decay = quant->distance_decay;
  }
  fftfree::FtContrast c;
  if (use_c) {
    c.kernel_time = contrast->kernel_time;
    c.kernel_freq = contrast->kernel_freq;
    c.brightness = contrast->brightness;
    c.contrast = contrast->contrast;
    c.gamma = contrast->gamma;
    c.clamp_zero = contrast->clamp_zero != 0;
  }
  // shapes through the pipeline; a result of another shape is refused (the reference
  // computes it and then returns 0 without writing, ft_tools_cffi.cpp:111-113)
  const bool up = ut > 1 || uf > 1, down = dt > 1 || df > 1;
  const long long t1 = up ? (long long)tb * ut : tb, f1 = up ? (long long)fb * uf : fb;
  const long long t2 = down ? (t1 + dt - 1) / dt : t1, f2 = down ? (f1 + df - 1) / df : f1;
  if (t2 != tb || f2 != fb) return 0;
  if (t1 * f1 >= (long long)INT_MAX) return 0;
  const size_t n0 = (size_t)tb * fb, n1 = (size_t)(t1 * f1);
  Buffers d;
  if (!d.alloc(n0, n1)) return 0;
  Step st;
  hipStream_t s = nullptr;
  if (!polar) {
    std::vector<float2> z(n0);
    for (size_t i = 0; i < n0; ++i) z[i] = make_float2(a[i], b[i]);
    st.check(hipMemcpy(d.za, z.data(), n0 * sizeof(float2), hipMemcpyHostToDevice));
    float2* cur = d.za;
    float2* other = d.zb;
    if (up) {
      st.check(fftfree::ft_upsample(cur, tb, fb, ut, uf, other, s));
      std::swap(cur, other);
      f_step /= (double)uf;
    }
    if (use_q && q.mix > 0.0f) st.check(fftfree::ft_quantize(cur, nullptr, (int)t1, (int)f1, q, f_origin, f_step, s));
    if (use_c) st.check(fftfree::ft_contrast(cur, nullptr, (int)t1, (int)f1, c, d.scratch, s));
    if (down) {
      st.check(fftfree::ft_downsample(cur, (int)t1, (int)f1, dt, df, preserve, other, s));
      std::swap(cur, other);
    }
    if (!st.ok) return 0;
    st.check(hipMemcpy(z.data(), cur, n0 * sizeof(float2), hipMemcpyDeviceToHost));
    if (!st.ok) return 0;
    for (size_t i = 0; i < n0; ++i) {
      a[i] = z[i].x;
      b[i] = z[i].y;
    }
    return 1;
  }
  // polar grid: magnitude and phase planes; resampling goes through complex values
  // (upsample_polar / downsample_polar, ft_tools.hpp:229-245)
  st.check(hipMemcpy(d.ma, a, n0 * sizeof(float), hipMemcpyHostToDevice));
  st.check(hipMemcpy(d.pa, b, n0 * sizeof(float), hipMemcpyHostToDevice));
  if (up) {
    st.check(fftfree::ft_polar_to_complex(d.ma, d.pa, n0, d.za, s));
    st.check(fftfree::ft_upsample(d.za, tb, fb, ut, uf, d.zb, s));
    st.check(fftfree::ft_complex_to_polar(d.zb, n1, d.ma, d.pa, s));
    f_step /= (double)uf;
  }
  if (use_q && q.mix > 0.0f) st.check(fftfree::ft_quantize(nullptr, d.ma, (int)t1, (int)f1, q, f_origin, f_step, s));
  if (use_c) st.check(fftfree::ft_contrast(nullptr, d.ma, (int)t1, (int)f1, c, d.scratch, s));
  if (down) {
    st.check(fftfree::ft_polar_to_complex(d.ma, d.pa, n1, d.za, s));
    st.check(fftfree::ft_downsample(d.za, (int)t1, (int)f1, dt, df, preserve, d.zb, s));
    st.check(fftfree::ft_complex_to_polar(d.zb, n0, d.ma, d.pa, s));
  }
  if (!st.ok) return 0;
  st.check(hipMemcpy(a, d.ma, n0 * sizeof(float), hipMemcpyDeviceToHost));
  st.check(hipMemcpy(b, d.pa, n0 * sizeof(float), hipMemcpyDeviceToHost));
  return st.ok ? 1 : 0;
}

}  // namespace

extern "C" {

int ft_apply_complex_pipeline(const ft_grid_spec* spec, const ft_resample_plan_c* plan, const ft_quantizer_config* quant,
                              const ft_contrast_config* contrast, float* real, float* imag, int time_bins,
                              int freq_bins) {
  try {
    return pipeline(false, spec, plan, quant, contrast, real, imag, time_bins, freq_bins);
  } catch (...) {
    return 0;
  }
}

int ft_apply_polar_pipeline(const ft_grid_spec* spec, const ft_resample_plan_c* plan, const ft_quantizer_config* quant,
                            const ft_contrast_config* contrast, float* magnitude, float* phase, int time_bins,
                            int freq_bins) {
  try {
    return pipeline(true, spec, plan, quant, contrast, magnitude, phase, time_bins, freq_bins);
  } catch (...) {
    return 0;
  }
}

}  // extern "C"

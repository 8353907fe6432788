/* ft_tools_cffi.h -- C ABI of the time-frequency grid pipeline in libfft_cffi.so.
 *
 * reference: ft_tools_cffi.hpp:9-58 (ft_tools_cffi.cpp:76-170, ft_tools.hpp:144-477), built
 * into the same library as fft_cffi (CMakeLists.txt:260). The grids are time-major
 * (index t * freq_bins + f), real/imag or magnitude/phase planes, changed in place. The
 * pipeline: bilinear upsampling, the exponential harmonic magnitude quantizer, the local
 * contrast booster, box downsampling (averaging when preserve_energy); the result must
 * have the input's shape, else the call returns 0 and leaves the planes as they were.
 * Returns 1 on success. Host pointers; the work runs on the GPU.
 */
#ifndef FFTFREE_AMD_FT_TOOLS_CFFI_H
#define FFTFREE_AMD_FT_TOOLS_CFFI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct ft_grid_spec {
  double time_origin;
  double freq_origin;
  double time_step; /* 0 reads as 1 */
  double freq_step; /* 0 reads as 1 */
};

struct ft_resample_plan_c {
  int upsample_time; /* <= 0 reads as 1 */
  int upsample_freq;
  int downsample_time;
  int downsample_freq;
  int preserve_energy;
};

struct ft_quantizer_config {
  int enabled;
  float base_frequency;
  float magnitude_step;
  float mix;
  float distance_decay;
};

struct ft_contrast_config {
  int enabled;
  int kernel_time; /* <= 0: the whole axis */
  int kernel_freq;
  float brightness;
  float contrast;
  float gamma;
  int clamp_zero;
};

int ft_apply_complex_pipeline(const struct ft_grid_spec* spec, const struct ft_resample_plan_c* plan,
                              const struct ft_quantizer_config* quant, const struct ft_contrast_config* contrast,
                              float* real, float* imag, int time_bins, int freq_bins);
int ft_apply_polar_pipeline(const struct ft_grid_spec* spec, const struct ft_resample_plan_c* plan,
                            const struct ft_quantizer_config* quant, const struct ft_contrast_config* contrast,
                            float* magnitude, float* phase, int time_bins, int freq_bins);

/* Additive (no counterpart in the reference): the same pipelines on device planes,
 * enqueued on `stream` (a hipStream_t; NULL = the default stream). The output planes may
 * be the input planes (in place) or separate buffers; other overlaps are not allowed.
 * Without resampling and with a local contrast box of at most 17 x 33 cells (or none),
 * the pipeline is one launch; otherwise it runs step by step on a workspace taken from
 * the stream-ordered allocator. Returns 1 once the work is enqueued, 0 on the argument
 * and shape errors of the host forms. */
int ft_apply_complex_pipeline_device(const struct ft_grid_spec* spec, const struct ft_resample_plan_c* plan,
                                     const struct ft_quantizer_config* quant,
                                     const struct ft_contrast_config* contrast, const float* d_real,
                                     const float* d_imag, float* d_out_real, float* d_out_imag, int time_bins,
                                     int freq_bins, void* stream);
int ft_apply_polar_pipeline_device(const struct ft_grid_spec* spec, const struct ft_resample_plan_c* plan,
                                   const struct ft_quantizer_config* quant, const struct ft_contrast_config* contrast,
                                   const float* d_magnitude, const float* d_phase, float* d_out_magnitude,
                                   float* d_out_phase, int time_bins, int freq_bins, void* stream);

#ifdef __cplusplus
}
#endif

#endif

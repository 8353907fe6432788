// ft_kernels.hpp -- launchers of the time-frequency grid pipeline (ft_tools.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace fftfree {

// ExponentialHarmonicQuantizer settings (ft_tools.hpp:262-268).
struct FtQuantizer {
  float base_frequency = 55.0f;
  float magnitude_step = 0.05f;
  float mix = 0.5f;
  float distance_decay = 8.0f;
};

// SpectralLocalContrastBooster settings (ft_tools.hpp:358-366).
struct FtContrast {
  int kernel_time = 3;
  int kernel_freq = 3;
  float brightness = 0.0f;
  float contrast = 0.5f;
  float gamma = 1.0f;
  int clamp_zero = 1;
};

// Grids are time-major complex (float2) or polar (two float planes).
hipError_t ft_upsample(const float2* src, int ts, int fs, int tf, int ff, float2* dst, hipStream_t s);
hipError_t ft_downsample(const float2* src, int ts, int fs, int tf, int ff, int preserve, float2* dst,
                         hipStream_t s);
hipError_t ft_polar_to_complex(const float* mag, const float* ph, size_t total, float2* z, hipStream_t s);
hipError_t ft_complex_to_polar(const float2* z, size_t total, float* mag, float* ph, hipStream_t s);
// z (complex grid) or mag (polar grid; z == nullptr)
hipError_t ft_quantize(float2* z, float* mag, int tb, int fb, const FtQuantizer& q, double freq_origin,
                       double freq_step, hipStream_t s);
// scratch: 2 * tb * fb + 1 floats
hipError_t ft_contrast(float2* z, float* mag, int tb, int fb, const FtContrast& c, float* scratch, hipStream_t s);

}  // namespace fftfree

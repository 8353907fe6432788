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
  // 1 / base_frequency and 1 / magnitude_step, rounded once on the host: the reference's
  // -ffast-math build hoists these reciprocals out of its loops and multiplies by them in
  // place of freq / base and magnitude / step.
  float inv_base = 1.0f / 55.0f;
  float inv_step = 1.0f / 0.05f;
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

// Tile of the one-pass kernel and the largest box-stencil radii it holds in LDS.
constexpr int kFtTileT = 16, kFtTileF = 128;
constexpr int kFtMaxRadiusT = 8, kFtMaxRadiusF = 16;
constexpr float kFtEpsilon = 1.1920928955078125e-07f;  // FLT_EPSILON: |z| at or below it is left alone

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

// Quantizer + local contrast (no resampling) in one pass from planes to planes. The output
// planes must not overlap the input planes when the contrast unit is on.
bool ft_fused_supported(int use_c, const FtContrast& c);
hipError_t ft_fused(bool polar, const float* in_a, const float* in_b, float* out_a, float* out_b, int tb, int fb,
                    const FtQuantizer& q, int use_q, double freq_origin, double freq_step, const FtContrast& c,
                    int use_c, hipStream_t s);
hipError_t ft_planes_to_complex(const float* a, const float* b, size_t total, float2* z, hipStream_t s);
hipError_t ft_complex_to_planes(const float2* z, size_t total, float* a, float* b, hipStream_t s);

}  // namespace fftfree

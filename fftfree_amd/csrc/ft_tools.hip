// ft_tools.hip -- the time-frequency grid pipeline of ft_apply_complex_pipeline /
// ft_apply_polar_pipeline (ft_tools_cffi.cpp:76-170 over ft_tools.hpp:144-477) on the
// device: bilinear upsampling, the exponential harmonic magnitude quantizer, the local
// contrast booster (a box stencil over magnitudes) and box downsampling. Grids are
// time-major (index t * freq_bins + f). Each step is one elementwise or stencil launch;
// every output cell is computed by one thread, in the reference's order of float
// operations within that cell.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ft_kernels.hpp"

namespace fftfree {

namespace {

__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + (b - a) * t; }

// |z| as the reference's build computes std::abs of a complex float: its -ffast-math
// flags turn it into sqrt(re * re + im * im) with the sum fused (mul, fma, sqrt), and a
// correctly rounded sqrtf of the same fma gives the same bits.
__device__ __forceinline__ float ft_abs(float x, float y) { return sqrtf(__builtin_fmaf(x, x, y * y)); }

// ft_tools.hpp:144-183: bilinear interpolation at (t / tf, f / ff).
__global__ void __launch_bounds__(256) ft_upsample_kernel(const float2* __restrict__ src, int ts, int fs, int tf,
                                                          int ff, float2* __restrict__ dst) {
  const int td = ts * tf, fd = fs * ff;
  const size_t total = (size_t)td * fd;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
    const int t = (int)(idx / (size_t)fd);
    const int f = (int)(idx - (size_t)t * fd);
    const double st = (double)t / (double)tf;
    const int t0 = (int)floor(st);
    const int t1 = min(t0 + 1, ts - 1);
    const float tw = (float)(st - (double)t0);
    const double sf = (double)f / (double)ff;
    const int f0 = (int)floor(sf);
    const int f1 = min(f0 + 1, fs - 1);
    const float fw = (float)(sf - (double)f0);
    const float2 c00 = src[(size_t)t0 * fs + f0], c10 = src[(size_t)t1 * fs + f0];
    const float2 c01 = src[(size_t)t0 * fs + f1], c11 = src[(size_t)t1 * fs + f1];
    const float2 a = make_float2(lerpf(c00.x, c10.x, tw), lerpf(c00.y, c10.y, tw));
    const float2 b = make_float2(lerpf(c01.x, c11.x, tw), lerpf(c01.y, c11.y, tw));
    dst[idx] = make_float2(lerpf(a.x, b.x, fw), lerpf(a.y, b.y, fw));
  }
}

// ft_tools.hpp:185-227: box average over factor x factor source cells (sum when
// preserve_energy is off), source rows outer, columns inner.
__global__ void __launch_bounds__(256) ft_downsample_kernel(const float2* __restrict__ src, int ts, int fs, int tf,
                                                            int ff, int preserve, float2* __restrict__ dst) {
  const int td = (ts + tf - 1) / tf, fd = (fs + ff - 1) / ff;
  const size_t total = (size_t)td * fd;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
    const int t = (int)(idx / (size_t)fd);
    const int f = (int)(idx - (size_t)t * fd);
    const int tb = t * tf, te = min(tb + tf, ts);
    const int fb = f * ff, fe = min(fb + ff, fs);
    float2 acc = make_float2(0.0f, 0.0f);
    int count = 0;
    for (int a = tb; a < te; ++a)
      for (int b = fb; b < fe; ++b) {
        const float2 v = src[(size_t)a * fs + b];
        acc.x += v.x;
        acc.y += v.y;
        ++count;
      }
    if (preserve && count > 0) {
      const float inv = 1.0f / (float)count;
      acc.x *= inv;
      acc.y *= inv;
    }
    dst[idx] = acc;
  }
}

// std::polar / std::abs / std::arg between the two grid forms (ft_tools.hpp:123-142).
__global__ void __launch_bounds__(256) ft_polar_to_complex_kernel(const float* __restrict__ mag,
                                                                  const float* __restrict__ ph, size_t total,
                                                                  float2* __restrict__ z) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    float s, c;
    sincosf(ph[i], &s, &c);
    z[i] = make_float2(mag[i] * c, mag[i] * s);
  }
}

__global__ void __launch_bounds__(256) ft_complex_to_polar_kernel(const float2* __restrict__ z, size_t total,
                                                                  float* __restrict__ mag, float* __restrict__ ph) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const float2 v = z[i];
    mag[i] = ft_abs(v.x, v.y);
    ph[i] = atan2f(v.y, v.x);
  }
}

// ExponentialHarmonicQuantizer (ft_tools.hpp:267-356).
__device__ float quant_weight(const FtQuantizer& q, float freq) {
  if (q.mix <= 0.0f) return 0.0f;
  if (q.base_frequency <= 0.0f) return fminf(fmaxf(q.mix, 0.0f), 1.0f);
  const float ratio = freq * q.inv_base;
  if (!isfinite((double)ratio)) return 0.0f;
  const float nearest = roundf(ratio);  // halves away from zero: std::round on the double, exactly
  const float distance = fabsf(ratio - nearest);
  const float attraction = expf(-q.distance_decay * distance);
  return fminf(fmaxf(q.mix * attraction, 0.0f), 1.0f);
}

__device__ __forceinline__ float quant_apply(const FtQuantizer& q, float m, float w) {
  if (w <= 0.0f) return m;
  float target = m;
  if (q.magnitude_step > 0.0f) target = roundf(m * q.inv_step) * q.magnitude_step;
  return lerpf(m, target, w);
}

__device__ float quant_magnitude(const FtQuantizer& q, float m, float freq) {
  return quant_apply(q, m, quant_weight(q, freq));
}

__global__ void __launch_bounds__(256) ft_quantize_kernel(float2* __restrict__ z, float* __restrict__ mag, int tb,
                                                          int fb, FtQuantizer q, double freq_origin,
                                                          double freq_step) {
  const size_t total = (size_t)tb * fb;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const int f = (int)(i % (size_t)fb);
    const float freq = (float)(freq_origin + freq_step * (double)f);
    if (mag) {  // polar grid: the magnitude alone
      mag[i] = quant_magnitude(q, mag[i], freq);
    } else {    // complex grid: |z| moved, the phase kept; tiny values untouched
      const float2 v = z[i];
      const float m = ft_abs(v.x, v.y);
      if (m <= kFtEpsilon) continue;
      const float adj = fmaxf(quant_magnitude(q, m, freq), 0.0f);
      float s, c;
      sincosf(atan2f(v.y, v.x), &s, &c);
      z[i] = make_float2(adj * c, adj * s);
    }
  }
}

__global__ void __launch_bounds__(256) ft_abs_kernel(const float2* __restrict__ z, size_t total,
                                                     float* __restrict__ mag) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256)
    mag[i] = ft_abs(z[i].x, z[i].y);
}

// Mean of all magnitudes (the global contrast mode): one workgroup.
__global__ void __launch_bounds__(256) ft_mean_kernel(const float* __restrict__ mag, size_t total,
                                                      float* __restrict__ out) {
  __shared__ float part[256];
  float s = 0.0f;
  for (size_t i = threadIdx.x; i < total; i += 256) s += mag[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0] / (float)total;
}

// SpectralLocalContrastBooster (ft_tools.hpp:358-461): local box mean, contrast about it,
// brightness, optional clamp at 0 and gamma.
__global__ void __launch_bounds__(256) ft_contrast_kernel(const float* __restrict__ mag, int tb, int fb,
                                                          FtContrast c, const float* __restrict__ global_mean,
                                                          float* __restrict__ out) {
  const bool gt = c.kernel_time <= 0, gf = c.kernel_freq <= 0;
  const int rt = gt ? tb : max(c.kernel_time / 2, 0);
  const int rf = gf ? fb : max(c.kernel_freq / 2, 0);
  const float gamma = c.gamma > 0.0f ? c.gamma : 1.0f;
  const bool use_gamma = fabsf(gamma - 1.0f) > 1e-5f;
  const float inv_gamma = use_gamma ? 1.0f / gamma : 1.0f;
  const size_t total = (size_t)tb * fb;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const int t = (int)(i / (size_t)fb);
    const int f = (int)(i - (size_t)t * fb);
    float mean;
    if (gt && gf) {
      mean = global_mean[0];
    } else {
      const int t0 = gt ? 0 : max(0, t - rt), t1 = gt ? tb - 1 : min(tb - 1, t + rt);
      const int f0 = gf ? 0 : max(0, f - rf), f1 = gf ? fb - 1 : min(fb - 1, f + rf);
      float sum = 0.0f;
      int count = 0;
      for (int a = t0; a <= t1; ++a)
        for (int b = f0; b <= f1; ++b) {
          sum += mag[(size_t)a * fb + b];
          ++count;
        }
      mean = count > 0 ? sum / (float)count : 0.0f;
    }
    float adj = mean + (mag[i] - mean) * (1.0f + c.contrast) + c.brightness;
    if (c.clamp_zero) adj = fmaxf(adj, 0.0f);
    if (use_gamma) adj = (float)pow((double)fmaxf(adj, 1e-12f), (double)inv_gamma);
    out[i] = adj;
  }
}

// z = std::polar(mag, arg z) (ft_tools.hpp:381-385).
__global__ void __launch_bounds__(256) ft_set_magnitude_kernel(float2* __restrict__ z, const float* __restrict__ mag,
                                                               size_t total) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    float s, c;
    sincosf(atan2f(z[i].y, z[i].x), &s, &c);
    z[i] = make_float2(mag[i] * c, mag[i] * s);
  }
}

// Contrast adjustment of one magnitude about its box mean (ft_tools.hpp:436-452).
__device__ __forceinline__ float contrast_adjust(const FtContrast& c, float m, float mean, bool use_gamma,
                                                 float inv_gamma) {
  float adj = mean + (m - mean) * (1.0f + c.contrast) + c.brightness;
  if (c.clamp_zero) adj = fmaxf(adj, 0.0f);
  if (use_gamma) adj = (float)pow((double)fmaxf(adj, 1e-12f), (double)inv_gamma);
  return adj;
}

__device__ __forceinline__ int xcd_order(int i, int G) {  // contiguous tile runs per XCD
  const int q = G >> 3, rem = G & 7;
  const int x = i & 7, s = i >> 3;
  return x * q + (x < rem ? x : rem) + s;
}

// The pipeline without resampling in one pass: quantizer, then the local contrast
// booster, read from and written to planes. A workgroup owns a tile of kFtTileT x
// kFtTileF cells. Each thread first issues every load it needs (its share of the tile
// plus the box-stencil halo, and the raw values of its own output cells), then the
// quantized magnitudes go to LDS once, and each output cell takes its box mean there.
// Complex grids: the result is polar(m', arg z_q) with z_q = polar(adj, arg z), as the
// two units compute it one after the other; arg z_q = arg z except where adj is 0, when
// z_q is a signed zero. Tiles are dealt to XCDs in runs along time, so vertical halos
// hit L2.
template <bool POLAR, int KPER>
__global__ void __launch_bounds__(256) ft_fused_kernel(const float* in_a, const float* in_b, float* out_a, float* out_b,
                                                       int tb, int fb, FtQuantizer q, int use_q, double freq_origin,
                                                       double freq_step, FtContrast c, int use_c, int tiles_t,
                                                       int tiles_f) {
  constexpr int kExt = (kFtTileT + 2 * kFtMaxRadiusT) * (kFtTileF + 2 * kFtMaxRadiusF);
  constexpr int kPer = KPER;                         // halo-tile cells per thread, at most
  constexpr int kOut = kFtTileT * kFtTileF / 256;    // output cells per thread
  constexpr int kRowStep = 256 / kFtTileF;
  static_assert(kExt % 256 == 0 && (kFtTileT * kFtTileF) % 256 == 0, "tile shape");
  static_assert(KPER * 256 <= kExt, "halo tile");
  __shared__ float mq[KPER * 256];                  // quantized magnitudes of the halo tile
  __shared__ float mabs[POLAR ? 1 : KPER * 256];    // |z| of the halo tile (complex grids)
  __shared__ float wcol[kFtTileF + 2 * kFtMaxRadiusF];
  const int tid = threadIdx.x;
  const int tile = xcd_order(blockIdx.x, tiles_t * tiles_f);
  const int tfi = tile / tiles_t, tti = tile - tfi * tiles_t;
  const int t0 = tti * kFtTileT, f0 = tfi * kFtTileF;
  const int rt = use_c ? max(c.kernel_time / 2, 0) : 0;
  const int rf = use_c ? max(c.kernel_freq / 2, 0) : 0;
  const int er = kFtTileT + 2 * rt, ec = kFtTileF + 2 * rf;
  const int n_ext = er * ec;
  const float inv_ec = 1.0f / (float)ec;  // (i + 0.5) * inv_ec truncates to i / ec for i < kExt
  // 1. loads, all issued before any is used: the halo tile and this thread's output
  //    cells. Halo cells outside the grid read element 0; their LDS values are never
  //    read, since the box is clipped to the grid.
  float xa[kPer], xb[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + 256 * k;
    const int r = (int)(((float)i + 0.5f) * inv_ec), j = i - r * ec;
    const int t = t0 - rt + r, f = f0 - rf + j;
    const bool valid = i < n_ext && t >= 0 && t < tb && f >= 0 && f < fb;
    const size_t at = valid ? (size_t)t * fb + f : 0;
    xa[k] = in_a[at];
    xb[k] = POLAR ? 0.0f : in_b[at];
  }
  const int jo = tid & (kFtTileF - 1), ro = tid / kFtTileF;
  const int fo = f0 + jo;
  const bool copy_phase = POLAR && out_b != in_b;
  float oa[kOut], ob[kOut];
#pragma unroll
  for (int m = 0; m < kOut; ++m) {
    const int t = t0 + ro + kRowStep * m;
    const bool valid = t < tb && fo < fb;
    const size_t at = valid ? (size_t)t * fb + fo : 0;
    oa[m] = POLAR ? 0.0f : in_a[at];
    ob[m] = (!POLAR || copy_phase) ? in_b[at] : 0.0f;
  }
  // 2. quantizer weights per column, then quantized magnitudes into LDS
  for (int j = tid; j < ec; j += 256) {
    const int f = f0 - rf + j;
    wcol[j] = (use_q && f >= 0 && f < fb) ? quant_weight(q, (float)(freq_origin + freq_step * (double)f)) : 0.0f;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + 256 * k;
    if (i < n_ext) {
      const int r = (int)(((float)i + 0.5f) * inv_ec), j = i - r * ec;
      const int t = t0 - rt + r, f = f0 - rf + j;
      const bool in_grid = t >= 0 && t < tb && f >= 0 && f < fb;
      const float w = wcol[j];
      float v;
      if (POLAR) {
        v = quant_apply(q, xa[k], w);
      } else {
        const float m = ft_abs(xa[k], xb[k]);
        v = (w > 0.0f && m > kFtEpsilon) ? fmaxf(quant_apply(q, m, w), 0.0f) : m;
        mabs[i] = m;
      }
      mq[i] = in_grid ? v : 0.0f;  // zeros outside the grid: the 3 x 3 sum below adds them
    }
  }
  __syncthreads();
  // 3. box mean, contrast, direction, stores
  if (fo >= fb) return;
  const float gamma = c.gamma > 0.0f ? c.gamma : 1.0f;
  const bool use_gamma = fabsf(gamma - 1.0f) > 1e-5f;
  const float inv_gamma = use_gamma ? 1.0f / gamma : 1.0f;
  const int f_lo = max(0, fo - rf) - f0 + rf, f_hi = min(fb - 1, fo + rf) - f0 + rf;
  const float wq = wcol[jo + rf];
#pragma unroll
  for (int m = 0; m < kOut; ++m) {
    const int r = ro + kRowStep * m;
    const int t = t0 + r;
    if (t >= tb) break;
    const float mq0 = mq[(r + rt) * ec + jo + rf];
    float mc = mq0;
    if (use_c) {
      const int t_lo = max(0, t - rt) - t0 + rt, t_hi = min(tb - 1, t + rt) - t0 + rt;
      float sum = 0.0f;
      if (rt == 1 && rf == 1) {
        // the 3 x 3 box unrolled, row-major as the reference sums it; cells outside the
        // grid hold 0 and adding 0 leaves every partial sum as the clipped loop has it
        const float* q0 = mq + r * ec + jo;
        sum = q0[0] + q0[1] + q0[2];
        sum = sum + q0[ec] + q0[ec + 1] + q0[ec + 2];
        sum = sum + q0[2 * ec] + q0[2 * ec + 1] + q0[2 * ec + 2];
      } else {
        for (int a = t_lo; a <= t_hi; ++a)
          for (int b = f_lo; b <= f_hi; ++b) sum += mq[a * ec + b];
      }
      // an IEEE divide, as the reference's build divides (local_mean /= count,
      // ft_tools.hpp:433-437): the divisor changes per cell, so -ffast-math keeps it a divide
      const float mean = sum / (float)((t_hi - t_lo + 1) * (f_hi - f_lo + 1));
      mc = contrast_adjust(c, mq0, mean, use_gamma, inv_gamma);
    }
    const size_t at = (size_t)t * fb + fo;
    if (POLAR) {
      out_a[at] = mc;
      if (copy_phase) out_b[at] = ob[m];
      continue;
    }
    const float x = oa[m], y = ob[m];
    const float m0 = mabs[(r + rt) * ec + jo + rf];
    const bool quantized = wq > 0.0f && m0 > kFtEpsilon;
    if (!quantized && !use_c) {
      out_a[at] = x;
      out_b[at] = y;
      continue;
    }
    float dx, dy;
    const float inv = __builtin_amdgcn_rcpf(m0);
    if (m0 > 0.0f && !(quantized && use_c && mq0 == 0.0f)) {
      if (inv <= 3.4028234663852886e38f) {  // cos(atan2(y, x)), sin(atan2(y, x))
        dx = x * inv;
        dy = y * inv;
      } else {                               // |z| below 1 / FLT_MAX
        dx = x / m0;
        dy = y / m0;
      }
    } else {
      // a zero: the argument follows atan2's signed-zero cases
      float sn, cs;
      sincosf(atan2f(y, x), &sn, &cs);
      if (m0 > 0.0f) sincosf(atan2f(mq0 * sn, mq0 * cs), &sn, &cs);
      dx = cs;
      dy = sn;
    }
    out_a[at] = mc * dx;
    out_b[at] = mc * dy;
  }
}

__global__ void __launch_bounds__(256) ft_planes_to_complex_kernel(const float* __restrict__ a,
                                                                   const float* __restrict__ b, size_t total,
                                                                   float2* __restrict__ z) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256)
    z[i] = make_float2(a[i], b[i]);
}

__global__ void __launch_bounds__(256) ft_complex_to_planes_kernel(const float2* __restrict__ z, size_t total,
                                                                   float* __restrict__ a, float* __restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const float2 v = z[i];
    a[i] = v.x;
    b[i] = v.y;
  }
}

unsigned grid_of(size_t work) {
  const size_t g = (work + 255) / 256;
  return (unsigned)(g == 0 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace

hipError_t ft_upsample(const float2* src, int ts, int fs, int tf, int ff, float2* dst, hipStream_t s) {
  hipLaunchKernelGGL(ft_upsample_kernel, dim3(grid_of((size_t)ts * tf * fs * ff)), dim3(256), 0, s, src, ts, fs, tf,
                     ff, dst);
  return hipGetLastError();
}

hipError_t ft_downsample(const float2* src, int ts, int fs, int tf, int ff, int preserve, float2* dst,
                         hipStream_t s) {
  const size_t out = (size_t)((ts + tf - 1) / tf) * ((fs + ff - 1) / ff);
  hipLaunchKernelGGL(ft_downsample_kernel, dim3(grid_of(out)), dim3(256), 0, s, src, ts, fs, tf, ff, preserve, dst);
  return hipGetLastError();
}

hipError_t ft_polar_to_complex(const float* mag, const float* ph, size_t total, float2* z, hipStream_t s) {
  hipLaunchKernelGGL(ft_polar_to_complex_kernel, dim3(grid_of(total)), dim3(256), 0, s, mag, ph, total, z);
  return hipGetLastError();
}

hipError_t ft_complex_to_polar(const float2* z, size_t total, float* mag, float* ph, hipStream_t s) {
  hipLaunchKernelGGL(ft_complex_to_polar_kernel, dim3(grid_of(total)), dim3(256), 0, s, z, total, mag, ph);
  return hipGetLastError();
}

hipError_t ft_quantize(float2* z, float* mag, int tb, int fb, const FtQuantizer& q, double freq_origin,
                       double freq_step, hipStream_t s) {
  hipLaunchKernelGGL(ft_quantize_kernel, dim3(grid_of((size_t)tb * fb)), dim3(256), 0, s, z, mag, tb, fb, q,
                     freq_origin, freq_step);
  return hipGetLastError();
}

hipError_t ft_contrast(float2* z, float* mag, int tb, int fb, const FtContrast& c, float* scratch, hipStream_t s) {
  const size_t total = (size_t)tb * fb;
  // scratch: total floats of magnitudes (complex grids), total floats of output, 1 mean
  float* m = z ? scratch : mag;
  float* out = scratch + total;
  float* mean = scratch + 2 * total;
  if (z) hipLaunchKernelGGL(ft_abs_kernel, dim3(grid_of(total)), dim3(256), 0, s, z, total, m);
  if (c.kernel_time <= 0 && c.kernel_freq <= 0)
    hipLaunchKernelGGL(ft_mean_kernel, dim3(1), dim3(256), 0, s, m, total, mean);
  hipLaunchKernelGGL(ft_contrast_kernel, dim3(grid_of(total)), dim3(256), 0, s, m, tb, fb, c, mean, out);
  if (z) {
    hipLaunchKernelGGL(ft_set_magnitude_kernel, dim3(grid_of(total)), dim3(256), 0, s, z, out, total);
  } else {
    const hipError_t e = hipMemcpyAsync(mag, out, total * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

bool ft_fused_supported(int use_c, const FtContrast& c) {
  if (!use_c) return true;
  return c.kernel_time > 0 && c.kernel_freq > 0 && c.kernel_time / 2 <= kFtMaxRadiusT &&
         c.kernel_freq / 2 <= kFtMaxRadiusF;
}

hipError_t ft_fused(bool polar, const float* in_a, const float* in_b, float* out_a, float* out_b, int tb, int fb,
                    const FtQuantizer& q, int use_q, double freq_origin, double freq_step, const FtContrast& c,
                    int use_c, hipStream_t s) {
  const int tiles_t = (tb + kFtTileT - 1) / kFtTileT, tiles_f = (fb + kFtTileF - 1) / kFtTileF;
  if ((long long)tiles_t * tiles_f > 0x7fffffffll) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(tiles_t * tiles_f));
  const int rt = use_c ? max(c.kernel_time / 2, 0) : 0, rf = use_c ? max(c.kernel_freq / 2, 0) : 0;
  const int n_ext = (kFtTileT + 2 * rt) * (kFtTileF + 2 * rf);
  constexpr int kSmall = 10, kLarge = (kFtTileT + 2 * kFtMaxRadiusT) * (kFtTileF + 2 * kFtMaxRadiusF) / 256;
  const bool small = n_ext <= 256 * kSmall;  // boxes up to 5 x 5
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, in_a, in_b, out_a, out_b, tb, fb, q, use_q, freq_origin,
                       freq_step, c, use_c, tiles_t, tiles_f);
  };
  if (polar)
    small ? go(ft_fused_kernel<true, kSmall>) : go(ft_fused_kernel<true, kLarge>);
  else
    small ? go(ft_fused_kernel<false, kSmall>) : go(ft_fused_kernel<false, kLarge>);
  return hipGetLastError();
}

hipError_t ft_planes_to_complex(const float* a, const float* b, size_t total, float2* z, hipStream_t s) {
  hipLaunchKernelGGL(ft_planes_to_complex_kernel, dim3(grid_of(total)), dim3(256), 0, s, a, b, total, z);
  return hipGetLastError();
}

hipError_t ft_complex_to_planes(const float2* z, size_t total, float* a, float* b, hipStream_t s) {
  hipLaunchKernelGGL(ft_complex_to_planes_kernel, dim3(grid_of(total)), dim3(256), 0, s, z, total, a, b);
  return hipGetLastError();
}

}  // namespace fftfree

// ref_ft_driver.cpp -- TEST INFRASTRUCTURE ONLY. A driver compiled together with the
// reference's own header-only ft_tools.hpp (included from where it lies, never copied),
// so that the oracle's restatement of the grid pipeline can be checked against the
// reference's code. ft_tools_cffi.cpp itself cannot be compiled here: it includes
// fft_cffi.hpp, which includes eigen_fft.hpp and so needs Eigen. This driver therefore
// takes the settings already mapped as ft_tools_cffi.cpp:18-58 maps them (steps of 0
// read as 1, factors <= 0 read as 1) and runs the reference's templates in the order of
// ft_tools_cffi.cpp:88-119: upsample_field, the quantizer, the contrast booster,
// downsample_field, then the shape check.
#include <cstddef>

#include "ft_tools.hpp"

namespace {

template <typename Field>
int run(Field& field, int ut, int uf, int dt, int df, int preserve, int q_on, float qb, float qs, float qm, float qd,
        int c_on, int ckt, int ckf, float cb, float cc, float cg, int cz) {
  if (ut > 1 || uf > 1) field = fftfree::ft::upsample_field(field, ut, uf);
  fftfree::ft::ExponentialHarmonicQuantizer<float> quantizer;
  quantizer.base_frequency = qb;
  quantizer.magnitude_step = qs;
  quantizer.mix = qm;
  quantizer.distance_decay = qd;
  fftfree::ft::SpectralLocalContrastBooster<float> contrast;
  contrast.kernel_time = ckt;
  contrast.kernel_freq = ckf;
  contrast.brightness = cb;
  contrast.contrast = cc;
  contrast.gamma = cg;
  contrast.clamp_zero = cz != 0;
  if (q_on) quantizer(field);
  if (c_on) contrast(field);
  if (dt > 1 || df > 1) field = fftfree::ft::downsample_field(field, dt, df, preserve != 0);
  return 1;
}

}  // namespace

extern "C" int refsrc_ft_pipeline(int polar, double f_origin, double f_step, int ut, int uf, int dt, int df,
                                  int preserve, int q_on, float qb, float qs, float qm, float qd, int c_on, int ckt,
                                  int ckf, float cb, float cc, float cg, int cz, float* a, float* b, int tb, int fb) {
  fftfree::ft::GridSpec spec;
  spec.freq_origin = f_origin;
  spec.freq_step = f_step != 0.0 ? f_step : 1.0;
  ut = ut > 0 ? ut : 1;
  uf = uf > 0 ? uf : 1;
  dt = dt > 0 ? dt : 1;
  df = df > 0 ? df : 1;
  const std::size_t total = static_cast<std::size_t>(tb) * static_cast<std::size_t>(fb);
  if (polar) {
    fftfree::ft::PolarGrid<float> field(tb, fb, spec);
    for (std::size_t i = 0; i < total; ++i) {
      field.magnitude[i] = a[i];
      field.phase[i] = b[i];
    }
    run(field, ut, uf, dt, df, preserve, q_on, qb, qs, qm, qd, c_on, ckt, ckf, cb, cc, cg, cz);
    if (field.time_bins != tb || field.freq_bins != fb) return 0;
    for (std::size_t i = 0; i < total; ++i) {
      a[i] = field.magnitude[i];
      b[i] = field.phase[i];
    }
    return 1;
  }
  fftfree::ft::ComplexGrid<float> field(tb, fb, spec);
  for (std::size_t i = 0; i < total; ++i) field.values[i] = std::complex<float>(a[i], b[i]);
  run(field, ut, uf, dt, df, preserve, q_on, qb, qs, qm, qd, c_on, ckt, ckf, cb, cc, cg, cz);
  if (field.time_bins != tb || field.freq_bins != fb) return 0;
  for (std::size_t i = 0; i < total; ++i) {
    a[i] = field.values[i].real();
    b[i] = field.values[i].imag();
  }
  return 1;
}

// fft_stft_mf.hip -- instantiations of the multi-frame STFT kernels
// (stft_r2c_mf_kernel, fft_stft_impl.hpp) and their rows of the variant table.
#include "fft_stft_impl.hpp"

namespace fftfree {
namespace {
const StftEntry kMf[] = {
    FFTFREE_STFT_MF(0, 256, 16, 4, "16x16 multiframe k=4", 16, 16),
    FFTFREE_STFT_PL(3, 512, 8, 6, "8x8x8 tile-preload k=6", 8, 8, 8),
    FFTFREE_STFT_MF(1, 1024, 16, 8, "16x16x4 multiframe k=8", 16, 16, 4),
};
}  // namespace

const StftEntry* stft_mf_entries(int* count) {
  *count = (int)(sizeof(kMf) / sizeof(kMf[0]));
  return kMf;
}

}  // namespace fftfree

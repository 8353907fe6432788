// fft_stft_lines_2048.hip -- the line-aligned STFT at N = 2048 (one wave per frame), the
// default at that size.
#include "fft_stft_impl.hpp"

namespace fftfree {
namespace {
const StftEntry kRows[] = {
    // N = 2048 (the default there): one wave per frame (16 values per lane), twiddles in
    // registers, window taps in LDS, 2 workgroups per CU (LDS: frame buffers and the
    // ring); 122.5 us polar against 188.6 us for the multi-frame kernel at 32765 frames.
    // Window taps in registers too: 130.2 us (256 VGPRs and a small spill).
    FFTFREE_STFT_LN(0, 1024, 16, 8, 0, 487, "16x16x4 lines ldswin wg=2 k=8", 16, 16, 4),
    FFTFREE_STFT_LN(4, 1024, 16, 8, 0, 231, "16x16x4 lines wg=2 k=8", 16, 16, 4),
};
}  // namespace

const StftEntry* stft_lines_2048_entries(int* count) {
  *count = (int)(sizeof(kRows) / sizeof(kRows[0]));
  return kRows;
}

}  // namespace fftfree

// fft_stft_lines_4096.hip -- the line-aligned STFT at N = 4096 (two waves per frame), the
// default at that size.
#include "fft_stft_impl.hpp"

namespace fftfree {
namespace {
const StftEntry kRows[] = {
    // N = 4096 (the default there): two waves per frame (16 values per lane, workgroup
    // barriers between the passes), 2 frames per iteration, 16 iterations per 32-frame
    // tile; window taps in registers, twiddles in LDS (beside the frame buffers and the
    // ring): 130.8 us polar at 16381 frames against 182.2 us for the tile kernel; both
    // tables in registers: 132.4 us
    FFTFREE_STFT_LN(0, 2048, 16, 16, 0, 739, "16x16x8 lines regwin wg=2 k=16", 16, 16, 8),
    FFTFREE_STFT_LN(3, 2048, 16, 16, 0, 231, "16x16x8 lines wg=2 k=16", 16, 16, 8),
};
}  // namespace

const StftEntry* stft_lines_4096_entries(int* count) {
  *count = (int)(sizeof(kRows) / sizeof(kRows[0]));
  return kRows;
}

}  // namespace fftfree

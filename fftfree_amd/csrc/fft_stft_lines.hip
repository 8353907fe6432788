// fft_stft_lines.hip -- instantiations of the line-aligned N = 1024 STFT kernel
// (stft_r2c_lines_kernel, fft_stft_impl.hpp): the default, its workgroups-per-CU and
// small-batch rows and plain launch order. The other rows live in fft_stft_lines_exch.hip
// (register tables and exchanges), fft_stft_lines_order.hip (tile orders),
// fft_stft_lines_2048.hip and fft_stft_lines_4096.hip, one translation unit each so that
// they compile side by side.
#include "fft_stft_impl.hpp"

namespace fftfree {
namespace {
const StftEntry kRows[] = {
    // the default: tiles dealt 8 consecutive per XCD, even tiles last frames first (reads
    // 1.02x the samples instead of 1.10x, the same time)
    FFTFREE_STFT_LN(0, 512, 8, 8, 0, 227, "8x8x8 lines wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(1, 512, 8, 8, 3, 3, "8x8x8 lines wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(2, 512, 8, 8, 3, 0, "8x8x8 lines-xcd-rounds wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(6, 512, 8, 2, 0, 3, "8x8x8 small-batch lines k=2", 8, 8, 8),
    FFTFREE_STFT_LN(7, 512, 8, 1, 0, 3, "8x8x8 small-batch lines k=1", 8, 8, 8),
    FFTFREE_STFT_LN(18, 512, 8, 8, 0, 3, "8x8x8 lines launch-order wg=4 k=8", 8, 8, 8),
};
}  // namespace

const StftEntry* stft_lines_entries(int* count) {
  *count = (int)(sizeof(kRows) / sizeof(kRows[0]));
  return kRows;
}

}  // namespace fftfree

// fft_stft_lines_order.hip -- N = 1024 line-aligned STFT rows with other tile orders
// (groups of 2, 4, 8 tiles per XCD, with and without the alternating walk).
#include "fft_stft_impl.hpp"

namespace fftfree {
namespace {
const StftEntry kRows[] = {
    FFTFREE_STFT_LN(13, 512, 8, 8, 0, 35, "8x8x8 lines pairs wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(14, 512, 8, 8, 0, 67, "8x8x8 lines quads wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(15, 512, 8, 8, 0, 99, "8x8x8 lines octets wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(16, 512, 8, 8, 0, 163, "8x8x8 lines pairs-alt wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(17, 512, 8, 8, 0, 195, "8x8x8 lines quads-alt wg=4 k=8", 8, 8, 8),
};
}  // namespace

const StftEntry* stft_lines_order_entries(int* count) {
  *count = (int)(sizeof(kRows) / sizeof(kRows[0]));
  return kRows;
}

}  // namespace fftfree

// fft_stft_lines_exch.hip -- N = 1024 line-aligned STFT rows with register tables or
// register exchanges between the passes (measured against the default, kept as variants).
#include "fft_stft_impl.hpp"

namespace fftfree {
namespace {
const StftEntry kRows[] = {
    FFTFREE_STFT_LN(8, 512, 8, 8, 3, 7, "8x8x8 lines regs wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(9, 512, 8, 8, 3, 1, "8x8x8 lines mixed wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(10, 512, 8, 8, 0, 11, "8x8x8 lines dpp wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(11, 512, 8, 8, 3, 15, "8x8x8 lines dpp regs wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(12, 512, 8, 8, 0, 27, "8x8x8 lines dpp-hi wg=4 k=8", 8, 8, 8),
};
}  // namespace

const StftEntry* stft_lines_exch_entries(int* count) {
  *count = (int)(sizeof(kRows) / sizeof(kRows[0]));
  return kRows;
}

}  // namespace fftfree

// fft_stft_lines.hip -- instantiations of the line-aligned N = 1024 and 2048 STFT kernel
// (stft_r2c_lines_kernel, fft_stft_impl.hpp) and their rows of the variant table.
#include "fft_stft_impl.hpp"

namespace fftfree {
namespace {
const StftEntry kLines[] = {
    // the default: tiles dealt 8 consecutive per XCD, even tiles last frames first (reads
    // 1.02x the samples instead of 1.10x, the same time)
    FFTFREE_STFT_LN(0, 512, 8, 8, 0, 227, "8x8x8 lines wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(1, 512, 8, 8, 3, 3, "8x8x8 lines wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(2, 512, 8, 8, 3, 0, "8x8x8 lines-xcd-rounds wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(6, 512, 8, 2, 0, 3, "8x8x8 small-batch lines k=2", 8, 8, 8),
    FFTFREE_STFT_LN(7, 512, 8, 1, 0, 3, "8x8x8 small-batch lines k=1", 8, 8, 8),
    FFTFREE_STFT_LN(8, 512, 8, 8, 3, 7, "8x8x8 lines regs wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(9, 512, 8, 8, 3, 1, "8x8x8 lines mixed wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(10, 512, 8, 8, 0, 11, "8x8x8 lines dpp wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(11, 512, 8, 8, 3, 15, "8x8x8 lines dpp regs wg=3 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(12, 512, 8, 8, 0, 27, "8x8x8 lines dpp-hi wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(13, 512, 8, 8, 0, 35, "8x8x8 lines pairs wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(14, 512, 8, 8, 0, 67, "8x8x8 lines quads wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(15, 512, 8, 8, 0, 99, "8x8x8 lines octets wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(16, 512, 8, 8, 0, 163, "8x8x8 lines pairs-alt wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(17, 512, 8, 8, 0, 195, "8x8x8 lines quads-alt wg=4 k=8", 8, 8, 8),
    FFTFREE_STFT_LN(18, 512, 8, 8, 0, 3, "8x8x8 lines launch-order wg=4 k=8", 8, 8, 8),
    // N = 2048 (the default there): one wave per frame (16 values per lane), twiddles in
    // registers, window taps in LDS, 2 workgroups per CU (LDS: frame buffers and the
    // ring); 122.5 us polar against 188.6 us for the multi-frame kernel at 32765 frames.
    // Window taps in registers too: 130.2 us (256 VGPRs and a small spill).
    FFTFREE_STFT_LN(0, 1024, 16, 8, 0, 487, "16x16x4 lines ldswin wg=2 k=8", 16, 16, 4),
    FFTFREE_STFT_LN(4, 1024, 16, 8, 0, 231, "16x16x4 lines wg=2 k=8", 16, 16, 4),
    // N = 4096 (the default there): two waves per frame (16 values per lane, workgroup
    // barriers between the passes), 2 frames per iteration, 16 iterations per 32-frame
    // tile; window taps in registers, twiddles in LDS (beside the frame buffers and the
    // ring): 130.8 us polar at 16381 frames against 182.2 us for the tile kernel; both
    // tables in registers: 132.4 us
    FFTFREE_STFT_LN(0, 2048, 16, 16, 0, 739, "16x16x8 lines regwin wg=2 k=16", 16, 16, 8),
    FFTFREE_STFT_LN(3, 2048, 16, 16, 0, 231, "16x16x8 lines wg=2 k=16", 16, 16, 8),
};
}  // namespace

const StftEntry* stft_lines_entries(int* count) {
  *count = (int)(sizeof(kLines) / sizeof(kLines[0]));
  return kLines;
}

}  // namespace fftfree

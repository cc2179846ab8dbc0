// lnorm_dary32_hi.hip -- the 32-bit tabulated d-ary walk's kernels for
// d = 5..10 (lnorm_blk16.h; the launch entry is in lnorm_dary32.hip).
#include "lnorm_blk16.h"

namespace lnorm {

hipError_t launch_blk32_hi(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.d) {
    case 5: return launch_blk<5, 32>(mode, a, st, block);
    case 6: return launch_blk<6, 32>(mode, a, st, block);
    case 7: return launch_blk<7, 32>(mode, a, st, block);
    case 8: return launch_blk<8, 32>(mode, a, st, block);
    case 9: return launch_blk<9, 32>(mode, a, st, block);
    case 10: return launch_blk<10, 32>(mode, a, st, block);
    default: return hipErrorInvalidValue;
  }
}

int blocks_per_cu_blk32_hi(int d, int words, int block) {
  switch (d) {
    case 5: return blocks_per_cu_blk<5, 32>(words, block);
    case 6: return blocks_per_cu_blk<6, 32>(words, block);
    case 7: return blocks_per_cu_blk<7, 32>(words, block);
    case 8: return blocks_per_cu_blk<8, 32>(words, block);
    case 9: return blocks_per_cu_blk<9, 32>(words, block);
    case 10: return blocks_per_cu_blk<10, 32>(words, block);
    default: return 0;
  }
}

}  // namespace lnorm

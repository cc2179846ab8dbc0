// lnorm_dary.hip -- the tabulated packed d-ary walk, one lane per segment
// (d = 2..10, the M fastest rows as per-label subset tables, walk_blk16):
// widths, the launch entry, and the kernels for d = 2..4 (d = 5..10 are in
// lnorm_dary_hi.hip).
#include "lnorm_blk16.h"

namespace lnorm {

// the smallest instantiated width (blk16_has_width) holding the columns
int blk16_words(int d, int ny) {
  if (d < 2 || d > 10) return 0;
  const int need = (ny + 1) / 2;
  for (int w = 4; w <= 96; w += 2)
    if (w >= need && blk16_has_width(d, w)) return w;
  return 0;
}

int blk16_rows(int d, int words) { return blk16_low_rows(d, words); }

hipError_t launch_blk16_lo(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.d) {
    case 2: return launch_blk<2, 16>(mode, a, st, block);
    case 3: return launch_blk<3, 16>(mode, a, st, block);
    case 4: return launch_blk<4, 16>(mode, a, st, block);
    default: return hipErrorInvalidValue;
  }
}

int blocks_per_cu_blk16_lo(int d, int words, int block) {
  switch (d) {
    case 2: return blocks_per_cu_blk<2, 16>(words, block);
    case 3: return blocks_per_cu_blk<3, 16>(words, block);
    case 4: return blocks_per_cu_blk<4, 16>(words, block);
    default: return 0;
  }
}

int blk16_blocks_per_cu(int d, int words, int block) {
  return d <= 4 ? blocks_per_cu_blk16_lo(d, words, block) : blocks_per_cu_blk16_hi(d, words, block);
}

hipError_t launch_blk16(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (block <= 0 || block % 64 || block > 256) block = 256;
  const int words = a.ny_pad / 2;
  if (blk16_words(a.d, a.ny_pad) != words || words * 2 != a.ny_pad) return hipErrorInvalidValue;
  if (a.q < blk16_low_rows(a.d, words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  return a.d <= 4 ? launch_blk16_lo(mode, a, st, block) : launch_blk16_hi(mode, a, st, block);
}

}  // namespace lnorm

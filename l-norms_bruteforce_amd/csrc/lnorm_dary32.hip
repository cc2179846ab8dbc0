// lnorm_dary32.hip -- the tabulated d-ary walk in 32-bit words (walk_blk16
// with WB = 32): one column per word, biased by 2^31, v_sad_u32.  It serves
// int32 inputs whose column sums exceed the 16-bit walks' bound (32767), at
// the same table rows per word count as the 16-bit walk, so at half its
// columns.  Widths, the launch entry and the kernels for d = 2..4 (d = 5..10 in
// lnorm_dary32_hi.hip, so the two halves compile in parallel).
#include "lnorm_blk16.h"

namespace lnorm {

// the smallest instantiated 32-bit width (blk32_has_width) holding ny columns
int blk32_words(int d, int ny) {
  if (d < 2 || d > 10) return 0;
  for (int w = 8; w <= 96; w += 8)
    if (w >= ny && blk32_has_width(d, w)) return w;
  return 0;
}

int blk32_rows(int d, int words) { return blk16_low_rows(d, words); }

hipError_t launch_blk32_lo(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.d) {
    case 2: return launch_blk<2, 32>(mode, a, st, block);
    case 3: return launch_blk<3, 32>(mode, a, st, block);
    case 4: return launch_blk<4, 32>(mode, a, st, block);
    default: return hipErrorInvalidValue;
  }
}

int blocks_per_cu_blk32_lo(int d, int words, int block) {
  switch (d) {
    case 2: return blocks_per_cu_blk<2, 32>(words, block);
    case 3: return blocks_per_cu_blk<3, 32>(words, block);
    case 4: return blocks_per_cu_blk<4, 32>(words, block);
    default: return 0;
  }
}

int blk32_blocks_per_cu(int d, int words, int block) {
  return d <= 4 ? blocks_per_cu_blk32_lo(d, words, block) : blocks_per_cu_blk32_hi(d, words, block);
}

hipError_t launch_blk32(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (block <= 0 || block % 64 || block > 256) block = 256;
  const int words = a.ny_pad;
  if (blk32_words(a.d, a.ny_pad) != words) return hipErrorInvalidValue;
  if (a.q < blk16_low_rows(a.d, words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  return a.d <= 4 ? launch_blk32_lo(mode, a, st, block) : launch_blk32_hi(mode, a, st, block);
}

}  // namespace lnorm

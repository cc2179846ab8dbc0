// lnorm_tab32_bin.hip -- the tabulated binary walk (L1 / LM) in 32-bit
// words (walk_tab16_bin with WB = 32): one |.| column per word, biased by
// 2^31, v_sad_u32, for int32 inputs whose |.| column sums exceed 32767.
// Widths, the launch entry and the kernels (lnorm_tab_bin.h).
#include "lnorm_tab_bin.h"

namespace lnorm {

// multiples of 8 up to 48 words, then 64, 80, 96, 128, 160 and 192
static const int kTab32BinWords[] = {8, 16, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192};

int tab32_bin_words(int nabs) {
  for (int w : kTab32BinWords) if (w >= nabs) return w;
  return 0;
}

#define LN_TB32_WIDTHS(X) X(8) X(16) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192)

template <bool MARG>
static hipError_t launch_tb32(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.ny_pad) {
#define X(n) case n: return launch_tb_nw<n, MARG, 32>(mode, a, st, block);
    LN_TB32_WIDTHS(X)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

template <bool MARG>
static int blocks_per_cu_tb32(int words, int block) {
  switch (words) {
#define X(n) case n: return blocks_per_cu_tb_nw<n, MARG, 32>(block);
    LN_TB32_WIDTHS(X)
#undef X
    default: break;
  }
  return 0;
}

int tab32_bin_blocks_per_cu(Kind kind, int words, int block) {
  if (kind == KIND_L1) return blocks_per_cu_tb32<false>(words, block);
  if (kind == KIND_LM) return blocks_per_cu_tb32<true>(words, block);
  return 0;
}

hipError_t launch_tab32_bin(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (block <= 0 || block % 64 || block > 256) block = 256;
  const int words = a.ny_pad;
  if (tab32_bin_words(words) != words) return hipErrorInvalidValue;
  if (a.q < tab32_bin_rows(words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  if (kind == KIND_L1) return launch_tb32<false>(mode, a, st, block);
  if (kind == KIND_LM) return launch_tb32<true>(mode, a, st, block);
  return hipErrorInvalidValue;
}

}  // namespace lnorm

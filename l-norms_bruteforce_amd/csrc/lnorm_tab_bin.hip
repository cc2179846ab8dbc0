// lnorm_tab_bin.hip -- the tabulated binary walk (L1 / LM) in 16-bit words:
// widths, the launch entry and the kernels (lnorm_tab_bin.h).
#include "lnorm_tab_bin.h"

namespace lnorm {

// every even width up to 32 words (the blocked kernel's widths), then steps of 8 and 16
static const int kTabBinWords[] = {4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 40, 48, 56, 64, 80, 96, 128, 160, 192};

int tab16_bin_words(int nabs) {
  const int need = (nabs + 1) / 2;
  for (int w : kTabBinWords) if (w >= need) return w;
  return 0;
}

#define LN_TB16_WIDTHS(X) X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) \
  X(40) X(48) X(56) X(64) X(80) X(96) X(128) X(160) X(192)

template <bool MARG>
static hipError_t launch_tb(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.ny_pad / 2) {
#define X(n) case n: return launch_tb_nw<n, MARG, 16>(mode, a, st, block);
    LN_TB16_WIDTHS(X)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

template <bool MARG>
static int blocks_per_cu_tb(int words, int block) {
  switch (words) {
#define X(n) case n: return blocks_per_cu_tb_nw<n, MARG, 16>(block);
    LN_TB16_WIDTHS(X)
#undef X
    default: break;
  }
  return 0;
}

int tab16_bin_blocks_per_cu(Kind kind, int words, int block) {
  if (kind == KIND_L1) return blocks_per_cu_tb<false>(words, block);
  if (kind == KIND_LM) return blocks_per_cu_tb<true>(words, block);
  return 0;
}

hipError_t launch_tab16_bin(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (block <= 0 || block % 64 || block > 256) block = 256;
  const int words = a.ny_pad / 2;
  if (words * 2 != a.ny_pad || tab16_bin_words(2 * words) != words) return hipErrorInvalidValue;
  if (a.q < tab_bin_rows(words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  if (kind == KIND_L1) return launch_tb<false>(mode, a, st, block);
  if (kind == KIND_LM) return launch_tb<true>(mode, a, st, block);
  return hipErrorInvalidValue;
}

}  // namespace lnorm

// lnorm_tab_bin.hip -- the tabulated binary walk (L1 / LM), one lane per
// segment: the 4 or 5 fastest rows as a sign-pattern table (walk_tab16_bin).
#include "lnorm_walks.h"

namespace lnorm {

template <int NW, bool MARG, int MODE>
__global__ __launch_bounds__(256) void k_tab16_bin(
    const uint32_t *__restrict__ Ap, const uint32_t *__restrict__ tab, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst) {
  walk_tab16_bin<NW, MARG, MODE>(as_const(Ap), as_const(tab), R, q, seg0, nseg, target, unit_best, gbest, gfirst,
                                 global_wave());
}

// every even width up to 32 words (the blocked kernel's widths), then steps of 8 and 16
static const int kTabBinWords[] = {4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 40, 48, 56, 64, 80, 96, 128, 160, 192};

int tab16_bin_words(int nabs) {
  const int need = (nabs + 1) / 2;
  for (int w : kTabBinWords) if (w >= need) return w;
  return 0;
}

template <int NW, bool MARG>
static hipError_t launch_tb_nw(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const int wpb = block / 64;
  const dim3 grid((unsigned)((waves + wpb - 1) / wpb)), blk(block);
  const uint32_t *Ap = (const uint32_t *)a.A;
  const uint32_t *tab = (const uint32_t *)a.D;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_tab16_bin<NW, MARG, MODE_MAX>), grid, blk, 0, st, Ap, tab, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_tab16_bin<NW, MARG, MODE_FIRST>), grid, blk, 0, st, Ap, tab, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

template <bool MARG>
static hipError_t launch_tb(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.ny_pad / 2) {
#define X(n) case n: return launch_tb_nw<n, MARG>(mode, a, st, block);
    X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) X(40) X(48) X(56) X(64) \
    X(80) X(96) X(128) X(160) X(192)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

// workgroups of k_tab16_bin<words, MARG, MODE_MAX> resident per CU at this
// block size (0: unknown)
template <bool MARG>
static int blocks_per_cu_tb(int words, int block) {
  int nb = 0;
  switch (words) {
#define X(n)                                                                                                      \
  case n:                                                                                                         \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tab16_bin<n, MARG, MODE_MAX>, block, 0) != hipSuccess) \
      nb = 0;                                                                                                     \
    break;
    X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) X(40) X(48) X(56) X(64) \
    X(80) X(96) X(128) X(160) X(192)
#undef X
    default: break;
  }
  return nb;
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

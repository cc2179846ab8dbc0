// lnorm_dary.hip -- the tabulated packed walks, one lane per segment: d-ary
// (d = 2..10, the M fastest rows as per-label subset tables, walk_blk16) and
// binary L1 / LM (the 5 fastest rows as a sign-pattern table, walk_tab16_bin).
#include "lnorm_walks.h"

namespace lnorm {

template <int D, int NW, int MODE>
__global__ __launch_bounds__(256) void k_blk16_dary(
    const uint32_t *__restrict__ Ap, const uint32_t *__restrict__ neg, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst) {
  walk_blk16<D, NW, blk16_low_rows(D, NW), MODE>(as_const(Ap), as_const(neg), R, q, seg0, nseg, target, unit_best,
                                                  gbest, gfirst, global_wave());
}

static const int kBlkWords[] = {8, 16, 24, 32, 40, 48, 56, 64, 80, 96};

int blk16_words(int d, int ny) {
  const int need = (ny + 1) / 2;
  const int mx = blk16_max_words(d);
  if (d < 2 || d > 10) return 0;
  for (int w : kBlkWords) if (w >= need && w <= mx) return w;
  return 0;
}

int blk16_rows(int d, int words) { return blk16_low_rows(d, words); }

template <int D, int NW>
static hipError_t launch_blk_nw(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const int wpb = block / 64;
  const dim3 grid((unsigned)((waves + wpb - 1) / wpb)), blk(block);
  const uint32_t *Ap = (const uint32_t *)a.A;
  const uint32_t *neg = (const uint32_t *)a.D;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_blk16_dary<D, NW, MODE_MAX>), grid, blk, 0, st, Ap, neg, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_blk16_dary<D, NW, MODE_FIRST>), grid, blk, 0, st, Ap, neg, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

template <int D>
static hipError_t launch_blk(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.ny_pad / 2) {
#define X(n) case n: if constexpr (n <= blk16_max_words(D)) return launch_blk_nw<D, n>(mode, a, st, block); break;
    X(8) X(16) X(24) X(32) X(40) X(48) X(56) X(64) X(80) X(96)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_blk16(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (block <= 0 || block % 64 || block > 256) block = 256;
  const int words = a.ny_pad / 2;
  if (blk16_words(a.d, a.ny_pad) != words || words * 2 != a.ny_pad) return hipErrorInvalidValue;
  if (a.q < blk16_low_rows(a.d, words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  switch (a.d) {
    case 2: return launch_blk<2>(mode, a, st, block);
    case 3: return launch_blk<3>(mode, a, st, block);
    case 4: return launch_blk<4>(mode, a, st, block);
    case 5: return launch_blk<5>(mode, a, st, block);
    case 6: return launch_blk<6>(mode, a, st, block);
    case 7: return launch_blk<7>(mode, a, st, block);
    case 8: return launch_blk<8>(mode, a, st, block);
    case 9: return launch_blk<9>(mode, a, st, block);
    case 10: return launch_blk<10>(mode, a, st, block);
    default: break;
  }
  return hipErrorInvalidValue;
}

// ---- tabulated binary walk (L1 / LM) -----------------------------------------
template <int NW, bool MARG, int MODE>
__global__ __launch_bounds__(256) void k_tab16_bin(
    const uint32_t *__restrict__ Ap, const uint32_t *__restrict__ tab, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst) {
  walk_tab16_bin<NW, MARG, MODE>(as_const(Ap), as_const(tab), R, q, seg0, nseg, target, unit_best, gbest, gfirst,
                                 global_wave());
}

static const int kTabBinWords[] = {4, 8, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64, 80, 96, 128, 160, 192};

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
    X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(40) X(48) X(56) X(64) X(80) X(96) X(128) X(160) X(192)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
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

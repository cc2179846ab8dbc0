// lnorm_blk16.h -- the tabulated d-ary walk, one lane per segment
// (walk_blk16): its kernel and launchers in both word formats, instantiated
// by lnorm_dary.hip (16-bit, d = 2..4), lnorm_dary_hi.hip (16-bit,
// d = 5..10) and lnorm_dary32.hip (32-bit), which compile in parallel.
#pragma once

#include "lnorm_walks.h"

namespace lnorm {

// WB: the word format (16: packed halves, NW words hold 2 NW columns; 32:
// one column per word).  The registers of a shape depend on NW alone, so
// both formats share the table rows blk16_low_rows(d, NW).
template <int D, int NW, int MODE, int WB>
__global__ __launch_bounds__(256) void k_blk16_dary(
    const uint32_t *__restrict__ Ap, const uint32_t *__restrict__ neg, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst) {
  walk_blk16<D, NW, blk16_low_rows(D, NW), MODE, WB>(as_const(Ap), as_const(neg), R, q, seg0, nseg, target,
                                                      unit_best, gbest, gfirst, global_wave());
}

template <int D, int NW, int WB>
static inline hipError_t launch_blk_nw(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const int wpb = block / 64;
  const dim3 grid((unsigned)((waves + wpb - 1) / wpb)), blk(block);
  const uint32_t *Ap = (const uint32_t *)a.A;
  const uint32_t *neg = (const uint32_t *)a.D;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_blk16_dary<D, NW, MODE_MAX, WB>), grid, blk, 0, st, Ap, neg, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_blk16_dary<D, NW, MODE_FIRST, WB>), grid, blk, 0, st, Ap, neg, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

// instantiated widths in words: multiples of 4 up to 32 (every even count for
// d <= 4), then 40..64 in steps of 8, 80 and 96; at most one wave's table
constexpr bool blk16_has_width(int d, int nw) {
  const bool listed = nw <= 32 ? nw >= 4 && (nw % 4 == 0 || (d <= 4 && nw % 2 == 0))
                               : (nw == 40 || nw == 48 || nw == 56 || nw == 64 || nw == 80 || nw == 96);
  return listed && nw <= blk16_max_words(d);
}
// 32-bit words (one column each; inputs whose column sums exceed 16 bits):
// 8, 16, 24, 32, 48, 64 and 96 words, and the widest word count of the
// 16-bit walk where that is not listed (40 at d = 5).  A coarser list than the
// 16-bit one: each instantiation compiles about 2.5x longer.
constexpr bool blk32_has_width(int d, int nw) {
  const bool listed = nw == 8 || nw == 16 || nw == 24 || nw == 32 || nw == 48 || nw == 64 || nw == 96 ||
                      nw == blk16_max_words(d);
  return listed && nw <= blk16_max_words(d);
}
template <int WB>
constexpr bool blk_has_width(int d, int nw) { return WB == 16 ? blk16_has_width(d, nw) : blk32_has_width(d, nw); }

template <int D, int WB>
static inline hipError_t launch_blk(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (WB == 16 ? a.ny_pad / 2 : a.ny_pad) {
#define X(n) case n: if constexpr (blk_has_width<WB>(D, n)) return launch_blk_nw<D, n, WB>(mode, a, st, block); break;
    X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) X(40) X(48) X(56) \
    X(64) X(80) X(96)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

// workgroups of k_blk16_dary<D, words, MODE_MAX, WB> resident per CU at this
// block size (0: unknown)
template <int D, int WB>
static inline int blocks_per_cu_blk(int words, int block) {
  int nb = 0;
  switch (words) {
#define X(n)                                                                                           \
  case n:                                                                                              \
    if constexpr (blk_has_width<WB>(D, n))                                                             \
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_blk16_dary<D, n, MODE_MAX, WB>, block,   \
                                                       0) != hipSuccess)                               \
        nb = 0;                                                                                        \
    break;
    X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) X(40) X(48) X(56) \
    X(64) X(80) X(96)
#undef X
    default: break;
  }
  return nb;
}

// one translation unit's share of launch_blk16 and blk16_blocks_per_cu
hipError_t launch_blk16_lo(Mode mode, const PlainArgs &a, hipStream_t st, int block);   // d = 2..4
hipError_t launch_blk16_hi(Mode mode, const PlainArgs &a, hipStream_t st, int block);   // d = 5..10
int blocks_per_cu_blk16_lo(int d, int words, int block);
int blocks_per_cu_blk16_hi(int d, int words, int block);
// the 32-bit walk's share (lnorm_dary32.hip)
hipError_t launch_blk32_lo(Mode mode, const PlainArgs &a, hipStream_t st, int block);   // d = 2..4
hipError_t launch_blk32_hi(Mode mode, const PlainArgs &a, hipStream_t st, int block);   // d = 5..10
int blocks_per_cu_blk32_lo(int d, int words, int block);
int blocks_per_cu_blk32_hi(int d, int words, int block);

}  // namespace lnorm

// lnorm_tab_bin.h -- the tabulated binary walk (L1 / LM), one lane per
// segment: the 4 or 5 fastest rows as a sign-pattern table
// (walk_tab16_bin).  Its kernel and launchers in both word formats,
// instantiated by lnorm_tab_bin.hip (16-bit) and lnorm_tab32_bin.hip
// (32-bit), which compile in parallel.
#pragma once

#include "lnorm_walks.h"

namespace lnorm {

template <int NW, bool MARG, int MODE, int WB>
__global__ __launch_bounds__(256) void k_tab16_bin(
    const uint32_t *__restrict__ Ap, const uint32_t *__restrict__ tab, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst) {
  walk_tab16_bin<NW, MARG, MODE, WB>(as_const(Ap), as_const(tab), R, q, seg0, nseg, target, unit_best, gbest, gfirst,
                                     global_wave());
}

template <int NW, bool MARG, int WB>
static hipError_t launch_tb_nw(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const int wpb = block / 64;
  const dim3 grid((unsigned)((waves + wpb - 1) / wpb)), blk(block);
  const uint32_t *Ap = (const uint32_t *)a.A;
  const uint32_t *tab = (const uint32_t *)a.D;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_tab16_bin<NW, MARG, MODE_MAX, WB>), grid, blk, 0, st, Ap, tab, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_tab16_bin<NW, MARG, MODE_FIRST, WB>), grid, blk, 0, st, Ap, tab, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

// workgroups of k_tab16_bin<NW, MARG, MODE_MAX, WB> resident per CU at this
// block size (0: unknown)
template <int NW, bool MARG, int WB>
static int blocks_per_cu_tb_nw(int block) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tab16_bin<NW, MARG, MODE_MAX, WB>, block, 0) != hipSuccess)
    nb = 0;
  return nb;
}

}  // namespace lnorm

// lnorm_batch.h -- what the batched walks' translation units share: the
// persistent waves' task counter and the launch grid.
#pragma once

#include "lnorm_walks.h"

namespace lnorm {

// The wave's next task.  A wave starts on the counter shard of its XCD
// (workgroups are dispatched to the 8 XCDs round-robin by index) and moves
// to the next shard when that one is exhausted, so every task is handed out
// once and the waves of an XCD that finishes early help the others.  Every
// lane executes the atomic (lane 0 adds 1, the others 0): no control flow
// depends on the lane index.  A `if (lane == 0)` here, loop-invariant but
// lane-divergent, let the compiler split the persistent loop by lane, and
// lanes 1..63 then looped on task 0 forever.
__device__ __forceinline__ bool next_task(unsigned int *counters, uint64_t ntasks, int &shard, int &left,
                                          uint32_t &t) {
  while (left > 0) {
    // readfirstlane: the shard is wave-uniform, and the compiler must see it
    // so; with a uniform address it issues one atomic per wave (the lane
    // addends are summed first) instead of 64 same-address atomics.
    const int us = __builtin_amdgcn_readfirstlane(shard);
    const uint32_t k = atomicAdd(counters + (size_t)us * kCounterStride, wave_lane() == 0 ? 1u : 0u);
    const uint64_t cand = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)k) * kCounterShards + shard;
    if (cand < ntasks) {
      t = (uint32_t)cand;
      return true;
    }
    shard = (shard + 1) & (kCounterShards - 1);
    --left;
  }
  return false;
}

// grid = min(tasks, resident waves); the waves loop over tasks
static inline unsigned batch_grid(uint64_t ntasks) {
  const uint64_t blocks = (ntasks + 3) / 4;            // 4 waves per 256-thread block
  const uint64_t cap = 256ull * 8;                      // 256 CUs x up to 8 blocks resident
  return (unsigned)(blocks < cap ? (blocks ? blocks : 1) : cap);
}

// The batched tabulated d-ary walk (walk_blk16 per task), instantiated by
// lnorm_batch_tab.hip (d = 2..4) and lnorm_batch_tab_hi.hip (d = 5..10).
template <int D, int NW, int MODE>
__global__ __launch_bounds__(256) void k_batch_t16(const BatchTask *__restrict__ tasks, int nitems,
                                                   uint64_t ntasks, unsigned int *__restrict__ counter) {
  int shard = (int)(blockIdx.x & (kCounterShards - 1)), left = kCounterShards;
  uint32_t t;
  while (next_task(counter, ntasks, shard, left, t)) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tasks[mid].task0 <= t) lo = mid;
      else hi = mid - 1;
    }
    const BatchTask &it = tasks[lo];
    uint64_t seg0 = 0;
    int32_t target = 0;
    if (MODE == MODE_FIRST) {
      seg0 = *it.unit * it.seg_base;
      target = *it.gbest;
    }
    walk_blk16<D, NW, blk16_low_rows(D, NW), MODE>(as_const((const uint32_t *)it.A), as_const((const uint32_t *)it.D),
                                                    it.R, it.q, seg0, it.nseg, target, it.unit_best, it.gbest,
                                                    it.gfirst, t - it.task0);
  }
}

template <int D, int NW>
static inline hipError_t lb_t16(Mode mode, const BatchLaunch &b, hipStream_t st) {
  const dim3 grid(batch_grid(b.ntasks)), blk(256);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_batch_t16<D, NW, MODE_MAX>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks, b.counter);
  else
    hipLaunchKernelGGL((k_batch_t16<D, NW, MODE_FIRST>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks, b.counter);
  return hipGetLastError();
}

hipError_t launch_batch_t16_hi(Mode mode, const BatchLaunch &b, hipStream_t st);   // d = 5..10

}  // namespace lnorm

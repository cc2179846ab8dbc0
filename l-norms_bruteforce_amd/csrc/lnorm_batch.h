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

}  // namespace lnorm

// lnorm_batch.hip -- many problems of uneven shape and d in one persistent
// launch per kernel instantiation.
//
// Every item of a batch is a calc_Lnorm problem (L_CUDA.cu:281-359).  Items
// that share an instantiation (walk, kind, width) form a group; one launch
// per group covers all of its items.  A task is one wave of one item (the 64
// lane segments lane_segment(w, .)); waves of the persistent grid take tasks
// from an atomic counter until none is left, so large and small items
// balance over the chip without a launch per item.  Items are ordered by
// decreasing work inside a group (longest first).
//
// The two phases of the single-problem search run device-side for all items:
// MODE_MAX records each item's unit maxima, k_batch_first_unit finds each
// item's first unit attaining its maximum, and MODE_FIRST re-walks that unit,
// reading the unit index and the target from device memory.  Nothing waits
// on the host in between, so the whole sequence can be captured in a HIP
// graph and replayed.
#include "lnorm_batch.h"

namespace lnorm {

template <int KIND, int NY, int MODE>
__global__ __launch_bounds__(256) void k_batch_bin(const BatchTask *__restrict__ tasks, int nitems,
                                                   uint64_t ntasks, unsigned int *__restrict__ counter,
                                                   uint32_t B) {
  int shard = (int)(blockIdx.x & (kCounterShards - 1)), left = kCounterShards;
  uint32_t t;
  while (next_task(counter, ntasks, shard, left, t)) {   // every wave ends when all shards are empty
    int lo = 0, hi = nitems - 1;                 // last item with task0 <= t
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
    walk_bin<KIND, NY, MODE>(as_const(it.A), as_const(it.D), it.R, it.q, seg0, it.nseg, target, it.unit_best, it.gbest, it.gfirst, B,
                             t - it.task0);
  }
}

template <int D, int NW, int MODE>
__global__ __launch_bounds__(256) void k_batch_p16(const BatchTask *__restrict__ tasks, int nitems,
                                                   uint64_t ntasks, unsigned int *__restrict__ counter,
                                                   uint32_t B) {
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
    walk_p16<D, NW, MODE>(as_const((const uint32_t *)it.A), it.R, it.q, seg0, it.nseg, target, it.unit_best, it.gbest,
                          it.gfirst, B, t - it.task0);
  }
}

template <int NW, bool MARG, int MODE>
__global__ __launch_bounds__(256) void k_batch_tb16(const BatchTask *__restrict__ tasks, int nitems,
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
    walk_tab16_bin<NW, MARG, MODE>(as_const((const uint32_t *)it.A), as_const((const uint32_t *)it.D), it.R, it.q,
                                   seg0, it.nseg, target, it.unit_best, it.gbest, it.gfirst, t - it.task0);
  }
}

// One workgroup per item: the smallest unit index whose best equals the
// item's maximum.
__global__ __launch_bounds__(256) void k_batch_first_unit(const BatchTask *__restrict__ tasks,
                                                          unsigned long long *__restrict__ first) {
  __shared__ unsigned long long red[4];
  const BatchTask &it = tasks[blockIdx.x];
  const int32_t target = *it.gbest;
  unsigned long long mine = ~0ull;
  for (uint64_t i = threadIdx.x; i < it.nseg; i += blockDim.x)
    if (it.unit_best[i] == target) {
      mine = i;
      break;
    }
  const uint64_t m = wave_min_u64(mine);
  if (wave_lane() == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = red[w] < r ? red[w] : r;
    first[blockIdx.x] = r;
  }
}

hipError_t launch_batch_first_unit(const BatchTask *tasks, int nitems, unsigned long long *first, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_first_unit, dim3((unsigned)nitems), dim3(256), 0, st, tasks, first);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers: grid = min(tasks, resident waves); the waves loop over tasks
// ---------------------------------------------------------------------------

template <int KIND, int NY>
static hipError_t lb_bin(Mode mode, const BatchLaunch &b, hipStream_t st) {
  const dim3 grid(batch_grid(b.ntasks)), blk(256);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_batch_bin<KIND, NY, MODE_MAX>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks, b.counter,
                       kBias32);
  else
    hipLaunchKernelGGL((k_batch_bin<KIND, NY, MODE_FIRST>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks,
                       b.counter, kBias32);
  return hipGetLastError();
}

template <int D, int NW>
static hipError_t lb_p16(Mode mode, const BatchLaunch &b, hipStream_t st) {
  const dim3 grid(batch_grid(b.ntasks)), blk(256);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_batch_p16<D, NW, MODE_MAX>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks, b.counter,
                       kBias16x2);
  else
    hipLaunchKernelGGL((k_batch_p16<D, NW, MODE_FIRST>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks,
                       b.counter, kBias16x2);
  return hipGetLastError();
}

template <int NW, bool MARG>
static hipError_t lb_tb16(Mode mode, const BatchLaunch &b, hipStream_t st) {
  const dim3 grid(batch_grid(b.ntasks)), blk(256);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_batch_tb16<NW, MARG, MODE_MAX>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks, b.counter);
  else
    hipLaunchKernelGGL((k_batch_tb16<NW, MARG, MODE_FIRST>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks,
                       b.counter);
  return hipGetLastError();
}

// Width in words of a batch item's tabulated L1 / LM walk: even counts from 4
// up to 32 words, then 40..64 in steps of 8 (the instantiations below).
int batch_tb16_words(int nabs) {
  const int need = std::max(1, (nabs + 1) / 2);
  if (need <= 32) return std::max(4, (need + 1) / 2 * 2);
  for (int w : {40, 48, 56, 64})
    if (need <= w) return w;
  return 0;
}

template <bool MARG>
static hipError_t lb_tb16_w(Mode mode, const BatchLaunch &b, hipStream_t st) {
  switch (b.width / 2) {
#define X(n) case n: return lb_tb16<n, MARG>(mode, b, st);
    X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) X(40) X(48) X(56) X(64)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

template <int KIND>
static hipError_t lb_bin_w(Mode mode, const BatchLaunch &b, hipStream_t st) {
  switch (b.width) {
#define X(n) case n: return lb_bin<KIND, n>(mode, b, st);
    X(8) X(16) X(24) X(32) X(40) X(48) X(56) X(64) X(80) X(96) X(112) X(128)
#undef X
    default: return hipErrorInvalidValue;
  }
}

template <int D>
static hipError_t lb_p16_w(Mode mode, const BatchLaunch &b, hipStream_t st) {
  switch (b.width / 2) {
#define X(n) case n: if constexpr (n <= dary16_max_words(D)) return lb_p16<D, n>(mode, b, st); break;
    X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(40) X(48) X(64)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_walk(Mode mode, const BatchLaunch &b, hipStream_t st) {
  if (b.nitems <= 0) return hipSuccess;
  if (!b.packed) {
    if (b.kind == KIND_L1) return lb_bin_w<KIND_L1>(mode, b, st);
    if (b.kind == KIND_LM) return lb_bin_w<KIND_LM>(mode, b, st);
    return hipErrorInvalidValue;
  }
  if (b.table && b.kind == KIND_L1) return lb_tb16_w<false>(mode, b, st);
  if (b.table && b.kind == KIND_LM) return lb_tb16_w<true>(mode, b, st);
  if (b.table) return launch_batch_t16(mode, b, st);
  switch (b.d) {
    case 2: return lb_p16_w<2>(mode, b, st);
    case 3: return lb_p16_w<3>(mode, b, st);
    case 4: return lb_p16_w<4>(mode, b, st);
    case 5: return lb_p16_w<5>(mode, b, st);
    case 6: return lb_p16_w<6>(mode, b, st);
    case 7: return lb_p16_w<7>(mode, b, st);
    case 8: return lb_p16_w<8>(mode, b, st);
    case 9: return lb_p16_w<9>(mode, b, st);
    case 10: return lb_p16_w<10>(mode, b, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace lnorm

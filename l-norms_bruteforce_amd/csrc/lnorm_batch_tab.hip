// lnorm_batch_tab.hip -- the batched tabulated d-ary walk (walk_blk16 per
// task; the task scheme is that of lnorm_batch.hip).  Batch items are small
// (at most 32 words), so widths step by 4 words: an item pads to at most 3
// words beyond its own columns.
#include "lnorm_batch.h"

namespace lnorm {

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

// Width in words of a batch item's tabulated walk: the smallest multiple of 4
// holding its columns, at most 32 and at most one wave's table width.
int batch_t16_words(int d, int ny) {
  if (d < 2 || d > 10) return 0;
  const int w = std::max(4, ((ny + 1) / 2 + 3) / 4 * 4);
  return w <= 32 && w <= blk16_max_words(d) ? w : 0;
}

template <int D, int NW>
static hipError_t lb_t16(Mode mode, const BatchLaunch &b, hipStream_t st) {
  const dim3 grid(batch_grid(b.ntasks)), blk(256);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_batch_t16<D, NW, MODE_MAX>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks, b.counter);
  else
    hipLaunchKernelGGL((k_batch_t16<D, NW, MODE_FIRST>), grid, blk, 0, st, b.tasks, b.nitems, b.ntasks, b.counter);
  return hipGetLastError();
}

template <int D>
static hipError_t lb_t16_w(Mode mode, const BatchLaunch &b, hipStream_t st) {
  if (batch_t16_words(D, b.width) * 2 != b.width) return hipErrorInvalidValue;
  switch (b.width / 2) {
#define X(n) case n: if constexpr (n <= blk16_max_words(D)) return lb_t16<D, n>(mode, b, st); break;
    X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_t16(Mode mode, const BatchLaunch &b, hipStream_t st) {
  switch (b.d) {
    case 2: return lb_t16_w<2>(mode, b, st);
    case 3: return lb_t16_w<3>(mode, b, st);
    case 4: return lb_t16_w<4>(mode, b, st);
    case 5: return lb_t16_w<5>(mode, b, st);
    case 6: return lb_t16_w<6>(mode, b, st);
    case 7: return lb_t16_w<7>(mode, b, st);
    case 8: return lb_t16_w<8>(mode, b, st);
    case 9: return lb_t16_w<9>(mode, b, st);
    case 10: return lb_t16_w<10>(mode, b, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace lnorm

// lnorm_batch_tab.hip -- the batched tabulated d-ary walk (k_batch_t16 in
// lnorm_batch.h; the task scheme is that of lnorm_batch.hip): widths, the
// launch entry, and the kernels for d = 2..4 (d = 5..10 are in
// lnorm_batch_tab_hi.hip).  Batch items are small (at most 32 words), so
// widths step by 2 words for d <= 4 and by 4 words beyond: an item pads to at
// most 1 or 3 words beyond its own columns.
#include "lnorm_batch.h"

namespace lnorm {

// Width in words of a batch item's tabulated walk: the smallest multiple of 2
// (d <= 4) or 4 (d >= 5) holding its columns, at least 4, at most 32 and at
// most one wave's table width.
int batch_t16_words(int d, int ny) {
  if (d < 2 || d > 10) return 0;
  const int step = d <= 4 ? 2 : 4;
  const int w = std::max(4, ((ny + 1) / 2 + step - 1) / step * step);
  return w <= 32 && w <= blk16_max_words(d) ? w : 0;
}

template <int D>
static hipError_t lb_t16_w(Mode mode, const BatchLaunch &b, hipStream_t st) {
  if (batch_t16_words(D, b.width) * 2 != b.width) return hipErrorInvalidValue;
  switch (b.width / 2) {
#define X(n) case n: return lb_t16<D, n>(mode, b, st);
    X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32)
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
    default: return b.d <= 10 ? launch_batch_t16_hi(mode, b, st) : hipErrorInvalidValue;
  }
}

}  // namespace lnorm

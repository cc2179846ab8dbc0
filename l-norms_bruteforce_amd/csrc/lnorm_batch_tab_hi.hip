// lnorm_batch_tab_hi.hip -- the batched tabulated d-ary walk's kernels for
// d = 5..10 (k_batch_t16 in lnorm_batch.h), widths of 4..32 words in steps of
// 4 up to one wave's table width.
#include "lnorm_batch.h"

namespace lnorm {

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

hipError_t launch_batch_t16_hi(Mode mode, const BatchLaunch &b, hipStream_t st) {
  switch (b.d) {
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

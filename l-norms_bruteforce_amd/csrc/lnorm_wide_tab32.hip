// lnorm_wide_tab32.hip -- the W-wave tabulated d-ary walk in 32-bit words
// (lnorm_wide_tab.h with WB = 32: one column per word, biased by 2^31,
// v_sad_u32) for int32 inputs whose column sums exceed 32767: d = 2..10, W
// waves of 32 words up to 256 columns at d <= 6, 128 at d = 7, 96 at d = 8,
// 9 and 64 at d = 10.  The kernels for d = 2..4 and the launch entry; d =
// 5..10 in lnorm_wide_tab32_hi.hip.  The L1 / LM walk is in
// lnorm_wide_tb32.hip.
#include "lnorm_wide_tab.h"

namespace lnorm {

int wt32_words(int d, int ny) { return wt_words_for(d, ny); }

int wt32_rows(int d, int words) { return wt_rows_for(d, words); }

hipError_t launch_wide_t32(Mode mode, const PlainArgs &a, hipStream_t st) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  const int words = a.ny_pad;
  if (wt32_words(a.d, a.ny_pad) != words) return hipErrorInvalidValue;
  if (a.q < wt32_rows(a.d, words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  switch (a.d) {
    case 2: return lwt_d<2, 32>(mode, a, st);
    case 3: return lwt_d<3, 32>(mode, a, st);
    case 4: return lwt_d<4, 32>(mode, a, st);
    default: return launch_wide_t32_hi(mode, a, st);
  }
}

}  // namespace lnorm

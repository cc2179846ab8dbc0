// lnorm_wide_tb32.hip -- the W-wave tabulated L1 / LM walk in 32-bit words
// (lnorm_wide_tab.h with WB = 32: one |.| column per word, biased by 2^31,
// v_sad_u32) for int32 inputs whose |.| column sums exceed 32767: 8 to 16
// waves of 32 words, 16 of 64; 193..1024 |.| columns.
#include "lnorm_wide_tab.h"

namespace lnorm {

int wtb32_words(int nabs) { return wtb_words_for(nabs); }

int wtb32_rows(int words) { return wtb_m(words == 1024 ? 16 : words / 32); }

hipError_t launch_wide_tb32(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  const int words = a.ny_pad;
  if (wtb32_words(words) != words) return hipErrorInvalidValue;
  if (a.q < wtb32_rows(words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  if (kind == KIND_L1) return lwtb_m<false, 32>(mode, a, st);
  if (kind == KIND_LM) return lwtb_m<true, 32>(mode, a, st);
  return hipErrorInvalidValue;
}

}  // namespace lnorm

// lnorm_wide_tab.hip -- the W-wave tabulated walks in 16-bit words
// (lnorm_wide_tab.h): d-ary (d = 2..10) and L1 / LM.
#include "lnorm_wide_tab.h"

namespace lnorm {

// Widths of the W-wave tabulated L1 / LM walk: 256, 320, 384, 512 or 1024
// words (8, 10, 12 or 16 waves of 32 words, 16 of 64; up to 2048 |.| columns).
int wtb16_words(int nabs) { return wtb_words_for((nabs + 1) / 2); }

int wtb16_rows(int words) { return wtb_m(words == 1024 ? 16 : words / 32); }

hipError_t launch_wide_tb16(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  const int words = a.ny_pad / 2;
  if (words * 2 != a.ny_pad || wtb16_words(2 * words) != words) return hipErrorInvalidValue;
  if (a.q < wtb16_rows(words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  if (kind == KIND_L1) return lwtb_m<false, 16>(mode, a, st);
  if (kind == KIND_LM) return lwtb_m<true, 16>(mode, a, st);
  return hipErrorInvalidValue;
}

int wt16_words(int d, int ny) { return wt_words_for(d, (ny + 1) / 2); }

int wt16_rows(int d, int words) { return wt_rows_for(d, words); }

hipError_t launch_wide_t16(Mode mode, const PlainArgs &a, hipStream_t st) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  const int words = a.ny_pad / 2;
  if (wt16_words(a.d, a.ny_pad) != words || words * 2 != a.ny_pad) return hipErrorInvalidValue;
  if (a.q < wt16_rows(a.d, words) || a.R - 1 < a.q) return hipErrorInvalidValue;
  return lwt_any<16>(mode, a, st);
}

}  // namespace lnorm

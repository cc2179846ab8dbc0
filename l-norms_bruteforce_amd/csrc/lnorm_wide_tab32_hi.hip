// lnorm_wide_tab32_hi.hip -- the W-wave tabulated d-ary walk in 32-bit words
// (lnorm_wide_tab.h with WB = 32) for d = 5..10; the launch entry, its
// checks and d = 2..4 are in lnorm_wide_tab32.hip.
#include "lnorm_wide_tab.h"

namespace lnorm {

hipError_t launch_wide_t32_hi(Mode mode, const PlainArgs &a, hipStream_t st) {
  switch (a.d) {
    case 5: return lwt_d<5, 32>(mode, a, st);
    case 6: return lwt_d<6, 32>(mode, a, st);
    case 7: return lwt_d<7, 32>(mode, a, st);
    case 8: return lwt_d<8, 32>(mode, a, st);
    case 9: return lwt_d<9, 32>(mode, a, st);
    case 10: return lwt_d<10, 32>(mode, a, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace lnorm

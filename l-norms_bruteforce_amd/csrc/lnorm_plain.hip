// lnorm_plain.hip -- the plain Gray-code walks on gfx950.
//
// Work split: the ranks of the Gray enumeration (the reference's
// total_num_to_calc, L_OpenMP.c:352-367) are cut into aligned segments of
// 2^q (binary) or d^q (d-ary) consecutive ranks; each lane of a wave walks
// one segment.  Because every segment starts at a multiple of its length,
// the row flipped at step t (ctz(t), or the lowest non-zero base-d digit
// of t) is the same for all 64 lanes, and lane_segment() gives the lanes of a
// wave segments of equal parity so the flip direction is uniform too.  The
// row delta therefore arrives through scalar loads and SGPR operands; the
// per-lane state (the partial sums y, or t_c for d labels) lives in VGPRs.
//
// Objectives (reference kernels L_OpenMP.c:15-239):
//   L1: sum_l |y_l|;   LM: y_0 + sum_{l>=1} |y_l|;   L2, Ld: sum_c sum_l |t_{c,l}|.
// |v| + acc is one v_sad_u32 on the value biased by 2^31.
//
// MODE_MAX records the best value of every segment and the global maximum;
// MODE_FIRST re-walks segments and returns the smallest rank whose value
// equals a target (the reference's strict '>' plus first-worker argmax,
// L_OpenMP.c:79,291-294, select exactly that rank).
#include "lnorm_walks.h"

namespace lnorm {


// ---------------------------------------------------------------------------
// single-problem kernels: the walks over the launch's own waves
// ---------------------------------------------------------------------------
template <int KIND, int NY, int MODE>
__global__ __launch_bounds__(256) void k_plain_bin(
    const int32_t *__restrict__ A, const int32_t *__restrict__ Dt, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst, uint32_t B) {
  walk_bin<KIND, NY, MODE>(as_const(A), as_const(Dt), R, q, seg0, nseg, target, unit_best, gbest, gfirst, B, global_wave());
}

template <int D, int NY, int MODE>
__global__ __launch_bounds__(256) void k_plain_dary(
    const int32_t *__restrict__ A, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst,
    uint32_t B) {
  walk_dary<D, NY, MODE>(as_const(A), R, q, seg0, nseg, target, unit_best, gbest, gfirst, B, global_wave());
}

template <int D, int NW, int MODE>
__global__ __launch_bounds__(256) void k_plain16_dary(
    const uint32_t *__restrict__ Ap, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst,
    uint32_t B) {
  walk_p16<D, NW, MODE>(as_const(Ap), R, q, seg0, nseg, target, unit_best, gbest, gfirst, B, global_wave());
}

static const int kPkWords[] = {4, 8, 12, 16, 20, 24, 28, 32, 40, 48, 64};

int plain16_words(int d, int ny) {
  const int need = (ny + 1) / 2;
  const int mx = dary16_max_words(d);
  for (int w : kPkWords) if (w >= need && w <= mx) return w;
  return 0;
}

template <int D, int NW>
static hipError_t launch_p16_nw(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const int wpb = block / 64;
  const dim3 grid((unsigned)((waves + wpb - 1) / wpb)), blk(block);
  const uint32_t *Ap = (const uint32_t *)a.A;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_plain16_dary<D, NW, MODE_MAX>), grid, blk, 0, st, Ap, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst, kBias16x2);
  else
    hipLaunchKernelGGL((k_plain16_dary<D, NW, MODE_FIRST>), grid, blk, 0, st, Ap, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst, kBias16x2);
  return hipGetLastError();
}

template <int D>
static hipError_t launch_p16(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.ny_pad / 2) {
#define X(n) case n: if constexpr (n <= dary16_max_words(D)) return launch_p16_nw<D, n>(mode, a, st, block); break;
    X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(40) X(48) X(64)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_plain16(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  if (!a.A || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (block <= 0 || block % 64 || block > 256) block = 256;
  if (!plain16_words(a.d, a.ny_pad) || plain16_words(a.d, a.ny_pad) * 2 != a.ny_pad) return hipErrorInvalidValue;
  switch (a.d) {
    case 2: return launch_p16<2>(mode, a, st, block);
    case 3: return launch_p16<3>(mode, a, st, block);
    case 4: return launch_p16<4>(mode, a, st, block);
    case 5: return launch_p16<5>(mode, a, st, block);
    case 6: return launch_p16<6>(mode, a, st, block);
    case 7: return launch_p16<7>(mode, a, st, block);
    case 8: return launch_p16<8>(mode, a, st, block);
    case 9: return launch_p16<9>(mode, a, st, block);
    case 10: return launch_p16<10>(mode, a, st, block);
    default: break;
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// instantiation tables and launchers
// ---------------------------------------------------------------------------
#define LN_BIN_WIDTHS(X) X(8) X(16) X(24) X(32) X(40) X(48) X(56) X(64) X(80) X(96) X(112) X(128)
#define LN_L2_WIDTHS(X) X(8) X(16) X(24) X(32) X(40) X(48) X(56) X(64)

static const int kBinWidths[] = {8, 16, 24, 32, 40, 48, 56, 64, 80, 96, 112, 128};
static const int kL2Widths[] = {8, 16, 24, 32, 40, 48, 56, 64};

// widths of the d-ary kernels (d * NY VGPRs of state)
static int dary_max_width(int d) {
  switch (d) {
    case 3: return 40;
    case 4: return 32;
    case 5: return 24;
    case 6: return 24;
    case 7: return 16;
    case 8: return 16;
    case 9: return 16;
    case 10: return 16;
    default: return 0;
  }
}

int plain_ny_pad(Kind kind, int d, int ny) {
  if (kind == KIND_L1 || kind == KIND_LM) {
    for (int w : kBinWidths) if (w >= ny) return w;
    return 0;
  }
  if (kind == KIND_L2) {
    for (int w : kL2Widths) if (w >= ny) return w;
    return 0;
  }
  const int mx = dary_max_width(d);
  for (int w = 8; w <= mx; w += 8) if (w >= ny) return w;
  return 0;
}

bool plain_supported(Kind kind, int d, int ny_pad) {
  return ny_pad > 0 && plain_ny_pad(kind, d, ny_pad) == ny_pad;
}

template <int KIND, int NY>
static hipError_t launch_bin_ny(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const int wpb = block / 64;
  const dim3 grid((unsigned)((waves + wpb - 1) / wpb)), blk(block);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_plain_bin<KIND, NY, MODE_MAX>), grid, blk, 0, st, a.A, a.D, a.R, a.q,
                       a.seg0, a.nseg, a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
  else
    hipLaunchKernelGGL((k_plain_bin<KIND, NY, MODE_FIRST>), grid, blk, 0, st, a.A, a.D, a.R, a.q,
                       a.seg0, a.nseg, a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
  return hipGetLastError();
}

template <int D, int NY>
static hipError_t launch_dary_ny(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const int wpb = block / 64;
  const dim3 grid((unsigned)((waves + wpb - 1) / wpb)), blk(block);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_plain_dary<D, NY, MODE_MAX>), grid, blk, 0, st, a.A, a.R, a.q, a.seg0,
                       a.nseg, a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
  else
    hipLaunchKernelGGL((k_plain_dary<D, NY, MODE_FIRST>), grid, blk, 0, st, a.A, a.R, a.q, a.seg0,
                       a.nseg, a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
  return hipGetLastError();
}

template <int D>
static hipError_t launch_dary(Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  switch (a.ny_pad) {
    case 8: return launch_dary_ny<D, 8>(mode, a, st, block);
    case 16: return launch_dary_ny<D, 16>(mode, a, st, block);
    case 24:
      if constexpr (D <= 6) return launch_dary_ny<D, 24>(mode, a, st, block);
      break;
    case 32:
      if constexpr (D <= 4) return launch_dary_ny<D, 32>(mode, a, st, block);
      break;
    case 40:
      if constexpr (D <= 3) return launch_dary_ny<D, 40>(mode, a, st, block);
      break;
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_plain(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st, int block) {
  if (!a.A || (kind != KIND_LD && !a.D) || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  // the kernels are compiled for at most 256 threads (__launch_bounds__)
  if (block <= 0 || block % 64 || block > 256) block = 256;
  if (!plain_supported(kind, a.d, a.ny_pad)) return hipErrorInvalidValue;
  if (kind == KIND_L1) {
    switch (a.ny_pad) {
#define X(n) case n: return launch_bin_ny<KIND_L1, n>(mode, a, st, block);
      LN_BIN_WIDTHS(X)
#undef X
    }
  } else if (kind == KIND_LM) {
    switch (a.ny_pad) {
#define X(n) case n: return launch_bin_ny<KIND_LM, n>(mode, a, st, block);
      LN_BIN_WIDTHS(X)
#undef X
    }
  } else if (kind == KIND_L2) {
    switch (a.ny_pad) {
#define X(n) case n: return launch_bin_ny<KIND_L2, n>(mode, a, st, block);
      LN_L2_WIDTHS(X)
#undef X
    }
  } else {
    switch (a.d) {
      case 3: return launch_dary<3>(mode, a, st, block);
      case 4: return launch_dary<4>(mode, a, st, block);
      case 5: return launch_dary<5>(mode, a, st, block);
      case 6: return launch_dary<6>(mode, a, st, block);
      case 7: return launch_dary<7>(mode, a, st, block);
      case 8: return launch_dary<8>(mode, a, st, block);
      case 9: return launch_dary<9>(mode, a, st, block);
      case 10: return launch_dary<10>(mode, a, st, block);
      default: break;
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace lnorm

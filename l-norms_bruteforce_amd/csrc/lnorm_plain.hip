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
#include "lnorm_device.h"

namespace lnorm {

// ---------------------------------------------------------------------------
// Binary family: L1, LM (one partial-sum vector y) and L2 (t0, t1).
// D holds [2][R-1][NY]: D[1][j] = +2A[j] (L1/LM) or +A[j] (L2), D[0][j] its
// negation; the new label of row j at rank N is 1 - bit_{j+1}(N).
// ---------------------------------------------------------------------------
template <int KIND, int NY, int MODE>
__global__ __launch_bounds__(256) void k_plain_bin(
    const int32_t *__restrict__ A, const int32_t *__restrict__ Dt, int R, int q,
    uint64_t seg0, uint64_t nseg, int32_t target, int32_t *__restrict__ unit_best,
    int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst, uint32_t B) {
  const int lane = wave_lane();
  const uint64_t w = global_wave();
  if (((w >> 1) << 7) >= nseg) return;  // the whole wave lies past the end
  const uint64_t local = lane_segment(w, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t par = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ w) & 1));

  constexpr bool TWO = (KIND == KIND_L2);
  uint32_t y[NY];
  uint32_t z[TWO ? NY : 1];

  // seed: the labels of rank N0 = seg * 2^q are the bits of N0 ^ (N0 >> 1)
  const uint64_t N0 = seg << q;
  const uint64_t g0 = N0 ^ (N0 >> 1);
#pragma unroll
  for (int l = 0; l < NY; ++l) y[l] = (uint32_t)A[(size_t)(R - 1) * NY + l] + kBias32;
  if constexpr (TWO) {
#pragma unroll
    for (int l = 0; l < NY; ++l) z[l] = kBias32;
  }
  for (int i = 0; i < R - 1; ++i) {
    const bool g = (g0 >> i) & 1;
    const int32_t *row = A + (size_t)i * NY;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const uint32_t a = (uint32_t)row[l];
      if constexpr (TWO) {
        z[l] += g ? a : 0u;
        y[l] += g ? 0u : a;
      } else {
        y[l] += g ? a : 0u - a;
      }
    }
  }

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  auto eval = [&](uint32_t t) {
    uint32_t acc;
    if constexpr (KIND == KIND_LM) {
      acc = y[0] ^ kBias32;
#pragma unroll
      for (int l = 1; l < NY; ++l) acc = sad32(y[l], B, acc);
    } else {
      acc = 0;
#pragma unroll
      for (int l = 0; l < NY; ++l) {
        acc = sad32(y[l], B, acc);
        if constexpr (TWO) acc = sad32(z[l], B, acc);
      }
    }
    if constexpr (MODE == MODE_MAX) {
      best = (int32_t)acc > best ? (int32_t)acc : best;
    } else {
      const uint32_t c = ((int32_t)acc == target) ? t : 0xFFFFFFFFu;
      first = c < first ? c : first;
    }
  };

  const uint32_t nsteps = 1u << q;
  const uint32_t nu = par << q;  // bit q of N: the parity of the segment index
  eval(0);
  for (uint32_t t = 1; t < nsteps; ++t) {
    const int j = __builtin_ctz(t);
    const uint32_t up = (((nu | t) >> (j + 1)) & 1u) ^ 1u;
    const int32_t *dp = Dt + (size_t)(up * (uint32_t)(R - 1) + (uint32_t)j) * NY;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const uint32_t dl = (uint32_t)dp[l];
      if constexpr (TWO) {
        z[l] += dl;
        y[l] -= dl;
      } else {
        y[l] += dl;
      }
    }
    eval(t);
  }

  if constexpr (MODE == MODE_MAX) {
    if (active) unit_best[local] = best;
    const int32_t m = wave_max_i32(active ? best : INT32_MIN);
    if (lane == 0) atomicMax(gbest, m);
  } else {
    const uint64_t r = (active && first != 0xFFFFFFFFu) ? ((seg << q) + first) : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (lane == 0 && m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
  }
}

// ---------------------------------------------------------------------------
// d-ary family (d = 3..10), reflected d-ary Gray code (L_OpenMP.c:126-239).
// The step counter t is kept in base d, one digit per nibble biased by
// 16 - d, so the changed digit j is the number of trailing 0xF nibbles and
// the carry of t + 1 rolls the lower digits over.  The parity of
// floor(N / d^(j+1)) decides the direction of digit j; it flips for every
// digit below the changed one.  Only the two label vectors that exchange
// row j are updated and re-summed.
// ---------------------------------------------------------------------------
template <int F, int T, int D, int NY>
__device__ __forceinline__ void apply_move(uint32_t (&t)[D][NY], uint32_t (&S)[D],
                                           const int32_t *__restrict__ row, uint32_t B) {
  uint32_t sf = 0, st = 0;
#pragma unroll
  for (int l = 0; l < NY; ++l) {
    const uint32_t a = (uint32_t)row[l];
    t[F][l] -= a;
    t[T][l] += a;
    sf = sad32(t[F][l], B, sf);
    st = sad32(t[T][l], B, st);
  }
  S[F] = sf;
  S[T] = st;
}

template <int C, int D, int NY>
struct Mover {
  __device__ __forceinline__ static void run(uint32_t (&t)[D][NY], uint32_t (&S)[D],
                                             const int32_t *__restrict__ row, int from, int up,
                                             uint32_t B) {
    if (from == C) {
      if (up) {
        if constexpr (C + 1 < D) apply_move<C, C + 1, D, NY>(t, S, row, B);
      } else {
        if constexpr (C > 0) apply_move<C, C - 1, D, NY>(t, S, row, B);
      }
    } else {
      if constexpr (C + 1 < D) Mover<C + 1, D, NY>::run(t, S, row, from, up, B);
    }
  }
};

template <int D, int NY, int MODE>
__global__ __launch_bounds__(256) void k_plain_dary(
    const int32_t *__restrict__ A, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst, uint32_t B) {
  const int lane = wave_lane();
  const uint64_t w = global_wave();
  if (((w >> 1) << 7) >= nseg) return;
  const uint64_t local = lane_segment(w, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ w) & 1));

  // labels of N0 = seg * d^q: digits below q are 0 and sit at an end of
  // their range (d-1 when floor(N0 / d^(i+1)) is odd); digits from q on are
  // the reflected code of seg.
  uint32_t par_mask;   // bit i: parity of floor(N / d^(i+1)), i < q
  if (D & 1) par_mask = pseg ? ((1u << q) - 1u) : 0u;
  else par_mask = (q > 0 && pseg) ? (1u << (q - 1)) : 0u;

  uint32_t t[D][NY];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int l = 0; l < NY; ++l)
      t[c][l] = (c == 0 ? (uint32_t)A[(size_t)(R - 1) * NY + l] : 0u) + kBias32;
  uint64_t x = seg;
  for (int i = 0; i < R - 1; ++i) {
    int lab;
    if (i < q) {
      lab = ((par_mask >> i) & 1) ? D - 1 : 0;
    } else {
      const uint32_t r = (uint32_t)(x % (uint64_t)(2 * D));
      lab = r < (uint32_t)D ? (int)r : (int)(2 * D - 1 - r);
      x /= (uint64_t)D;
    }
    const int32_t *row = A + (size_t)i * NY;
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
      for (int l = 0; l < NY; ++l) t[c][l] += (lab == c) ? (uint32_t)row[l] : 0u;
  }
  uint32_t S[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    uint32_t s = 0;
#pragma unroll
    for (int l = 0; l < NY; ++l) s = sad32(t[c][l], B, s);
    S[c] = s;
  }

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  auto eval = [&](uint32_t tt) {
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) acc += S[c];
    if constexpr (MODE == MODE_MAX) {
      best = (int32_t)acc > best ? (int32_t)acc : best;
    } else {
      const uint32_t cnd = ((int32_t)acc == target) ? tt : 0xFFFFFFFFu;
      first = cnd < first ? cnd : first;
    }
  };

  // nibble counter: digit value v stored as v + (16 - d)
  uint64_t biaspat = 0;
  for (int i = 0; i < q; ++i) biaspat |= (uint64_t)(16 - D) << (4 * i);
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;
  uint32_t nsteps = 1;
  for (int i = 0; i < q; ++i) nsteps *= (uint32_t)D;

  eval(0);
  for (uint32_t tt = 1; tt < nsteps; ++tt) {
    const int j = __builtin_ctzll(~cnt) >> 2;
    const uint64_t low = (1ull << (4 * j)) - 1ull;
    cnt = (cnt + 1ull) | (biaspat & low);
    const int v = (int)((cnt >> (4 * j)) & 15ull) - (16 - D);
    const int p = (pm >> j) & 1;
    const int from = p ? D - v : v - 1;
    pm ^= (1u << j) - 1u;
    Mover<0, D, NY>::run(t, S, A + (size_t)j * NY, from, p ? 0 : 1, B);
    eval(tt);
  }

  if constexpr (MODE == MODE_MAX) {
    if (active) unit_best[local] = best;
    const int32_t m = wave_max_i32(active ? best : INT32_MIN);
    if (lane == 0) atomicMax(gbest, m);
  } else {
    const uint64_t r = (active && first != 0xFFFFFFFFu) ? (seg * (uint64_t)nsteps + first) : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (lane == 0 && m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
  }
}

// ---------------------------------------------------------------------------
// Packed 16-bit d-ary family (d = 2..10): two columns per VGPR, each half
// biased by 2^15.  A row move is one v_pk_sub_u16 and one v_pk_add_u16 per
// two columns, a re-sum one v_sad_u16 per two columns: half the VALU work of
// the 32-bit walk.  Exact whenever every column's absolute sum is <= 2^15 - 1
// (the host checks), since each label sum t_c,l is bounded by it.  d = 2 is
// L_2: the reflected binary code P = 0,1,1,0 is the BRGC (L_OpenMP.c:101).
// ---------------------------------------------------------------------------
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b));
}

template <int F, int T, int D, int NW>
__device__ __forceinline__ void apply_move16(uint32_t (&t)[D][NW], uint32_t (&S)[D],
                                             const uint32_t *__restrict__ row, uint32_t B) {
  uint32_t sf = 0, st = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint32_t a = row[w];
    t[F][w] = pk_sub(t[F][w], a);
    t[T][w] = pk_add(t[T][w], a);
    sf = sad16(t[F][w], B, sf);
    st = sad16(t[T][w], B, st);
  }
  S[F] = sf;
  S[T] = st;
}

template <int C, int D, int NW>
struct Mover16 {
  __device__ __forceinline__ static void run(uint32_t (&t)[D][NW], uint32_t (&S)[D],
                                             const uint32_t *__restrict__ row, int from, int up,
                                             uint32_t B) {
    if (from == C) {
      if (up) {
        if constexpr (C + 1 < D) apply_move16<C, C + 1, D, NW>(t, S, row, B);
      } else {
        if constexpr (C > 0) apply_move16<C, C - 1, D, NW>(t, S, row, B);
      }
    } else {
      if constexpr (C + 1 < D) Mover16<C + 1, D, NW>::run(t, S, row, from, up, B);
    }
  }
};

// Ap: R x NW packed rows, word w = (A[i][2w+1] & 0xFFFF) << 16 | (A[i][2w] & 0xFFFF).
template <int D, int NW, int MODE>
__global__ __launch_bounds__(256) void k_plain16_dary(
    const uint32_t *__restrict__ Ap, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst, uint32_t B) {
  const int lane = wave_lane();
  const uint64_t w = global_wave();
  if (((w >> 1) << 7) >= nseg) return;
  const uint64_t local = lane_segment(w, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ w) & 1));

  uint32_t par_mask;
  if (D & 1) par_mask = pseg ? ((1u << q) - 1u) : 0u;
  else par_mask = (q > 0 && pseg) ? (1u << (q - 1)) : 0u;

  uint32_t t[D][NW];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int l = 0; l < NW; ++l) t[c][l] = pk_add(kBias16x2, c == 0 ? Ap[(size_t)(R - 1) * NW + l] : 0u);
  uint64_t x = seg;
  for (int i = 0; i < R - 1; ++i) {
    int lab;
    if (i < q) {
      lab = ((par_mask >> i) & 1) ? D - 1 : 0;
    } else {
      const uint32_t r = (uint32_t)(x % (uint64_t)(2 * D));
      lab = r < (uint32_t)D ? (int)r : (int)(2 * D - 1 - r);
      x /= (uint64_t)D;
    }
    const uint32_t *row = Ap + (size_t)i * NW;
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
      for (int l = 0; l < NW; ++l) t[c][l] = pk_add(t[c][l], (lab == c) ? row[l] : 0u);
  }
  uint32_t S[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    uint32_t s = 0;
#pragma unroll
    for (int l = 0; l < NW; ++l) s = sad16(t[c][l], B, s);
    S[c] = s;
  }

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  auto eval = [&](uint32_t tt) {
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) acc += S[c];
    if constexpr (MODE == MODE_MAX) {
      best = (int32_t)acc > best ? (int32_t)acc : best;
    } else {
      const uint32_t cnd = ((int32_t)acc == target) ? tt : 0xFFFFFFFFu;
      first = cnd < first ? cnd : first;
    }
  };

  uint64_t biaspat = 0;
  for (int i = 0; i < q; ++i) biaspat |= (uint64_t)(16 - D) << (4 * i);
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;
  uint32_t nsteps = 1;
  for (int i = 0; i < q; ++i) nsteps *= (uint32_t)D;

  eval(0);
  for (uint32_t tt = 1; tt < nsteps; ++tt) {
    const int j = __builtin_ctzll(~cnt) >> 2;
    const uint64_t low = (1ull << (4 * j)) - 1ull;
    cnt = (cnt + 1ull) | (biaspat & low);
    const int v = (int)((cnt >> (4 * j)) & 15ull) - (16 - D);
    const int p = (pm >> j) & 1;
    const int from = p ? D - v : v - 1;
    pm ^= (1u << j) - 1u;
    Mover16<0, D, NW>::run(t, S, Ap + (size_t)j * NW, from, p ? 0 : 1, B);
    eval(tt);
  }

  if constexpr (MODE == MODE_MAX) {
    if (active) unit_best[local] = best;
    const int32_t m = wave_max_i32(active ? best : INT32_MIN);
    if (lane == 0) atomicMax(gbest, m);
  } else {
    const uint64_t r = (active && first != 0xFFFFFFFFu) ? (seg * (uint64_t)nsteps + first) : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (lane == 0 && m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
  }
}

// largest packed width instantiated per d (d * NW <= 160 words of state)
static constexpr int dary16_max_words(int d) {
  return d == 2 ? 64 : d == 3 ? 48 : (d == 4 || d == 5) ? 32 : d == 6 ? 24 : (d >= 7 && d <= 10) ? 16 : 0;
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

// lnorm_colwalk.hip -- the column walk: one wave per segment, the columns
// over the lanes, every label sum in VGPRs.  For widths no register-resident
// lane-per-segment walk holds (up to the reference's `length` = 4096 columns
// for L1/LM and L2, 2048 at d = 3, 4 ... 768 at d = 9, 10), with any int32
// entries (32-bit words biased by 2^31, v_sad_u32).
//
// Lane L owns columns k * 64 + L (k < CPL); label c's sum of column k sits
// in t[c][k].  Every Gray step is wave-uniform (the segment is the wave's):
// the changed row and its direction come from a scalar step counter, as in
// the lane-per-segment walks (L1/LM: the binary reflected code; d >= 2: the
// reflected d-ary code in a nibble counter, d - 1 moves of row 0 unrolled
// per block of d steps).  The rows sit in LDS (the workgroup copies the
// matrix once), the ones most steps move also in registers (row 0 for
// d >= 2, which moves in (d - 1) / d of the steps; rows 0..H-1 for L1/LM):
// a step takes the changed row's CPL words per lane (from registers, or
// ds_read_b32, lane-consecutive, no bank conflicts).  d >= 2: it moves them
// between the two labels that exchange the row and re-sums those two labels'
// |.| over the lane's columns; L1/LM: it adds 2 A or -2 A to the one sum and
// re-sums its |.|.  Six DPP adds total the lanes' partial objectives.
//   Ops per step and lane: CPL x (2 moves + 2 |.|) for d >= 2, CPL x 2 for
//   L1/LM, plus the wave total (6 DPP adds) and the running maximum.
#include "lnorm_colwalk.h"
#include "lnorm_walks.h"

namespace lnorm {

// columns per lane of the instantiated widths (64 * CPL columns)
constexpr int kColCpl[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64};
// waves (segments) per workgroup, which share one copy of the rows in LDS:
// 8 for d >= 2 (up to 209 VGPRs: 2 waves per SIMD each), 16 for L1/LM (at
// most 128 VGPRs), whose wide rows would otherwise leave 2 waves per SIMD
// on a CU to hide the latency of each step's wave total
constexpr int kColWavesDary = 8, kColWavesBin = 16;

// the widest label state a lane holds: D x CPL words (binary: CPL)
constexpr bool col_fits(int D, int CPL) { return (D <= 1 ? CPL : D * CPL) <= 128; }

// The sum over the 64 lanes by DPP: pairs and quads (quad_perm), 8 and 16
// lanes (row_half_mirror, row_mirror), then rows 1 and 3 add the last lane
// of rows 0 and 2 (row_bcast:15) and rows 2 and 3 that of row 1
// (row_bcast:31); lane 63 ends with the total.  Six VALU adds with a DPP
// source and one readlane, where the shuffle butterfly takes six
// ds_bpermute round trips.
__device__ __forceinline__ uint32_t wave_total(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);   // row_half_mirror
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);   // row_mirror
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);   // row_bcast:15, rows 1, 3
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);   // row_bcast:31, rows 2, 3
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// A row's words of the lane's columns: in LDS (LdsRow) or, for the rows most
// steps move, held in registers (RegRow).
struct LdsRow {
  const uint32_t *p;   // the row's word of column lane, k * 64 apart
  __device__ __forceinline__ uint32_t operator()(int k) const { return p[k * 64]; }
};
template <int CPL>
struct RegRow {
  const uint32_t (&r)[CPL];
  __device__ __forceinline__ uint32_t operator()(int k) const { return r[k]; }
};

// ---- d >= 2: the label sums t[c][k] -----------------------------------------
template <int F, int T, int D, int CPL, class Row>
__device__ __forceinline__ void col_move(uint32_t (&t)[D][CPL], uint32_t (&S)[D], const Row &row, uint32_t B,
                                         uint32_t z) {
  uint32_t sf = z, st = z;
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const uint32_t a = row(k);
    t[F][k] -= a;
    t[T][k] += a;
    sf = sad32(t[F][k], B, sf);
    st = sad32(t[T][k], B, st);
  }
  S[F] = sf;
  S[T] = st;
}

template <int C, int D, int CPL>
struct ColMover {
  __device__ __forceinline__ static void run(uint32_t (&t)[D][CPL], uint32_t (&S)[D], const LdsRow &row,
                                             int from, int up, uint32_t B, uint32_t z) {
    if (from == C) {
      if (up) {
        if constexpr (C + 1 < D) col_move<C, C + 1, D, CPL>(t, S, row, B, z);
      } else {
        if constexpr (C > 0) col_move<C, C - 1, D, CPL>(t, S, row, B, z);
      }
    } else {
      if constexpr (C + 1 < D) ColMover<C + 1, D, CPL>::run(t, S, row, from, up, B, z);
    }
  }
};

// the d - 1 moves of row 0 inside one block of d steps (row 0 moves in
// (d - 1) / d of all steps: its words are held in registers where the state
// leaves room)
template <int K, int D, int CPL, bool UP>
struct ColDigit0Run {
  template <class Row, class Eval>
  __device__ __forceinline__ static void run(uint32_t (&t)[D][CPL], uint32_t (&S)[D], const Row &row0,
                                             uint32_t B, uint32_t z, uint32_t &tt, Eval &eval) {
    if constexpr (K < D) {
      constexpr int F = UP ? K - 1 : D - K;
      constexpr int T = UP ? K : D - K - 1;
      col_move<F, T, D, CPL>(t, S, row0, B, z);
      eval(++tt);
      ColDigit0Run<K + 1, D, CPL, UP>::run(t, S, row0, B, z, tt, eval);
    }
  }
};

// seed: row `row` joins label `lab` (wave-uniform)
template <int C, int D, int CPL>
struct ColAdd {
  __device__ __forceinline__ static void run(uint32_t (&t)[D][CPL], const uint32_t *row, int lab) {
    if (lab == C) {
#pragma unroll
      for (int k = 0; k < CPL; ++k) t[C][k] += row[k * 64];
    } else {
      if constexpr (C + 1 < D) ColAdd<C + 1, D, CPL>::run(t, row, lab);
    }
  }
};

// Copies the R x Cp matrix (lane-consecutive words) into LDS.
__device__ __forceinline__ void col_stage(uint32_t *lds, const int32_t *__restrict__ src, int words) {
  for (int i = threadIdx.x; i < words; i += blockDim.x) lds[i] = (uint32_t)src[i];
  __syncthreads();
}

template <int D, int CPL, int MODE>
__global__ __launch_bounds__(64 * kColWavesDary) void k_colwalk_dary(
    const int32_t *__restrict__ A, int R, int Cp, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst,
    uint32_t B) {
  extern __shared__ uint32_t rows[];
  col_stage(rows, A, R * Cp);
  const int lane = wave_lane();
  const uint64_t w = global_wave();
  if (w >= nseg) return;                  // wave-uniform; no barrier follows
  const uint64_t seg = seg0 + w;
  const uint32_t z = opaque_zero();
  const uint32_t *mine = rows + lane;     // this lane's column k of row i: mine[i * Cp + k * 64]

  // labels of N0 = seg * d^q: digits below q sit at an end of their range
  // (d - 1 where floor(N0 / d^(i+1)) is odd); digits from q on are the
  // reflected code of seg
  const uint32_t pseg = (uint32_t)(seg & 1);
  uint32_t par_mask;
  if (D & 1) par_mask = pseg ? ((1u << q) - 1u) : 0u;
  else par_mask = (q > 0 && pseg) ? (1u << (q - 1)) : 0u;

  uint32_t t[D][CPL];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int k = 0; k < CPL; ++k) t[c][k] = (c == 0 ? mine[(R - 1) * Cp + k * 64] : 0u) + B;
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
    ColAdd<0, D, CPL>::run(t, mine + (size_t)i * Cp, lab);
  }
  // row 0 in registers up to 96 state words (beyond, the extra CPL
  // registers made the widest d = 8 .. 10 shapes spill)
  constexpr bool kRow0Regs = D * CPL <= 96;
  uint32_t r0[kRow0Regs ? CPL : 1];
#pragma unroll
  for (int k = 0; k < (kRow0Regs ? CPL : 1); ++k) r0[k] = mine[k * 64];
  const auto row0 = [&]() {
    if constexpr (kRow0Regs) return RegRow<CPL>{r0};
    else return LdsRow{mine};
  }();
  uint32_t S[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    uint32_t s = z;
#pragma unroll
    for (int k = 0; k < CPL; ++k) s = sad32(t[c][k], B, s);
    S[c] = s;
  }

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  auto eval = [&](uint32_t tt) {
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) acc += S[c];
    const int32_t v = (int32_t)wave_total(acc);
    if constexpr (MODE == MODE_MAX) {
      best = v > best ? v : best;
    } else {
      if (v == target && first == 0xFFFFFFFFu) first = tt;
    }
  };

  // blocks of d steps: step b * d changes a digit j >= 1 (nibble counter,
  // digit value v stored as v + 16 - d); the d - 1 moves of row 0 between
  // are unrolled
  uint64_t biaspat = 0;
  for (int i = 1; i < q; ++i) biaspat |= (uint64_t)(16 - D) << (4 * (i - 1));
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;
  uint32_t nblocks = 1;
  for (int i = 1; i < q; ++i) nblocks *= (uint32_t)D;

  eval(0);
  uint32_t tt = 0;
  if (q > 0) {
    if (pm & 1u) ColDigit0Run<1, D, CPL, false>::run(t, S, row0, B, z, tt, eval);
    else ColDigit0Run<1, D, CPL, true>::run(t, S, row0, B, z, tt, eval);
    for (uint32_t b = 1; b < nblocks; ++b) {
      const int jj = __builtin_ctzll(~cnt) >> 2;
      const uint64_t low = (1ull << (4 * jj)) - 1ull;
      cnt = (cnt + 1ull) | (biaspat & low);
      const int v = (int)((cnt >> (4 * jj)) & 15ull) - (16 - D);
      const int j = jj + 1;
      const int p = (pm >> j) & 1;
      const int from = p ? D - v : v - 1;
      pm ^= (1u << j) - 1u;
      ColMover<0, D, CPL>::run(t, S, LdsRow{mine + (size_t)j * Cp}, from, p ? 0 : 1, B, z);
      eval(++tt);
      if (pm & 1u) ColDigit0Run<1, D, CPL, false>::run(t, S, row0, B, z, tt, eval);
      else ColDigit0Run<1, D, CPL, true>::run(t, S, row0, B, z, tt, eval);
    }
  }
  const uint64_t nsteps = q > 0 ? (uint64_t)nblocks * (uint64_t)D : 1ull;
  if constexpr (MODE == MODE_MAX) {
    if (lane == 0) unit_best[w] = best;
    if (best > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, best);
  } else {
    if (first != 0xFFFFFFFFu) atomicMin(gfirst, (unsigned long long)(seg * nsteps + first));
  }
}

// ---- L1 / LM: one sum per column -----------------------------------------------
// rows: the 2 A rows of the |.| columns (LDS); A: the rows themselves (the
// seed); marg: LM's column 0 (scalar loads), which enters without |.|.
template <int CPL, int MARG, int MODE>
__global__ __launch_bounds__(64 * kColWavesBin) void k_colwalk_bin(
    const int32_t *__restrict__ A, const int32_t *__restrict__ A2, const int32_t *__restrict__ margp, int R,
    int Cp, int q, uint64_t seg0, uint64_t nseg, int32_t target, int32_t *__restrict__ unit_best,
    int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst, uint32_t B) {
  extern __shared__ uint32_t rows[];
  col_stage(rows, A2, R * Cp);
  const cint32_t *marg = as_const(margp);   // scalar loads: the row index is wave-uniform
  const int lane = wave_lane();
  const uint64_t w = global_wave();
  if (w >= nseg) return;
  const uint64_t seg = seg0 + w;
  const uint64_t N0 = seg << q;
  const uint64_t g0 = N0 ^ (N0 >> 1);
  const uint32_t z = opaque_zero();
  const uint32_t *mine = rows + lane;

  // seed: y = A[R-1] + sum_i s_i A[i], s_i = +1 where Gray bit i of N0 is set
  uint32_t y[CPL];
  const int32_t *a = A + lane;
#pragma unroll
  for (int k = 0; k < CPL; ++k) y[k] = (uint32_t)a[(size_t)(R - 1) * Cp + k * 64] + B;
  int32_t y0 = MARG ? marg[R - 1] : 0;
  for (int i = 0; i < R - 1; ++i) {
    const bool plus = (g0 >> i) & 1;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const uint32_t v = (uint32_t)a[(size_t)i * Cp + k * 64];
      y[k] += plus ? v : 0u - v;
    }
    if (MARG) y0 += plus ? marg[i] : -marg[i];
  }

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  auto eval = [&](uint32_t tt) {
    uint32_t acc = z;
#pragma unroll
    for (int k = 0; k < CPL; ++k) acc = sad32(y[k], B, acc);
    const int32_t v = (int32_t)wave_total(acc) + y0;
    if constexpr (MODE == MODE_MAX) {
      best = v > best ? v : best;
    } else {
      if (v == target && first == 0xFFFFFFFFu) first = tt;
    }
  };

  // rows 0 .. H-1 move in 1 - 2^-H of all steps: their 2 A words sit in
  // registers, the others are read from LDS (one word per column and step
  // would otherwise take the CU's whole LDS bandwidth at 2 VALU ops a word)
  constexpr int H = CPL <= 16 ? 3 : CPL <= 32 ? 2 : CPL <= 48 ? 1 : 0;
  uint32_t hr[H > 0 ? H : 1][CPL];
#pragma unroll
  for (int h = 0; h < H; ++h)
#pragma unroll
    for (int k = 0; k < CPL; ++k) hr[h][k] = h < R ? mine[(size_t)h * Cp + k * 64] : 0u;

  eval(0);
  const uint32_t nsteps = 1u << q;
  for (uint32_t tt = 1; tt < nsteps; ++tt) {
    const int j = __builtin_ctz(tt);
    const bool up = !(((N0 + tt) >> (j + 1)) & 1);   // Gray bit j becomes 1: s_j from -1 to +1
    // `up` and j are wave-uniform: branches, not per-column selects (which
    // cost a negate and a v_cndmask besides the add)
    if (j < H) {
#pragma unroll
      for (int h = 0; h < H; ++h) {
        if (j == h) {
          if (up) {
#pragma unroll
            for (int k = 0; k < CPL; ++k) y[k] += hr[h][k];
          } else {
#pragma unroll
            for (int k = 0; k < CPL; ++k) y[k] -= hr[h][k];
          }
        }
      }
    } else {
      const uint32_t *row = mine + (size_t)j * Cp;
      if (up) {
#pragma unroll
        for (int k = 0; k < CPL; ++k) y[k] += row[k * 64];
      } else {
#pragma unroll
        for (int k = 0; k < CPL; ++k) y[k] -= row[k * 64];
      }
    }
    if (MARG) y0 += up ? 2 * marg[j] : -2 * marg[j];
    eval(tt);
  }
  if constexpr (MODE == MODE_MAX) {
    if (lane == 0) unit_best[w] = best;
    if (best > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, best);
  } else {
    if (first != 0xFFFFFFFFu) atomicMin(gfirst, (unsigned long long)(N0 + first));
  }
}

// ---- host side -------------------------------------------------------------------
int colwalk_cpl(int d, int ncols) {
  for (int c : kColCpl)
    if (64 * c >= ncols) return col_fits(d, c) ? c : 0;
  return 0;
}

size_t colwalk_lds_bytes(int R, int Cp) { return (size_t)R * Cp * 4; }

// Launches one wave per segment, W per workgroup, with the rows' LDS; above
// the default 64 KiB the kernel is allowed the larger size first.
template <int W, class K, class... Args>
static hipError_t lcol_launch(K kernel, uint64_t nseg, size_t lds, hipStream_t st, Args... args) {
  if (lds > (64u << 10)) {
    const hipError_t e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
    if (e != hipSuccess) return e;
  }
  const dim3 grid((unsigned)((nseg + W - 1) / W)), blk(64 * W);
  hipLaunchKernelGGL(kernel, grid, blk, lds, st, args...);
  return hipGetLastError();
}

template <int D, int CPL>
static hipError_t lcol_dary(Mode mode, const PlainArgs &a, hipStream_t st) {
  const size_t lds = colwalk_lds_bytes(a.R, a.ny_pad);
  if (mode == MODE_MAX)
    return lcol_launch<kColWavesDary>(k_colwalk_dary<D, CPL, MODE_MAX>, a.nseg, lds, st, a.A, a.R, a.ny_pad, a.q,
                                      a.seg0, a.nseg, a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
  return lcol_launch<kColWavesDary>(k_colwalk_dary<D, CPL, MODE_FIRST>, a.nseg, lds, st, a.A, a.R, a.ny_pad, a.q,
                                    a.seg0, a.nseg, a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
}

template <int CPL, int MARG>
static hipError_t lcol_bin(Mode mode, const PlainArgs &a, hipStream_t st) {
  const size_t lds = colwalk_lds_bytes(a.R, a.ny_pad);
  // the 2 A rows follow the rows in D; LM's column 0 follows them
  const int32_t *A2 = a.D;
  const int32_t *marg = a.D + (size_t)a.R * a.ny_pad;
  if (mode == MODE_MAX)
    return lcol_launch<kColWavesBin>(k_colwalk_bin<CPL, MARG, MODE_MAX>, a.nseg, lds, st, a.A, A2, marg, a.R,
                                     a.ny_pad, a.q, a.seg0, a.nseg, a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
  return lcol_launch<kColWavesBin>(k_colwalk_bin<CPL, MARG, MODE_FIRST>, a.nseg, lds, st, a.A, A2, marg, a.R,
                                   a.ny_pad, a.q, a.seg0, a.nseg, a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
}

template <int D>
static hipError_t lcol_d(int cpl, Mode mode, const PlainArgs &a, hipStream_t st) {
  switch (cpl) {
#define X(c) \
  case c:    \
    if constexpr (col_fits(D, c)) return lcol_dary<D, c>(mode, a, st); \
    break;
    X(1) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

template <int MARG>
static hipError_t lcol_b(int cpl, Mode mode, const PlainArgs &a, hipStream_t st) {
  switch (cpl) {
#define X(c) \
  case c: return lcol_bin<c, MARG>(mode, a, st);
    X(1) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_colwalk(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (a.ny_pad % 64 != 0 || a.R < 1) return hipErrorInvalidValue;
  const int cpl = a.ny_pad / 64;
  const int d = (kind == KIND_L1 || kind == KIND_LM) ? 1 : a.d;
  if (colwalk_cpl(d, a.ny_pad) != cpl) return hipErrorInvalidValue;
  if (colwalk_lds_bytes(a.R, a.ny_pad) > kColwalkLdsMax) return hipErrorInvalidValue;
  if (kind == KIND_L1) return lcol_b<0>(cpl, mode, a, st);
  if (kind == KIND_LM) return lcol_b<1>(cpl, mode, a, st);
  switch (a.d) {
    case 2: return lcol_d<2>(cpl, mode, a, st);
    case 3: return lcol_d<3>(cpl, mode, a, st);
    case 4: return lcol_d<4>(cpl, mode, a, st);
    case 5: return lcol_d<5>(cpl, mode, a, st);
    case 6: return lcol_d<6>(cpl, mode, a, st);
    case 7: return lcol_d<7>(cpl, mode, a, st);
    case 8: return lcol_d<8>(cpl, mode, a, st);
    case 9: return lcol_d<9>(cpl, mode, a, st);
    case 10: return lcol_d<10>(cpl, mode, a, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace lnorm

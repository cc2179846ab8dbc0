// lnorm_walks.h -- the Gray-code walks as device functions.
//
// Each walk_* runs one wave's share of a segment walk: wave index w of a
// launch (or of one item of a batch) owns the 64 segments lane_segment(w, .)
// and writes unit_best / gbest (MODE_MAX) or gfirst (MODE_FIRST).  The
// single-problem kernels (lnorm_plain.hip) call them with global_wave(); the
// persistent batch kernels (lnorm_batch.hip) with the wave index of a task.
#pragma once
#include "lnorm_device.h"

#include <utility>

namespace lnorm {

// Matrix and delta rows are read through the constant address space: every
// such load in a walk has a wave-uniform address, and constant-space loads
// become scalar loads (SGPR operands) even where the compiler cannot prove
// the memory unwritten (the persistent batch kernels, which store results in
// the same loop).  All device allocations are addressable this way.
//
// The walks take no branch on the lane index (lane 0 only): inside the
// persistent batch loop such a loop-invariant, lane-divergent condition has
// been seen to let the compiler split the loop by lane, so that lanes 1..63
// re-entered the loop without the task fetch and never left it.  The final
// atomics run on every lane with the wave-reduced value instead.
typedef __attribute__((address_space(4))) const int32_t cint32_t;
typedef __attribute__((address_space(4))) const uint32_t cuint32_t;
__device__ __forceinline__ const cint32_t *as_const(const int32_t *p) { return (const cint32_t *)(uintptr_t)p; }
__device__ __forceinline__ const cuint32_t *as_const(const uint32_t *p) { return (const cuint32_t *)(uintptr_t)p; }
// ---------------------------------------------------------------------------
// Binary family: L1, LM (one partial-sum vector y) and L2 (t0, t1).
// D holds [2][R-1][NY]: D[1][j] = +2A[j] (L1/LM) or +A[j] (L2), D[0][j] its
// negation; the new label of row j at rank N is 1 - bit_{j+1}(N).
// ---------------------------------------------------------------------------
template <int KIND, int NY, int MODE>
__device__ __forceinline__ void walk_bin(
    const cint32_t *A, const cint32_t *Dt, int R, int q,
    uint64_t seg0, uint64_t nseg, int32_t target, int32_t *__restrict__ unit_best,
    int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst, uint32_t B,
    uint64_t w) {
  const int lane = wave_lane();
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
    const cint32_t *row = A + (size_t)i * NY;
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
  const uint32_t zero = opaque_zero();
  auto eval = [&](uint32_t t) {
    uint32_t acc;
    if constexpr (KIND == KIND_LM) {
      acc = y[0] ^ kBias32;
#pragma unroll
      for (int l = 1; l < NY; ++l) acc = sad32(y[l], B, acc);
    } else {
      acc = zero;
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
    const cint32_t *dp = Dt + (size_t)(up * (uint32_t)(R - 1) + (uint32_t)j) * NY;
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
    // gbest only grows, so any value read from it is a lower bound of the
    // current one: skip the (device-scope, serialising) atomic when it
    // cannot raise the maximum.  The value read is made wave-uniform
    // (readfirstlane) so the branch is uniform: a lane-divergent branch here
    // made the compiler structurise the whole walk loop again, with one
    // register copy per state word and step.
    if (m > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, m);
  } else {
    const uint64_t r = (active && first != 0xFFFFFFFFu) ? ((seg << q) + first) : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
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
                                           const cint32_t *row, uint32_t B) {
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
                                             const cint32_t *row, int from, int up,
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

// The d-1 moves of row 0 inside one block of d steps (32-bit label vectors).
template <int K, int D, int NY, bool UP>
struct Digit0Run32 {
  template <class Eval>
  __device__ __forceinline__ static void run(uint32_t (&t)[D][NY], uint32_t (&S)[D], const cint32_t *row0,
                                             uint32_t B, uint32_t &tt, Eval &eval) {
    if constexpr (K < D) {
      constexpr int F = UP ? K - 1 : D - K;
      constexpr int T = UP ? K : D - K - 1;
      apply_move<F, T, D, NY>(t, S, row0, B);
      eval(++tt);
      Digit0Run32<K + 1, D, NY, UP>::run(t, S, row0, B, tt, eval);
    }
  }
};

// Seeding a lane-per-segment d-ary walk: row `row` joins label `lab`, which
// differs per lane.  Each state word adds the scalar row word ANDed with a
// per-label lane mask.  Written as `(lab == c) ? row : 0` per state word (a
// lane-varying select), the compiler held a second copy of the state, about
// twice the registers, and spilled the wider shapes.
template <int D, int NY>
__device__ __forceinline__ void seed_add32(uint32_t (&t)[D][NY], const cint32_t *row, int lab) {
  uint32_t mask[D];
#pragma unroll
  for (int c = 0; c < D; ++c) mask[c] = (lab == c) ? 0xFFFFFFFFu : 0u;
#pragma unroll
  for (int l = 0; l < NY; ++l) {
    const uint32_t v = (uint32_t)row[l];
#pragma unroll
    for (int c = 0; c < D; ++c) t[c][l] += v & mask[c];
  }
}

template <int D, int NY, int MODE>
__device__ __forceinline__ void walk_dary(
    const cint32_t *A, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst, uint32_t B,
    uint64_t w) {
  const int lane = wave_lane();
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
    seed_add32<D, NY>(t, A + (size_t)i * NY, lab);
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

  // Blocks of d steps, as in walk_p16: step b*d changes a digit j >= 1 (the
  // dispatched move); the d-1 moves of row 0 in between are unrolled.  The
  // nibble counter (digit value v stored as v + (16 - d)) covers digits
  // 1..q-1.
  uint64_t biaspat = 0;
  for (int i = 1; i < q; ++i) biaspat |= (uint64_t)(16 - D) << (4 * (i - 1));
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;
  uint32_t nblocks = 1;
  for (int i = 1; i < q; ++i) nblocks *= (uint32_t)D;

  eval(0);
  uint32_t tt = 0;
  if (q > 0) {
    if (pm & 1u) Digit0Run32<1, D, NY, false>::run(t, S, A, B, tt, eval);
    else Digit0Run32<1, D, NY, true>::run(t, S, A, B, tt, eval);
    for (uint32_t b = 1; b < nblocks; ++b) {
      const int jj = __builtin_ctzll(~cnt) >> 2;
      const uint64_t low = (1ull << (4 * jj)) - 1ull;
      cnt = (cnt + 1ull) | (biaspat & low);
      const int v = (int)((cnt >> (4 * jj)) & 15ull) - (16 - D);
      const int j = jj + 1;
      const int p = (pm >> j) & 1;
      const int from = p ? D - v : v - 1;
      pm ^= (1u << j) - 1u;
      Mover<0, D, NY>::run(t, S, A + (size_t)j * NY, from, p ? 0 : 1, B);
      eval(++tt);
      if (pm & 1u) Digit0Run32<1, D, NY, false>::run(t, S, A, B, tt, eval);
      else Digit0Run32<1, D, NY, true>::run(t, S, A, B, tt, eval);
    }
  }
  const uint32_t nsteps = q > 0 ? nblocks * (uint32_t)D : 1u;

  if constexpr (MODE == MODE_MAX) {
    if (active) unit_best[local] = best;
    const int32_t m = wave_max_i32(active ? best : INT32_MIN);
    // gbest only grows, so any value read from it is a lower bound of the
    // current one: skip the (device-scope, serialising) atomic when it
    // cannot raise the maximum.  The value read is made wave-uniform
    // (readfirstlane) so the branch is uniform: a lane-divergent branch here
    // made the compiler structurise the whole walk loop again, with one
    // register copy per state word and step.
    if (m > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, m);
  } else {
    const uint64_t r = (active && first != 0xFFFFFFFFu) ? (seg * (uint64_t)nsteps + first) : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
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

// Word formats of the tabulated walks.
//   WB = 16: two columns per word, each half biased by 2^15: v_pk_add_u16,
//            v_pk_sub_u16, v_sad_u16 (two |.| per instruction).  Exact while
//            every column's absolute sum is <= 2^15 - 1.
//   WB = 32: one column per word, biased by 2^31: v_add_u32, v_sub_u32,
//            v_sad_u32.  Exact under the int32 contract sum |A| <= 2^31 - 1,
//            which bounds every label sum, subset sum and objective.
// w_zero: the first |.| of a 32-bit chain accumulates onto a zero the
// compiler cannot see (a literal 0 turns v_sad_u32 into three instructions).
template <int WB>
__device__ __forceinline__ uint32_t w_add(uint32_t a, uint32_t b) {
  if constexpr (WB == 16) return pk_add(a, b);
  else return a + b;
}
template <int WB>
__device__ __forceinline__ uint32_t w_sub(uint32_t a, uint32_t b) {
  if constexpr (WB == 16) return pk_sub(a, b);
  else return a - b;
}
template <int WB>
__device__ __forceinline__ uint32_t w_sad(uint32_t a, uint32_t b, uint32_t c) {
  if constexpr (WB == 16) return sad16(a, b, c);
  else return sad32(a, b, c);
}
template <int WB>
__device__ __forceinline__ uint32_t w_zero() {
  if constexpr (WB == 16) return 0u;
  else return opaque_zero();
}
template <int WB>
constexpr uint32_t w_bias() { return WB == 16 ? kBias16x2 : kBias32; }
// A finished 32-bit |.| sum, made opaque.  v_sad_u16 is an intrinsic the
// compiler cannot look into; the 32-bit |a - b| + c is plain arithmetic, and
// without this the compiler re-associated the labelings' sums of table
// entries into the |.| chains themselves and kept hundreds of partial sums
// live (spills from d = 4 on).
template <int WB>
__device__ __forceinline__ uint32_t w_done(uint32_t x) {
  if constexpr (WB == 32) asm("" : "+v"(x));
  return x;
}

// The packed counterpart of seed_add32 (WB = 32: the same in 32-bit words).
template <int D, int NW, int WB = 16>
__device__ __forceinline__ void seed_add16(uint32_t (&t)[D][NW], const cuint32_t *row, int lab) {
  uint32_t mask[D];
#pragma unroll
  for (int c = 0; c < D; ++c) mask[c] = (lab == c) ? 0xFFFFFFFFu : 0u;
#pragma unroll
  for (int l = 0; l < NW; ++l) {
    const uint32_t v = row[l];
#pragma unroll
    for (int c = 0; c < D; ++c) t[c][l] = w_add<WB>(t[c][l], v & mask[c]);
  }
}

template <int F, int T, int D, int NW>
__device__ __forceinline__ void apply_move16(uint32_t (&t)[D][NW], uint32_t (&S)[D],
                                             const cuint32_t *row, uint32_t B) {
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
                                             const cuint32_t *row, int from, int up,
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

// The d-1 moves of row 0 inside one block of d steps: up 0->1->..->d-1, or
// down d-1->..->0, each followed by an evaluation of step ++tt.
template <int K, int D, int NW, bool UP>
struct Digit0Run {
  template <class Eval>
  __device__ __forceinline__ static void run(uint32_t (&t)[D][NW], uint32_t (&S)[D], const cuint32_t *row0,
                                             uint32_t B, uint32_t &tt, Eval &eval) {
    if constexpr (K < D) {
      constexpr int F = UP ? K - 1 : D - K;
      constexpr int T = UP ? K : D - K - 1;
      apply_move16<F, T, D, NW>(t, S, row0, B);
      eval(++tt);
      Digit0Run<K + 1, D, NW, UP>::run(t, S, row0, B, tt, eval);
    }
  }
};

// Ap: R x NW packed rows, word w = (A[i][2w+1] & 0xFFFF) << 16 | (A[i][2w] & 0xFFFF).
template <int D, int NW, int MODE>
__device__ __forceinline__ void walk_p16(
    const cuint32_t *Ap, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst, uint32_t B,
    uint64_t w) {
  const int lane = wave_lane();
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
    seed_add16<D, NW>(t, Ap + (size_t)i * NW, lab);
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

  // Steps come in blocks of d: step b*d changes a digit j >= 1 (the generic
  // move, dispatched on its labels), and steps b*d+1 .. b*d+d-1 move row 0
  // through all d labels in the direction of digit 0's parity, which is fixed
  // inside the block.  Those d-1 moves are unrolled with compile-time labels
  // (Digit0Run), so the digit counter and the dispatch run once per d steps.
  // The counter below covers digits 1..q-1 (block index b).
  uint64_t biaspat = 0;
  for (int i = 1; i < q; ++i) biaspat |= (uint64_t)(16 - D) << (4 * (i - 1));
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;
  uint32_t nblocks = 1;
  for (int i = 1; i < q; ++i) nblocks *= (uint32_t)D;

  eval(0);
  uint32_t tt = 0;
  if (q > 0) {
    if (pm & 1u) Digit0Run<1, D, NW, false>::run(t, S, Ap, B, tt, eval);
    else Digit0Run<1, D, NW, true>::run(t, S, Ap, B, tt, eval);
    for (uint32_t b = 1; b < nblocks; ++b) {
      const int jj = __builtin_ctzll(~cnt) >> 2;      // digit jj + 1 of the step counter changes
      const uint64_t low = (1ull << (4 * jj)) - 1ull;
      cnt = (cnt + 1ull) | (biaspat & low);
      const int v = (int)((cnt >> (4 * jj)) & 15ull) - (16 - D);
      const int j = jj + 1;
      const int p = (pm >> j) & 1;
      const int from = p ? D - v : v - 1;
      pm ^= (1u << j) - 1u;
      Mover16<0, D, NW>::run(t, S, Ap + (size_t)j * NW, from, p ? 0 : 1, B);
      eval(++tt);
      if (pm & 1u) Digit0Run<1, D, NW, false>::run(t, S, Ap, B, tt, eval);
      else Digit0Run<1, D, NW, true>::run(t, S, Ap, B, tt, eval);
    }
  }
  const uint32_t nsteps = q > 0 ? nblocks * (uint32_t)D : 1u;

  if constexpr (MODE == MODE_MAX) {
    if (active) unit_best[local] = best;
    const int32_t m = wave_max_i32(active ? best : INT32_MIN);
    // gbest only grows, so any value read from it is a lower bound of the
    // current one: skip the (device-scope, serialising) atomic when it
    // cannot raise the maximum.  The value read is made wave-uniform
    // (readfirstlane) so the branch is uniform: a lane-divergent branch here
    // made the compiler structurise the whole walk loop again, with one
    // register copy per state word and step.
    if (m > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, m);
  } else {
    const uint64_t r = (active && first != 0xFFFFFFFFu) ? (seg * (uint64_t)nsteps + first) : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
  }
}

// ---------------------------------------------------------------------------
// Blocked packed d-ary walk (d = 2..10): the M fastest rows are tabulated.
//
// Rank N = h * d^M + l.  The high index h walks rows M..R-2 by the reflected
// d-ary code (digit j of h is row M + j); the low index l runs through all
// d^M labelings of rows 0..M-1.  Digit i < M of N is P[(N / d^i) mod 2d]
// (L_OpenMP.c:127, :211-214), and N / d^i = h * d^(M-i) + l / d^i, where
// h * d^(M-i) = d * (h & 1) mod 2d for odd d, and for even d when i = M-1
// (else 0 mod 2d).  So the low labels of rank N are those of rank l,
// complemented (c -> d-1-c) when h is odd: every low label for odd d, the
// top one (row M-1) for even d.
//
// For label c let u_c be its packed sum over the rows outside the table
// (rows M..R-2 by their labels, and row R-1, which has label 0).  For a
// subset S of the low rows let
//   F_c(S) = sum_l |u_c,l + A_S,l| = sum_w sad16(u_c[w], neg_S[w]),
// with neg_S = -A_S + 2^15 per 16-bit half (a host-built table, read as
// scalar operands; S = {} gives the plain |u_c|).  A labeling puts the low
// rows into disjoint subsets S_c, and its objective is sum_c F_c(S_c).
// A high step moves one row between two labels and changes the tables of
// those two labels only: 2 * 2^M * NW v_sad_u16 for the d^M labelings of
// the block.  Each labeling then costs one sum of table entries at compile-
// time indices and half a v_max3:
//   d = 2, 3: F_0 + F_1 (+ F_2), one add;
//   d >= 4: sum_c |u_c| + sum over the labels with S_c != {} of
//     G_c(S_c) = F_c(S_c) - F_c({}), at most M terms.
// Exact under the packed walk's bound (every column's absolute sum <= 32767),
// which bounds every |u_c,l + A_S,l|.
// ---------------------------------------------------------------------------
constexpr int ipow_c(int b, int e) { return e == 0 ? 1 : b * ipow_c(b, e - 1); }

// Widest instantiation per d (16-bit words), and the rows tabulated at a
// width: d^M labelings per high step against 2^(M+1) table sums per state
// word, with the d * NW state and the d * 2^M table held in registers.
// Every (d, NW, M) below compiles without VGPR spills (checked in the
// register report of the build); M shrinks as the state grows.  (d = 2:
// both labels change at every high step, 2 sums per word and labeling
// whatever M.)
constexpr int blk16_max_words(int d) {
  return d == 2 ? 96 : d == 3 ? 64 : d == 4 ? 48 : d == 5 ? 40 : (d == 6 || d == 7) ? 32
                                              : (d == 8 || d == 9) ? 24 : d == 10 ? 16 : 0;
}
constexpr int blk16_low_rows(int d, int nw) {
  if (nw > blk16_max_words(d)) return 0;
  switch (d) {
    case 2: return nw <= 32 ? 5 : 4;
    case 3: return nw <= 16 ? 6 : nw <= 32 ? 5 : 4;   // 6 rows: 729 labelings per block (247 VGPRs at 16 words)
    case 4: return nw <= 32 ? 4 : nw <= 40 ? 3 : 2;
    case 5: return nw <= 16 ? 4 : nw <= 32 ? 3 : 2;   // 4 rows: 625 labelings per block (234 VGPRs at 16 words)
    case 6: return nw <= 8 ? 4 : nw <= 24 ? 3 : 2;   // 4 rows: 1296 labelings per block (230 VGPRs at 8 words)
    case 7: return nw <= 16 ? 3 : nw <= 24 ? 2 : 1;   // 3 rows: 343 labelings per block (192 VGPRs at 16 words)
    case 8: return nw <= 16 ? 3 : 2;   // 3 rows: 512 labelings per block (216 VGPRs at 16 words)
    case 9: return nw <= 16 ? 3 : 1;   // 729 labelings per block (244 VGPRs at 16 words)
    case 10: return nw <= 8 ? 3 : 2;   // 1000 labelings per block up to 8 words
    default: return 0;
  }
}

template <int D, int M>
struct LowCode {
  static constexpr int NPAT = ipow_c(D, M);
  // label of low row i at low index l, complemented for odd h
  static constexpr int label(int l, int i, bool odd) {
    int x = l;
    for (int k = 0; k < i; ++k) x /= D;
    const int r = x % (2 * D);
    const int c = r < D ? r : 2 * D - 1 - r;
    return (odd && ((D & 1) || i == M - 1)) ? D - 1 - c : c;
  }
  static constexpr int mask(int l, int c, bool odd) {
    int m = 0;
    for (int i = 0; i < M; ++i)
      if (label(l, i, odd) == c) m |= 1 << i;
    return m;
  }
};

template <int D, int M, bool ODD, int L, int... Cs>
__device__ __forceinline__ int32_t blk16_obj(const uint32_t (&T)[D][1 << M], uint32_t totU,
                                             std::integer_sequence<int, Cs...>) {
  using LC = LowCode<D, M>;
  if constexpr (D <= 3)
    return (int32_t)(0u + ... + T[Cs][LC::mask(L, Cs, ODD)]);
  else
    return (int32_t)(totU + ... + (LC::mask(L, Cs, ODD) ? T[Cs][LC::mask(L, Cs, ODD)] : 0u));
}

// MODE_MAX: two labelings and the running maximum meet in one v_max3.
// MODE_FIRST: fb keeps the smallest low index l in the block whose objective
// equals the target (l is a literal).
template <int D, int M, bool ODD, int MODE, int L>
__device__ __forceinline__ void blk16_eval_pair(const uint32_t (&T)[D][1 << M], uint32_t totU, int32_t &best,
                                                uint32_t &fb, int32_t target) {
  constexpr int NPAT = LowCode<D, M>::NPAT;
  constexpr int L2 = L + 1 < NPAT ? L + 1 : L;
  const auto cs = std::make_integer_sequence<int, D>{};
  const int32_t a = blk16_obj<D, M, ODD, L>(T, totU, cs);
  const int32_t b = blk16_obj<D, M, ODD, L2>(T, totU, cs);
  if constexpr (MODE == MODE_MAX) {
    const int32_t m = a > b ? a : b;
    best = m > best ? m : best;
  } else {
    const uint32_t ca = (a == target) ? (uint32_t)L : 0xFFFFFFFFu;
    const uint32_t cb = (b == target) ? (uint32_t)L2 : 0xFFFFFFFFu;
    const uint32_t c = ca < cb ? ca : cb;
    fb = c < fb ? c : fb;
  }
}

template <int D, int M, bool ODD, int MODE, int... Ps>
__device__ __forceinline__ void blk16_eval(const uint32_t (&T)[D][1 << M], uint32_t totU, int32_t &best,
                                           uint32_t &fb, int32_t target, std::integer_sequence<int, Ps...>) {
  (blk16_eval_pair<D, M, ODD, MODE, 2 * Ps>(T, totU, best, fb, target), ...);
}

// Table of label C from its sums u[C]: T[C][0] = |u_C|, and T[C][S] = F_C(S)
// (d <= 3) or F_C(S) - |u_C| (d >= 4, the accumulator starts at -|u_C|).
template <int C, int D, int NW, int M, int WB>
__device__ __forceinline__ void blk16_table(const uint32_t (&u)[D][NW], uint32_t (&T)[D][1 << M],
                                            const cuint32_t *neg) {
  const uint32_t z = w_zero<WB>();
  uint32_t U = z;
#pragma unroll
  for (int w = 0; w < NW; ++w) U = w_sad<WB>(u[C][w], neg[w], U);
  U = w_done<WB>(U);
  T[C][0] = U;
  const uint32_t base = D <= 3 ? z : z - U;
#pragma unroll
  for (int S = 1; S < (1 << M); ++S) {
    uint32_t acc = base;
#pragma unroll
    for (int w = 0; w < NW; ++w) acc = w_sad<WB>(u[C][w], neg[(size_t)S * NW + w], acc);
    T[C][S] = w_done<WB>(acc);
  }
}

template <int D, int NW, int M, int WB, int... Cs>
__device__ __forceinline__ void blk16_tables_all(const uint32_t (&u)[D][NW], uint32_t (&T)[D][1 << M],
                                                 const cuint32_t *neg, std::integer_sequence<int, Cs...>) {
  (blk16_table<Cs, D, NW, M, WB>(u, T, neg), ...);
}

// Tables of the two labels a high step changes, C and C + 1, in one pass
// over the subset table: each scalar word feeds two v_sad_u16 (one per
// label), half the scalar loads of two single-label passes, and the two
// accumulator chains are independent.
template <int C, int D, int NW, int M, int WB>
__device__ __forceinline__ void blk16_table_pair(const uint32_t (&u)[D][NW], uint32_t (&T)[D][1 << M],
                                                 const cuint32_t *neg) {
  const uint32_t z = w_zero<WB>();
  uint32_t U0 = z, U1 = z;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint32_t n = neg[w];
    U0 = w_sad<WB>(u[C][w], n, U0);
    U1 = w_sad<WB>(u[C + 1][w], n, U1);
  }
  U0 = w_done<WB>(U0);
  U1 = w_done<WB>(U1);
  T[C][0] = U0;
  T[C + 1][0] = U1;
  const uint32_t b0 = D <= 3 ? z : z - U0, b1 = D <= 3 ? z : z - U1;
#pragma unroll
  for (int S = 1; S < (1 << M); ++S) {
    uint32_t a0 = b0, a1 = b1;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t n = neg[(size_t)S * NW + w];
      a0 = w_sad<WB>(u[C][w], n, a0);
      a1 = w_sad<WB>(u[C + 1][w], n, a1);
    }
    T[C][S] = w_done<WB>(a0);
    T[C + 1][S] = w_done<WB>(a1);
  }
}

template <int C, int D, int NW, int M, int WB>
struct Blk16TablePair {
  __device__ __forceinline__ static void run(int lo, const uint32_t (&u)[D][NW], uint32_t (&T)[D][1 << M],
                                             const cuint32_t *neg) {
    if constexpr (C + 1 < D) {
      if (lo == C) blk16_table_pair<C, D, NW, M, WB>(u, T, neg);
      else Blk16TablePair<C + 1, D, NW, M, WB>::run(lo, u, T, neg);
    }
  }
};

// Moves one row (WB-bit words) from label `from` to from + 1 (up) or from - 1.
template <int C, int D, int NW, int WB = 16>
struct Move16 {
  __device__ __forceinline__ static void run(uint32_t (&u)[D][NW], const cuint32_t *row, int from, int up) {
    if (from == C) {
      constexpr int TU = C + 1 < D ? C + 1 : C, TD = C > 0 ? C - 1 : C;
      if (up) {
        if constexpr (TU != C) {
#pragma unroll
          for (int w = 0; w < NW; ++w) {
            u[C][w] = w_sub<WB>(u[C][w], row[w]);
            u[TU][w] = w_add<WB>(u[TU][w], row[w]);
          }
        }
      } else {
        if constexpr (TD != C) {
#pragma unroll
          for (int w = 0; w < NW; ++w) {
            u[C][w] = w_sub<WB>(u[C][w], row[w]);
            u[TD][w] = w_add<WB>(u[TD][w], row[w]);
          }
        }
      }
    } else if constexpr (C + 1 < D) {
      Move16<C + 1, D, NW, WB>::run(u, row, from, up);
    }
  }
};

// Ap: R x NW packed rows (as walk_p16); neg: 2^M x NW packed words -A_S +
// 2^15.  q >= M: a segment is d^(q-M) high steps of d^M labelings each.
// WB = 32: Ap holds R x NW int32 columns and neg the words -A_S + 2^31.
template <int D, int NW, int M, int MODE, int WB = 16>
__device__ __forceinline__ void walk_blk16(
    const cuint32_t *Ap, const cuint32_t *neg, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst,
    uint64_t w) {
  constexpr int NPAT = LowCode<D, M>::NPAT;
  const int lane = wave_lane();
  if (((w >> 1) << 7) >= nseg) return;
  const uint64_t local = lane_segment(w, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ w) & 1));
  const int qh = q - M;   // high digits inside the segment

  uint32_t par_mask;
  if (D & 1) par_mask = pseg ? ((1u << qh) - 1u) : 0u;
  else par_mask = (qh > 0 && pseg) ? (1u << (qh - 1)) : 0u;
  // parity of h at the segment's first high step
  const uint32_t hpar0 = ((D & 1) || qh == 0) ? pseg : 0u;

  uint32_t u[D][NW];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int l = 0; l < NW; ++l) u[c][l] = w_add<WB>(w_bias<WB>(), c == 0 ? Ap[(size_t)(R - 1) * NW + l] : 0u);
  uint64_t x = seg;
  for (int i = 0; i < R - 1 - M; ++i) {
    int lab;
    if (i < qh) {
      lab = ((par_mask >> i) & 1) ? D - 1 : 0;
    } else {
      const uint32_t r = (uint32_t)(x % (uint64_t)(2 * D));
      lab = r < (uint32_t)D ? (int)r : (int)(2 * D - 1 - r);
      x /= (uint64_t)D;
    }
    seed_add16<D, NW, WB>(u, Ap + (size_t)(M + i) * NW, lab);
  }
  uint32_t T[D][1 << M];
  blk16_tables_all<D, NW, M, WB>(u, T, neg, std::make_integer_sequence<int, D>{});
  uint32_t totU = 0;
#pragma unroll
  for (int c = 0; c < D; ++c) totU += T[c][0];

  uint64_t biaspat = 0;
  for (int i = 0; i < qh; ++i) biaspat |= (uint64_t)(16 - D) << (4 * i);
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;
  uint32_t nblk = 1;
  for (int i = 0; i < qh; ++i) nblk *= (uint32_t)D;

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  const auto ps = std::make_integer_sequence<int, (NPAT + 1) / 2>{};
  for (uint32_t k = 0; k < nblk; ++k) {
    if (k > 0) {
      const int j = __builtin_ctzll(~cnt) >> 2;      // high digit j (row M + j) changes
      const uint64_t low = (1ull << (4 * j)) - 1ull;
      cnt = (cnt + 1ull) | (biaspat & low);
      const int v = (int)((cnt >> (4 * j)) & 15ull) - (16 - D);
      const int p = (pm >> j) & 1;
      const int from = p ? D - v : v - 1;
      const int up = p ? 0 : 1;
      pm ^= (1u << j) - 1u;
      Move16<0, D, NW, WB>::run(u, Ap + (size_t)(M + j) * NW, from, up);
      // The table's scalar loads are loop-invariant; hoisted out of the
      // loop, 2^M * NW values would not fit in SGPRs and be spilled to VGPR
      // lanes (one v_readlane per use), so the pointer is laundered here.
      const cuint32_t *negk = neg;
      asm volatile("" : "+s"(negk));
      Blk16TablePair<0, D, NW, M, WB>::run(up ? from : from - 1, u, T, negk);
      totU = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) totU += T[c][0];
    }
    if constexpr (MODE == MODE_MAX) {
      // A block holds all d^M low labelings whatever the parity of h; only
      // their order differs, and the maximum does not depend on it.  (Both
      // orders in one loop also let the compiler merge their identical sums
      // into one set of d^M live values, which spilled.)
      uint32_t unused = 0;
      blk16_eval<D, M, false, MODE>(T, totU, best, unused, target, ps);
    } else {
      uint32_t fb = 0xFFFFFFFFu;
      if ((hpar0 ^ k) & 1u) {
        // The two orders sum the same table entries; an identity barrier on
        // the table keeps the compiler from hoisting all d^M sums out of
        // both branches as live values.
#pragma unroll
        for (int c = 0; c < D; ++c)
#pragma unroll
          for (int S = 0; S < (1 << M); ++S) asm volatile("" : "+v"(T[c][S]));
        asm volatile("" : "+v"(totU));
        blk16_eval<D, M, true, MODE>(T, totU, best, fb, target, ps);
      } else {
        blk16_eval<D, M, false, MODE>(T, totU, best, fb, target, ps);
      }
      const uint32_t c = fb != 0xFFFFFFFFu ? k * (uint32_t)NPAT + fb : 0xFFFFFFFFu;
      first = c < first ? c : first;
    }
  }
  const uint32_t nsteps = nblk * (uint32_t)NPAT;

  if constexpr (MODE == MODE_MAX) {
    if (active) unit_best[local] = best;
    const int32_t m = wave_max_i32(active ? best : INT32_MIN);
    if (m > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, m);
  } else {
    const uint64_t r = (active && first != 0xFFFFFFFFu) ? (seg * (uint64_t)nsteps + first) : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
  }
}

// ---------------------------------------------------------------------------
// Tabulated binary walk for L1 / LM (packed 16-bit).  Rank N = h * 2^M + l:
// BRGC bit i of N (L_OpenMP.c:58-69, bit i of N ^ (N >> 1)) is bit i of
// l ^ (l >> 1) for i < M - 1, that bit xor (h & 1) for i = M - 1, and bit
// i - M of h ^ (h >> 1) for the high rows.  With the lane's high sum
//   u = A_{R-1} + sum_{i >= M} s_i A_i   (s_i = 2 b_i - 1),
// low sign pattern S (bit i = b_i; M = 5 rows, 4 beyond 64 words) has the
// objective sum_l |u_l + A_S,l|,
// A_S = sum_{i < M} (2 S_i - 1) A_i: one v_sad_u16 per two columns against
// the scalar table word -A_S + 2^15.  LM adds column 0 without |.|
// (L_OpenMP.c:30, :40-41): u_0 + A_S,0 in 32 bits starts the accumulator.
// A high step adds +-2 A_r to u, one packed add per word for 2^M sign
// vectors.  Exact when every |.| column's absolute sum is <= 32767.
//
// Ap: R x NW packed |.| columns (LM: columns 1..C-1), then R x NW packed
// 2 A, then (LM) the R column-0 entries.  Tab: 2^M x NW words -A_S + 2^15,
// then (LM) the 2^M column-0 sums A_S,0.
// ---------------------------------------------------------------------------

template <int NW, int M, bool MARG, int WB, int MODE, int S>
__device__ __forceinline__ int32_t tb16_obj(const uint32_t (&u)[NW], int32_t u0, const cuint32_t *neg,
                                            const cint32_t *pos0) {
  uint32_t acc = MARG ? (uint32_t)(u0 + pos0[S]) : w_zero<WB>();
#pragma unroll
  for (int w = 0; w < NW; ++w) acc = w_sad<WB>(u[w], neg[(size_t)S * NW + w], acc);
  return (int32_t)acc;
}

// MODE_MAX: patterns in table order, two per v_max3.
template <int NW, int M, bool MARG, int WB, int... Ps>
__device__ __forceinline__ void tb16_max(const uint32_t (&u)[NW], int32_t u0, const cuint32_t *neg,
                                         const cint32_t *pos0, int32_t &best, std::integer_sequence<int, Ps...>) {
  auto pair = [&](auto P) {
    constexpr int S = decltype(P)::value;
    const int32_t a = tb16_obj<NW, M, MARG, WB, MODE_MAX, 2 * S>(u, u0, neg, pos0);
    const int32_t b = tb16_obj<NW, M, MARG, WB, MODE_MAX, 2 * S + 1>(u, u0, neg, pos0);
    const int32_t m = a > b ? a : b;
    best = m > best ? m : best;
  };
  (pair(std::integral_constant<int, Ps>{}), ...);
}

// MODE_MAX of the L1 walk at 64 and 128 16-bit words: the patterns one at a
// time in table order, each pair's maximum by one v_max3, the table words read
// in batches of KB, each batch's loads issued right after the first operation
// of the batch before it (the one that waits for that batch: scalar loads
// return out of order, so a wait covers every load in flight), pinned by
// scheduling barriers.  1.5-1.9% ahead of the compiler's order there; LM ran
// 5.5% slower, and at 96, 160 and 192 words the compiler did not unroll the
// batches (scratch spills), so those keep tb16_max.
template <int NW, int M, int WB, int KB>
__device__ __forceinline__ void tb16_max_batched(const uint32_t (&u)[NW], const cuint32_t *neg, int32_t &best) {
  constexpr int NB = (1 << M) * NW / KB;
  static_assert(NW % KB == 0, "batches tile a pattern's words");
  uint32_t buf[2][KB];
#pragma unroll
  for (int i = 0; i < KB; ++i) buf[0][i] = neg[i];
  uint32_t acc = 0;
  int32_t prev = 0;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int f = b * KB + i, S = f / NW, w = f % NW;
      acc = w_sad<WB>(u[w], buf[b & 1][i], w ? acc : w_zero<WB>());
      if (i == 0) {
        __builtin_amdgcn_sched_barrier(0);
        if (b + 1 < NB) {
#pragma unroll
          for (int j = 0; j < KB; ++j) buf[(b + 1) & 1][j] = neg[(size_t)(b + 1) * KB + j];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      if (w == NW - 1) {
        const int32_t t = (int32_t)w_done<WB>(acc);
        if (S & 1) {
          const int32_t m = prev > t ? prev : t;
          best = m > best ? m : best;
        } else {
          prev = t;
        }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// MODE_FIRST: patterns in the order of l; fb = the first l matching.  It
// re-walks a single unit, so the patterns run as a loop (the wave-uniform
// pattern index addresses the scalar table) rather than unrolled: half the
// code of the two unrolled orders.
template <int NW, int M, bool MARG, int WB>
__device__ __forceinline__ void tb16_first(const uint32_t (&u)[NW], int32_t u0, const cuint32_t *neg,
                                           const cint32_t *pos0, int32_t target, uint32_t odd, uint32_t &fb) {
  const uint32_t flip = odd ? (1u << (M - 1)) : 0u;
#pragma unroll 1
  for (uint32_t l = 0; l < (1u << M); ++l) {
    const uint32_t S = (l ^ (l >> 1)) ^ flip;
    const cuint32_t *row = neg + (size_t)S * NW;
    uint32_t acc = MARG ? (uint32_t)(u0 + pos0[S]) : w_zero<WB>();
#pragma unroll
    for (int w = 0; w < NW; ++w) acc = w_sad<WB>(u[w], row[w], acc);
    const uint32_t c = ((int32_t)acc == target) ? l : 0xFFFFFFFFu;
    fb = c < fb ? c : fb;
  }
}

template <int NW, bool MARG, int MODE, int WB = 16>
__device__ __forceinline__ void walk_tab16_bin(
    const cuint32_t *Ap, const cuint32_t *Tab, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst,
    uint64_t w) {
  constexpr int M = WB == 16 ? tab_bin_rows(NW) : tab32_bin_rows(NW);
  const int lane = wave_lane();
  if (((w >> 1) << 7) >= nseg) return;
  const uint64_t local = lane_segment(w, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ w) & 1));
  const int qh = q - M;
  const cuint32_t *A2 = Ap + (size_t)R * NW;                       // packed 2 A
  const cint32_t *A0 = (const cint32_t *)(Ap + (size_t)2 * R * NW);  // LM: column 0
  const cint32_t *pos0 = (const cint32_t *)(Tab + ((size_t)NW << M));

  // seed: the high rows' BRGC bits at h0 = seg * 2^qh
  const uint64_t h0 = seg << qh;
  const uint64_t g0 = h0 ^ (h0 >> 1);
  uint32_t u[NW];
  int32_t u0 = MARG ? A0[R - 1] : 0;
#pragma unroll
  for (int l = 0; l < NW; ++l) u[l] = w_add<WB>(w_bias<WB>(), Ap[(size_t)(R - 1) * NW + l]);
  for (int j = 0; j < R - 1 - M; ++j) {
    const bool b = (g0 >> j) & 1;
    const cuint32_t *row = Ap + (size_t)(M + j) * NW;
#pragma unroll
    for (int l = 0; l < NW; ++l) u[l] = b ? w_add<WB>(u[l], row[l]) : w_sub<WB>(u[l], row[l]);
    if (MARG) u0 += b ? A0[M + j] : -A0[M + j];
  }

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  const uint32_t nblk = 1u << qh;
  for (uint32_t k = 0; k < nblk; ++k) {
    if (k > 0) {
      // high digit j flips to 1 ^ (bit j + 1 of h); bit qh of h is the
      // segment's parity
      const int j = __builtin_ctz(k);
      const uint32_t hb = (j + 1 < qh) ? ((k >> (j + 1)) & 1u) : pseg;
      const cuint32_t *row2 = A2 + (size_t)(M + j) * NW;
      if (hb == 0) {
#pragma unroll
        for (int l = 0; l < NW; ++l) u[l] = w_add<WB>(u[l], row2[l]);
        if (MARG) u0 += 2 * A0[M + j];
      } else {
#pragma unroll
        for (int l = 0; l < NW; ++l) u[l] = w_sub<WB>(u[l], row2[l]);
        if (MARG) u0 -= 2 * A0[M + j];
      }
    }
    // the table's 2^M * NW scalar loads stay inside the loop (hoisted, they
    // would not fit in SGPRs)
    const cuint32_t *neg = Tab;
    asm volatile("" : "+s"(neg));
    if constexpr (MODE == MODE_MAX && !MARG && WB == 16 && (NW == 64 || NW == 128)) {
      tb16_max_batched<NW, M, WB, 32>(u, neg, best);
    } else if constexpr (MODE == MODE_MAX) {
      tb16_max<NW, M, MARG, WB>(u, u0, neg, pos0, best, std::make_integer_sequence<int, (1 << M) / 2>{});
    } else {
      uint32_t fb = 0xFFFFFFFFu;
      const uint32_t hp = qh > 0 ? (k & 1u) : pseg;
      tb16_first<NW, M, MARG, WB>(u, u0, neg, pos0, target, hp, fb);
      const uint32_t c = fb != 0xFFFFFFFFu ? (k << M) + fb : 0xFFFFFFFFu;
      first = c < first ? c : first;
    }
  }
  const uint32_t nsteps = nblk << M;

  if constexpr (MODE == MODE_MAX) {
    if (active) unit_best[local] = best;
    const int32_t m = wave_max_i32(active ? best : INT32_MIN);
    if (m > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, m);
  } else {
    const uint64_t r = (active && first != 0xFFFFFFFFu) ? (seg * (uint64_t)nsteps + first) : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
  }
}

}  // namespace lnorm

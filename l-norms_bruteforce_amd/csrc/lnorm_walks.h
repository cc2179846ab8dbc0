// lnorm_walks.h -- the Gray-code walks as device functions.
//
// Each walk_* runs one wave's share of a segment walk: wave index w of a
// launch (or of one item of a batch) owns the 64 segments lane_segment(w, .)
// and writes unit_best / gbest (MODE_MAX) or gfirst (MODE_FIRST).  The
// single-problem kernels (lnorm_plain.hip) call them with global_wave(); the
// persistent batch kernels (lnorm_batch.hip) with the wave index of a task.
#pragma once
#include "lnorm_device.h"

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
    const cint32_t *row = A + (size_t)i * NY;
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
    const cuint32_t *row = Ap + (size_t)i * NW;
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

// largest packed width instantiated per d (d * NW <= 160 words of state)

}  // namespace lnorm

// lnorm_int64.hip -- the int_type = long long variant of the enumeration.
//
// The reference can be rebuilt with 64-bit matrix entries and sums
// (README.md:100-105: int_type = long long, abs -> llabs).  These kernels
// compute that variant for matrices whose sums do not fit int32
// (sum |A| <= INT64_MAX keeps every partial sum and objective in range);
// problems inside the int32 bound take the int32 kernels, which give the
// same result.
//
// Same decomposition as the int32 walks: aligned segments, one per lane,
// wave-uniform flipped row and direction (lane_segment), row deltas through
// scalar loads, partial sums in VGPR pairs.  A 64-bit row update is an add
// with carry (2 ops) and |y| + acc a compare, negate and select plus a
// 64-bit add (about 7 ops): roughly 4.5x the int32 cost per row-update.  The
// MFMA-blocked path has no 64-bit counterpart.
//
//   k_plain64_bin<KIND, NY>  L1 / LM / L2, NY up to 64 (L2: 32) columns
//   k_plainf64_bin<KIND, NY> the same walk in doubles, for sum |A| < 2^52
//   k_widef64_bin<KIND,NY,W> the walk in doubles over W waves: L1 / LM up to
//                            512 columns, L2 up to 256
//   k_widef64_dary<D,NY,W>   the d-ary walk in doubles over 2, 4 or 8 waves:
//                            up to 192 columns at d = 3, 128 at 4, 64 beyond
//   k_plain64_dary<D, NY, V> L3..L10, d * NY <= 80 words of state (V = double
//                            while sum |A| < 2^52, else uint64_t)
//   k_general64              any width: wave per segment, columns over lanes
#include "lnorm_device.h"

namespace lnorm {

typedef __attribute__((address_space(4))) const int64_t cint64_t;
__device__ __forceinline__ const cint64_t *as_const64(const int64_t *p) { return (const cint64_t *)(uintptr_t)p; }
typedef __attribute__((address_space(4))) const double cdouble_t;

__device__ __forceinline__ int64_t abs64(int64_t v) { return v < 0 ? -v : v; }

__device__ __forceinline__ int64_t wave_max_i64(int64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int64_t o = __shfl_xor(v, off, 64);
    v = v > o ? v : o;
  }
  return (int64_t)uniform_u64((uint64_t)v);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return uniform_u64(v);
}

// MODE_MAX tail: unit maxima and the global maximum (skipping the atomic
// when a uniform read shows it cannot raise it); MODE_FIRST tail: the
// smallest matching rank.
template <int MODE>
__device__ __forceinline__ void finish64(bool active, uint64_t local, int64_t best, uint64_t rank_of_first,
                                         bool found, int64_t *unit_best, long long *gbest,
                                         unsigned long long *gfirst) {
  if constexpr (MODE == MODE_MAX) {
    if (active) unit_best[local] = best;
    const int64_t m = wave_max_i64(active ? best : INT64_MIN);
    const int64_t cur = (int64_t)uniform_u64((uint64_t)*(volatile long long *)gbest);
    if (m > cur) atomicMax(gbest, (long long)m);
  } else {
    const uint64_t r = (active && found) ? rank_of_first : ~0ull;
    const uint64_t m = wave_min_u64(r);
    if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
  }
}

// ---------------------------------------------------------------------------
// Binary family: D holds [2][R-1][NY]: D[1][j] = +2A[j] (L1/LM) or +A[j]
// (L2), D[0][j] its negation (L_OpenMP.c:40-41, 77-78, 116-117).
// ---------------------------------------------------------------------------
template <int KIND, int NY, int MODE>
__global__ __launch_bounds__(256) void k_plain64_bin(const int64_t *__restrict__ A_, const int64_t *__restrict__ D_,
                                                     int R, int q, uint64_t seg0, uint64_t nseg, int64_t target,
                                                     int64_t *__restrict__ unit_best, long long *__restrict__ gbest,
                                                     unsigned long long *__restrict__ gfirst) {
  const cint64_t *A = as_const64(A_), *Dt = as_const64(D_);
  const uint64_t w = global_wave();
  const int lane = wave_lane();
  if (((w >> 1) << 7) >= nseg) return;
  const uint64_t local = lane_segment(w, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t par = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ w) & 1));

  constexpr bool TWO = (KIND == KIND_L2);
  uint64_t y[NY];
  uint64_t z[TWO ? NY : 1];
  const uint64_t N0 = seg << q;
  const uint64_t g0 = N0 ^ (N0 >> 1);
#pragma unroll
  for (int l = 0; l < NY; ++l) y[l] = (uint64_t)A[(size_t)(R - 1) * NY + l];
  if constexpr (TWO) {
#pragma unroll
    for (int l = 0; l < NY; ++l) z[l] = 0;
  }
  for (int i = 0; i < R - 1; ++i) {
    const bool g = (g0 >> i) & 1;
    const cint64_t *row = A + (size_t)i * NY;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const uint64_t a = (uint64_t)row[l];
      if constexpr (TWO) {
        z[l] += g ? a : 0ull;
        y[l] += g ? 0ull : a;
      } else {
        y[l] += g ? a : 0ull - a;
      }
    }
  }

  int64_t best = INT64_MIN;
  uint32_t first = 0xFFFFFFFFu;
  auto eval = [&](uint32_t t) {
    int64_t acc;
    if constexpr (KIND == KIND_LM) {
      acc = (int64_t)y[0];
#pragma unroll
      for (int l = 1; l < NY; ++l) acc += abs64((int64_t)y[l]);
    } else {
      acc = 0;
#pragma unroll
      for (int l = 0; l < NY; ++l) {
        acc += abs64((int64_t)y[l]);
        if constexpr (TWO) acc += abs64((int64_t)z[l]);
      }
    }
    if constexpr (MODE == MODE_MAX) {
      best = acc > best ? acc : best;
    } else {
      const uint32_t c = (acc == target) ? t : 0xFFFFFFFFu;
      first = c < first ? c : first;
    }
  };

  const uint32_t nsteps = 1u << q;
  const uint32_t nu = par << q;
  eval(0);
  for (uint32_t t = 1; t < nsteps; ++t) {
    const int j = __builtin_ctz(t);
    const uint32_t up = (((nu | t) >> (j + 1)) & 1u) ^ 1u;
    const cint64_t *dp = Dt + (size_t)(up * (uint32_t)(R - 1) + (uint32_t)j) * NY;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const uint64_t dl = (uint64_t)dp[l];
      if constexpr (TWO) {
        z[l] += dl;
        y[l] -= dl;
      } else {
        y[l] += dl;
      }
    }
    eval(t);
  }
  finish64<MODE>(active, local, best, (seg << q) + first, first != 0xFFFFFFFFu, unit_best, gbest, gfirst);
}

// ---------------------------------------------------------------------------
// Binary family in double precision, for sum |A| < 2^52: every entry, delta
// (f A, f <= 2), partial sum and objective is then an integer of magnitude
// below 2^53, which a double holds exactly, and every add of two of them
// whose result is such an integer is exact, so the walk computes the same
// integers as k_plain64_bin.  A row update is one v_add_f64 and |y| + acc one
// v_add_f64 with the |.| source modifier: 2 ops per row-update against about
// 9 for the 64-bit integer walk.  A and D hold the doubles of the same
// layout (the host converts them).
// ---------------------------------------------------------------------------
template <int KIND, int NY, int MODE>
__global__ __launch_bounds__(256) void k_plainf64_bin(const double *__restrict__ A_, const double *__restrict__ D_,
                                                      int R, int q, uint64_t seg0, uint64_t nseg, int64_t target,
                                                      int64_t *__restrict__ unit_best, long long *__restrict__ gbest,
                                                      unsigned long long *__restrict__ gfirst) {
  const cdouble_t *A = (const cdouble_t *)(uintptr_t)A_, *Dt = (const cdouble_t *)(uintptr_t)D_;
  const uint64_t w = global_wave();
  const int lane = wave_lane();
  if (((w >> 1) << 7) >= nseg) return;
  const uint64_t local = lane_segment(w, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t par = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ w) & 1));

  constexpr bool TWO = (KIND == KIND_L2);
  double y[NY];
  double z[TWO ? NY : 1];
  const uint64_t N0 = seg << q;
  const uint64_t g0 = N0 ^ (N0 >> 1);
#pragma unroll
  for (int l = 0; l < NY; ++l) y[l] = A[(size_t)(R - 1) * NY + l];
  if constexpr (TWO) {
#pragma unroll
    for (int l = 0; l < NY; ++l) z[l] = 0.0;
  }
  for (int i = 0; i < R - 1; ++i) {
    const bool g = (g0 >> i) & 1;
    const cdouble_t *row = A + (size_t)i * NY;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const double a = row[l];
      if constexpr (TWO) {
        z[l] += g ? a : 0.0;
        y[l] += g ? 0.0 : a;
      } else {
        y[l] += g ? a : -a;
      }
    }
  }

  const double dtarget = (double)target;
  double best = -__builtin_inf();   // eval(0) below makes it finite
  uint32_t first = 0xFFFFFFFFu;
  auto eval = [&](uint32_t t) {
    double acc;
    if constexpr (KIND == KIND_LM) {
      acc = y[0];
#pragma unroll
      for (int l = 1; l < NY; ++l) acc += __builtin_fabs(y[l]);
    } else {
      acc = 0.0;
#pragma unroll
      for (int l = 0; l < NY; ++l) {
        acc += __builtin_fabs(y[l]);
        if constexpr (TWO) acc += __builtin_fabs(z[l]);
      }
    }
    if constexpr (MODE == MODE_MAX) {
      best = __builtin_fmax(best, acc);
    } else {
      const uint32_t c = (acc == dtarget) ? t : 0xFFFFFFFFu;
      first = c < first ? c : first;
    }
  };

  const uint32_t nsteps = 1u << q;
  const uint32_t nu = par << q;
  eval(0);
  for (uint32_t t = 1; t < nsteps; ++t) {
    const int j = __builtin_ctz(t);
    const uint32_t up = (((nu | t) >> (j + 1)) & 1u) ^ 1u;
    const cdouble_t *dp = Dt + (size_t)(up * (uint32_t)(R - 1) + (uint32_t)j) * NY;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const double dl = dp[l];
      if constexpr (TWO) {
        z[l] += dl;
        y[l] -= dl;
      } else {
        y[l] += dl;
      }
    }
    eval(t);
  }
  finish64<MODE>(active, local, (int64_t)best, (seg << q) + first, first != 0xFFFFFFFFu, unit_best, gbest, gfirst);
}

// ---------------------------------------------------------------------------
// The binary walk in doubles over W waves, for widths beyond one wave's
// registers (k_wide_bin's decomposition): wave wv keeps columns
// [wv*NY, (wv+1)*NY) of the 64 lane segments' sums.  Each wave's partial |.|
// sum of every step goes to LDS; after K steps (or the segment's end, so any
// segment length works) every wave totals K/W steps over the W partials and
// keeps their maximum (MODE_MAX) or first match (MODE_FIRST).  Exact while
// sum |A| < 2^52, as k_plainf64_bin.
// ---------------------------------------------------------------------------
template <int KIND, int NY, int W, int MODE>
__global__ __launch_bounds__(64 * W) void k_widef64_bin(const double *__restrict__ A_, const double *__restrict__ D_,
                                                        int R, int q, uint64_t seg0, uint64_t nseg, int64_t target,
                                                        int64_t *__restrict__ unit_best,
                                                        long long *__restrict__ gbest,
                                                        unsigned long long *__restrict__ gfirst) {
  constexpr int NYT = NY * W;
  constexpr int K = 8;   // steps per LDS combine
  static_assert(K % W == 0, "steps per combine must split evenly over the waves");
  constexpr bool TWO = (KIND == KIND_L2);
  __shared__ double part[W * K * 64];
  __shared__ double fin_best[W * 64];
  __shared__ uint32_t fin_first[W * 64];
  const cdouble_t *A = (const cdouble_t *)(uintptr_t)A_, *Dt = (const cdouble_t *)(uintptr_t)D_;
  const int lane = wave_lane();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t g = blockIdx.x;
  if (((g >> 1) << 7) >= nseg) return;   // uniform over the workgroup
  const uint64_t local = lane_segment(g, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t par = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ g) & 1));
  const int wbase = wv * NY;
  const bool mw = KIND == KIND_LM && wv == 0;   // the wave holding the marginal column 0

  double y[NY];
  double z[TWO ? NY : 1];
  const uint64_t N0 = seg << q;
  const uint64_t g0 = N0 ^ (N0 >> 1);
#pragma unroll
  for (int l = 0; l < NY; ++l) y[l] = A[(size_t)(R - 1) * NYT + wbase + l];
  if constexpr (TWO) {
#pragma unroll
    for (int l = 0; l < NY; ++l) z[l] = 0.0;
  }
  for (int i = 0; i < R - 1; ++i) {
    const bool gi = (g0 >> i) & 1;
    const cdouble_t *row = A + (size_t)i * NYT + wbase;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const double a = row[l];
      if constexpr (TWO) {
        z[l] += gi ? a : 0.0;
        y[l] += gi ? 0.0 : a;
      } else {
        y[l] += gi ? a : -a;
      }
    }
  }

  const double dtarget = (double)target;
  double best = -__builtin_inf();
  uint32_t first = 0xFFFFFFFFu;
  const uint32_t nsteps = 1u << q;
  const uint32_t nu = par << q;
  for (uint32_t t0 = 0; t0 < nsteps; t0 += K) {
    const uint32_t glen = nsteps - t0 < (uint32_t)K ? nsteps - t0 : (uint32_t)K;   // workgroup-uniform
#pragma unroll 1
    for (uint32_t b = 0; b < glen; ++b) {
      const uint32_t t = t0 + b;
      if (t > 0) {
        const int j = __builtin_ctz(t);
        const uint32_t up = (((nu | t) >> (j + 1)) & 1u) ^ 1u;
        const cdouble_t *dp = Dt + (size_t)(up * (uint32_t)(R - 1) + (uint32_t)j) * NYT + wbase;
#pragma unroll
        for (int l = 0; l < NY; ++l) {
          const double dl = dp[l];
          if constexpr (TWO) {
            z[l] += dl;
            y[l] -= dl;
          } else {
            y[l] += dl;
          }
        }
      }
      double acc;
      if constexpr (KIND == KIND_LM) {
        // column 0 enters without |.| (L_OpenMP.c:40-41); it is wave 0's first
        acc = mw ? y[0] : __builtin_fabs(y[0]);
#pragma unroll
        for (int l = 1; l < NY; ++l) acc += __builtin_fabs(y[l]);
      } else {
        acc = 0.0;
#pragma unroll
        for (int l = 0; l < NY; ++l) {
          acc += __builtin_fabs(y[l]);
          if constexpr (TWO) acc += __builtin_fabs(z[l]);
        }
      }
      part[((size_t)wv * K + b) * 64 + lane] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int bb = 0; bb < K / W; ++bb) {
      const uint32_t b = (uint32_t)(wv * (K / W) + bb);
      if (b < glen) {
        double tot = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < W; ++w2) tot += part[((size_t)w2 * K + b) * 64 + lane];
        if constexpr (MODE == MODE_MAX) {
          best = __builtin_fmax(best, tot);
        } else {
          const uint32_t c = (tot == dtarget) ? t0 + b : 0xFFFFFFFFu;
          first = c < first ? c : first;
        }
      }
    }
    __syncthreads();
  }

  // combine the waves' results per lane segment (wave 0 holds step 0's, so
  // its maximum is finite)
  fin_best[wv * 64 + lane] = best;
  fin_first[wv * 64 + lane] = first;
  __syncthreads();
  if (wv == 0) {
    double m = fin_best[lane];
    uint32_t f = fin_first[lane];
#pragma unroll
    for (int w2 = 1; w2 < W; ++w2) {
      m = __builtin_fmax(m, fin_best[w2 * 64 + lane]);
      f = fin_first[w2 * 64 + lane] < f ? fin_first[w2 * 64 + lane] : f;
    }
    finish64<MODE>(active, local, (int64_t)m, (seg << q) + f, f != 0xFFFFFFFFu, unit_best, gbest, gfirst);
  }
}

// ---------------------------------------------------------------------------
// d-ary family: reflected d-ary Gray code (L_OpenMP.c:126-239); the step
// counter and direction logic of the int32 walk, with 64-bit label vectors:
// V = uint64_t (two's-complement adds, |.| of the signed value) or, while
// sum |A| < 2^52, V = double (every value an exact integer below 2^53, as in
// k_plainf64_bin; |.| a source modifier of the add).
// ---------------------------------------------------------------------------
template <class V>
struct Num64;
template <>
struct Num64<uint64_t> {
  typedef int64_t Sum;
  typedef cint64_t CV;
  __device__ __forceinline__ static Sum mag(uint64_t x) { return abs64((int64_t)x); }
  __device__ __forceinline__ static Sum max(Sum a, Sum b) { return a > b ? a : b; }
  __device__ __forceinline__ static uint64_t masked(uint64_t v, uint64_t m) { return v & m; }
};
template <>
struct Num64<double> {
  typedef double Sum;
  typedef cdouble_t CV;
  __device__ __forceinline__ static Sum mag(double x) { return __builtin_fabs(x); }
  __device__ __forceinline__ static Sum max(Sum a, Sum b) { return __builtin_fmax(a, b); }
  // v or +0.0, by the bits (exact)
  __device__ __forceinline__ static double masked(double v, uint64_t m) {
    return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) & m);
  }
};

template <int F, int T, int D, int NY, class V>
__device__ __forceinline__ void apply_move64(V (&t)[D][NY], typename Num64<V>::Sum (&S)[D],
                                             const typename Num64<V>::CV *row) {
  typename Num64<V>::Sum sf = 0, st = 0;
#pragma unroll
  for (int l = 0; l < NY; ++l) {
    const V a = row[l];
    t[F][l] -= a;
    t[T][l] += a;
    sf += Num64<V>::mag(t[F][l]);
    st += Num64<V>::mag(t[T][l]);
  }
  S[F] = sf;
  S[T] = st;
}

template <int C, int D, int NY, class V>
struct Mover64 {
  __device__ __forceinline__ static void run(V (&t)[D][NY], typename Num64<V>::Sum (&S)[D],
                                             const typename Num64<V>::CV *row, int from, int up) {
    if (from == C) {
      if (up) {
        if constexpr (C + 1 < D) apply_move64<C, C + 1, D, NY, V>(t, S, row);
      } else {
        if constexpr (C > 0) apply_move64<C, C - 1, D, NY, V>(t, S, row);
      }
    } else {
      if constexpr (C + 1 < D) Mover64<C + 1, D, NY, V>::run(t, S, row, from, up);
    }
  }
};

template <int K, int D, int NY, bool UP, class V>
struct Digit0Run64 {
  template <class Eval>
  __device__ __forceinline__ static void run(V (&t)[D][NY], typename Num64<V>::Sum (&S)[D],
                                             const typename Num64<V>::CV *row0, uint32_t &tt, Eval &eval) {
    if constexpr (K < D) {
      constexpr int F = UP ? K - 1 : D - K;
      constexpr int T = UP ? K : D - K - 1;
      apply_move64<F, T, D, NY, V>(t, S, row0);
      eval(++tt);
      Digit0Run64<K + 1, D, NY, UP, V>::run(t, S, row0, tt, eval);
    }
  }
};

template <int D, int NY, int MODE, class V>
__global__ __launch_bounds__(256) void k_plain64_dary(const V *__restrict__ A_, int R, int q, uint64_t seg0,
                                                      uint64_t nseg, int64_t target, int64_t *__restrict__ unit_best,
                                                      long long *__restrict__ gbest,
                                                      unsigned long long *__restrict__ gfirst) {
  typedef typename Num64<V>::Sum Sum;
  typedef typename Num64<V>::CV CV;
  const CV *A = (const CV *)(uintptr_t)A_;
  const uint64_t w = global_wave();
  const int lane = wave_lane();
  if (((w >> 1) << 7) >= nseg) return;
  const uint64_t local = lane_segment(w, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ w) & 1));

  uint32_t par_mask;
  if (D & 1) par_mask = pseg ? ((1u << q) - 1u) : 0u;
  else par_mask = (q > 0 && pseg) ? (1u << (q - 1)) : 0u;

  V t[D][NY];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int l = 0; l < NY; ++l) t[c][l] = c == 0 ? (V)A[(size_t)(R - 1) * NY + l] : (V)0;
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
    // per-label lane masks, as seed_add32 (a lane-varying select per state
    // word doubles the registers)
    const CV *row = A + (size_t)i * NY;
    uint64_t mask[D];
#pragma unroll
    for (int c = 0; c < D; ++c) mask[c] = (lab == c) ? ~0ull : 0ull;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const V v = row[l];
#pragma unroll
      for (int c = 0; c < D; ++c) t[c][l] += Num64<V>::masked(v, mask[c]);
    }
  }
  Sum S[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    Sum s = 0;
#pragma unroll
    for (int l = 0; l < NY; ++l) s += Num64<V>::mag(t[c][l]);
    S[c] = s;
  }

  const Sum starget = (Sum)target;
  Sum best = 0;   // every objective is a sum of |.|
  uint32_t first = 0xFFFFFFFFu;
  auto eval = [&](uint32_t tt) {
    Sum acc = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) acc += S[c];
    if constexpr (MODE == MODE_MAX) {
      best = Num64<V>::max(best, acc);
    } else {
      const uint32_t cnd = (acc == starget) ? tt : 0xFFFFFFFFu;
      first = cnd < first ? cnd : first;
    }
  };

  // blocks of d steps: a dispatched move of digit j >= 1, then the d-1
  // moves of row 0 unrolled (the counter covers digits 1..q-1)
  uint64_t biaspat = 0;
  for (int i = 1; i < q; ++i) biaspat |= (uint64_t)(16 - D) << (4 * (i - 1));
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;
  uint32_t nblocks = 1;
  for (int i = 1; i < q; ++i) nblocks *= (uint32_t)D;

  eval(0);
  uint32_t tt = 0;
  if (q > 0) {
    if (pm & 1u) Digit0Run64<1, D, NY, false, V>::run(t, S, A, tt, eval);
    else Digit0Run64<1, D, NY, true, V>::run(t, S, A, tt, eval);
    for (uint32_t b = 1; b < nblocks; ++b) {
      const int jj = __builtin_ctzll(~cnt) >> 2;
      const uint64_t low = (1ull << (4 * jj)) - 1ull;
      cnt = (cnt + 1ull) | (biaspat & low);
      const int v = (int)((cnt >> (4 * jj)) & 15ull) - (16 - D);
      const int j = jj + 1;
      const int p = (pm >> j) & 1;
      const int from = p ? D - v : v - 1;
      pm ^= (1u << j) - 1u;
      Mover64<0, D, NY, V>::run(t, S, A + (size_t)j * NY, from, p ? 0 : 1);
      eval(++tt);
      if (pm & 1u) Digit0Run64<1, D, NY, false, V>::run(t, S, A, tt, eval);
      else Digit0Run64<1, D, NY, true, V>::run(t, S, A, tt, eval);
    }
  }
  const uint32_t nsteps = q > 0 ? nblocks * (uint32_t)D : 1u;
  finish64<MODE>(active, local, (int64_t)best, seg * (uint64_t)nsteps + first, first != 0xFFFFFFFFu, unit_best,
                 gbest, gfirst);
}

// ---------------------------------------------------------------------------
// The d-ary walk in doubles over W waves, beyond one wave's d * NY <= 80
// doubles (sum |A| < 2^52): wave wv keeps columns [wv*NY, (wv+1)*NY) of every
// label sum and its labels' partial |.| sums S.  The step counter and the
// moves are workgroup-uniform, as in k_plain64_dary; each step's partial
// objective (sum of S) goes to LDS, and after every K steps (or the
// segment's end) each wave totals K/W of them over the W partials.
// ---------------------------------------------------------------------------
constexpr int kWideDaryK = 8;

template <int D, int NY, int W, int MODE>
__global__ __launch_bounds__(64 * W) void k_widef64_dary(const double *__restrict__ A_, int R, int q, uint64_t seg0,
                                                         uint64_t nseg, int64_t target,
                                                         int64_t *__restrict__ unit_best,
                                                         long long *__restrict__ gbest,
                                                         unsigned long long *__restrict__ gfirst) {
  constexpr int NYT = NY * W;
  constexpr int K = kWideDaryK;
  static_assert(K % W == 0, "steps per combine must split evenly over the waves");
  __shared__ double part[W * K * 64];
  __shared__ double fin_best[W * 64];
  __shared__ uint32_t fin_first[W * 64];
  const cdouble_t *A = (const cdouble_t *)(uintptr_t)A_;
  const int lane = wave_lane();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t g = blockIdx.x;
  if (((g >> 1) << 7) >= nseg) return;   // uniform over the workgroup
  const uint64_t local = lane_segment(g, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ g) & 1));
  const int wbase = wv * NY;
  const cdouble_t *Aw = A + wbase;   // this wave's columns of row 0; row i at Aw + i * NYT

  uint32_t par_mask;
  if (D & 1) par_mask = pseg ? ((1u << q) - 1u) : 0u;
  else par_mask = (q > 0 && pseg) ? (1u << (q - 1)) : 0u;

  double t[D][NY];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int l = 0; l < NY; ++l) t[c][l] = c == 0 ? Aw[(size_t)(R - 1) * NYT + l] : 0.0;
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
    const cdouble_t *row = Aw + (size_t)i * NYT;
    uint64_t mask[D];
#pragma unroll
    for (int c = 0; c < D; ++c) mask[c] = (lab == c) ? ~0ull : 0ull;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const double v = row[l];
#pragma unroll
      for (int c = 0; c < D; ++c) t[c][l] += Num64<double>::masked(v, mask[c]);
    }
  }
  double S[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    double sc = 0.0;
#pragma unroll
    for (int l = 0; l < NY; ++l) sc += __builtin_fabs(t[c][l]);
    S[c] = sc;
  }

  uint32_t nblocks = 1;
  for (int i = 1; i < q; ++i) nblocks *= (uint32_t)D;
  const uint32_t nsteps = q > 0 ? nblocks * (uint32_t)D : 1u;
  const double dtarget = (double)target;
  double best = -__builtin_inf();
  uint32_t first = 0xFFFFFFFFu;
  // step tt's partial objective into LDS; after every K steps, or the last
  // one, each wave totals its share (the step index is workgroup-uniform)
  auto eval = [&](uint32_t tt) {
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) acc += S[c];
    const uint32_t b = tt % K;
    part[((size_t)wv * K + b) * 64 + lane] = acc;
    if (b == K - 1 || tt == nsteps - 1) {
      const uint32_t t0 = tt - b, glen = b + 1;
      __syncthreads();
#pragma unroll
      for (int bb = 0; bb < K / W; ++bb) {
        const uint32_t e = (uint32_t)(wv * (K / W) + bb);
        if (e < glen) {
          double tot = 0.0;
#pragma unroll
          for (int w2 = 0; w2 < W; ++w2) tot += part[((size_t)w2 * K + e) * 64 + lane];
          if constexpr (MODE == MODE_MAX) {
            best = __builtin_fmax(best, tot);
          } else {
            const uint32_t c = (tot == dtarget) ? t0 + e : 0xFFFFFFFFu;
            first = c < first ? c : first;
          }
        }
      }
      __syncthreads();
    }
  };

  uint64_t biaspat = 0;
  for (int i = 1; i < q; ++i) biaspat |= (uint64_t)(16 - D) << (4 * (i - 1));
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;

  eval(0);
  uint32_t tt = 0;
  if (q > 0) {
    if (pm & 1u) Digit0Run64<1, D, NY, false, double>::run(t, S, Aw, tt, eval);
    else Digit0Run64<1, D, NY, true, double>::run(t, S, Aw, tt, eval);
    for (uint32_t b = 1; b < nblocks; ++b) {
      const int jj = __builtin_ctzll(~cnt) >> 2;
      const uint64_t low = (1ull << (4 * jj)) - 1ull;
      cnt = (cnt + 1ull) | (biaspat & low);
      const int v = (int)((cnt >> (4 * jj)) & 15ull) - (16 - D);
      const int j = jj + 1;
      const int p = (pm >> j) & 1;
      const int from = p ? D - v : v - 1;
      pm ^= (1u << j) - 1u;
      Mover64<0, D, NY, double>::run(t, S, Aw + (size_t)j * NYT, from, p ? 0 : 1);
      eval(++tt);
      if (pm & 1u) Digit0Run64<1, D, NY, false, double>::run(t, S, Aw, tt, eval);
      else Digit0Run64<1, D, NY, true, double>::run(t, S, Aw, tt, eval);
    }
  }

  fin_best[wv * 64 + lane] = best;
  fin_first[wv * 64 + lane] = first;
  __syncthreads();
  if (wv == 0) {
    double m = fin_best[lane];
    uint32_t f = fin_first[lane];
#pragma unroll
    for (int w2 = 1; w2 < W; ++w2) {
      m = __builtin_fmax(m, fin_best[w2 * 64 + lane]);
      f = fin_first[w2 * 64 + lane] < f ? fin_first[w2 * 64 + lane] : f;
    }
    finish64<MODE>(active, local, (int64_t)m, seg * (uint64_t)nsteps + f, f != 0xFFFFFFFFu, unit_best, gbest,
                   gfirst);
  }
}

// ---------------------------------------------------------------------------
// Any width: one wave per segment, lanes own columns, label vectors in a
// per-wave global scratch slice (the 64-bit twin of k_general).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int gray_label64(uint64_t N, int i, int d) {
  if (d <= 2) return (int)(((N ^ (N >> 1)) >> i) & 1);
  uint64_t x = N;
  for (int k = 0; k < i; ++k) x /= (uint64_t)d;
  const uint32_t r = (uint32_t)(x % (uint64_t)(2 * d));
  return r < (uint32_t)d ? (int)r : (int)(2 * d - 1 - r);
}

template <int MODE>
__global__ __launch_bounds__(64) void k_general64(const int64_t *__restrict__ A, int R, int C, int kind, int d,
                                                  uint64_t seglen, uint64_t seg0, uint64_t nseg,
                                                  uint64_t *__restrict__ state, int64_t target,
                                                  int64_t *__restrict__ unit_best, long long *__restrict__ gbest,
                                                  unsigned long long *__restrict__ gfirst) {
  const int lane = threadIdx.x & 63;
  const bool l1 = (kind == KIND_L1 || kind == KIND_LM);
  const int nvec = l1 ? 1 : d;
  uint64_t *st = state + (size_t)blockIdx.x * (size_t)nvec * (size_t)C;
  int64_t S[16];

  for (uint64_t sidx = blockIdx.x; sidx < nseg; sidx += gridDim.x) {
    const uint64_t N0 = (seg0 + sidx) * seglen;
    for (int l = lane; l < C; l += 64) {
      const uint64_t last = (uint64_t)A[(size_t)(R - 1) * C + l];
      for (int c = 0; c < nvec; ++c) st[(size_t)c * C + l] = c == 0 ? last : 0ull;
    }
    for (int i = 0; i < R - 1; ++i) {
      const int lab = gray_label64(N0, i, d < 2 ? 2 : d);
      for (int l = lane; l < C; l += 64) {
        const uint64_t a = (uint64_t)A[(size_t)i * C + l];
        if (l1) st[l] += lab ? a : 0ull - a;
        else st[(size_t)lab * C + l] += a;
      }
    }
    auto vec_sum = [&](int c) -> int64_t {
      uint64_t p = 0;
      for (int l = lane; l < C; l += 64) {
        const int64_t y = (int64_t)st[(size_t)c * C + l];
        p += (uint64_t)((kind == KIND_LM && l == 0) ? y : abs64(y));
      }
      return (int64_t)wave_sum_u64(p);
    };
    for (int c = 0; c < nvec; ++c) S[c] = vec_sum(c);

    int64_t best = INT64_MIN;
    uint64_t first = ~0ull;
    auto eval = [&](uint64_t N) {
      int64_t acc = 0;
      for (int c = 0; c < nvec; ++c) acc += S[c];
      if (MODE == MODE_MAX) best = acc > best ? acc : best;
      else if (acc == target && first == ~0ull) first = N;
    };
    eval(N0);
    for (uint64_t t = 1; t < seglen; ++t) {
      const uint64_t N = N0 + t;
      int j, from = 0, to = 0, up = 0;
      if (d <= 2) {
        j = __builtin_ctzll(N);
        up = (int)(((N >> (j + 1)) & 1) ^ 1);
        from = up ? 0 : 1;
        to = up ? 1 : 0;
      } else {
        uint64_t x = N;
        j = 0;
        while (x % (uint64_t)d == 0) { x /= (uint64_t)d; ++j; }
        const int v = (int)(x % (uint64_t)d);
        const int p = (int)((x / (uint64_t)d) & 1);
        from = p ? d - v : v - 1;
        to = p ? d - 1 - v : v;
      }
      const int64_t *row = A + (size_t)j * C;
      if (l1) {
        const uint64_t m = up ? 2ull : (uint64_t)-2;
        for (int l = lane; l < C; l += 64) st[l] += m * (uint64_t)row[l];
        S[0] = vec_sum(0);
      } else {
        for (int l = lane; l < C; l += 64) {
          const uint64_t a = (uint64_t)row[l];
          st[(size_t)from * C + l] -= a;
          st[(size_t)to * C + l] += a;
        }
        const int64_t sf = vec_sum(from), stt = vec_sum(to);
        for (int c = 0; c < nvec; ++c) {
          if (c == from) S[c] = sf;
          if (c == to) S[c] = stt;
        }
      }
      eval(N);
    }
    // best and first are wave-uniform (built from wave sums): every lane
    // writes the same value, no branch on the lane index
    if (MODE == MODE_MAX) {
      unit_best[sidx] = best;
      atomicMax(gbest, (long long)best);
    } else if (first != ~0ull) {
      atomicMin(gfirst, (unsigned long long)first);
    }
  }
}

// ---------------------------------------------------------------------------
// helpers and launchers
// ---------------------------------------------------------------------------
__global__ void k_fill_i64(long long *p, long long v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ __launch_bounds__(256) void k_first_unit64(const int64_t *__restrict__ ub, uint64_t n,
                                                      const long long *__restrict__ gbest,
                                                      unsigned long long *__restrict__ first) {
  const int64_t target = (int64_t)*gbest;
  unsigned long long mine = ~0ull;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (ub[i] == target) {
      mine = i;
      break;
    }
  const uint64_t m = wave_min_u64(mine);
  if (m != ~0ull) atomicMin(first, (unsigned long long)m);
}

static unsigned grid256(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return (unsigned)(b < 4096 ? (b ? b : 1) : 4096);
}

hipError_t launch_fill_i64(long long *p, long long v, uint64_t n, hipStream_t st) {
  if (!p && n) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill_i64, dim3(grid256(n)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

hipError_t launch_first_unit64(const int64_t *ub, uint64_t n, const long long *gbest, unsigned long long *first,
                               hipStream_t st) {
  if (!ub || !gbest || !first) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_first_unit64, dim3(grid256(n)), dim3(256), 0, st, ub, n, gbest, first);
  return hipGetLastError();
}

template <int KIND, int NY>
static hipError_t l64_bin(Mode mode, const Plain64Args &a, hipStream_t st) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)((waves + 3) / 4)), blk(256);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_plain64_bin<KIND, NY, MODE_MAX>), grid, blk, 0, st, a.A, a.D, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_plain64_bin<KIND, NY, MODE_FIRST>), grid, blk, 0, st, a.A, a.D, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

template <int KIND, int NY>
static hipError_t lf64_bin(Mode mode, const Plain64Args &a, hipStream_t st) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)((waves + 3) / 4)), blk(256);
  const double *A = (const double *)a.A, *D = (const double *)a.D;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_plainf64_bin<KIND, NY, MODE_MAX>), grid, blk, 0, st, A, D, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_plainf64_bin<KIND, NY, MODE_FIRST>), grid, blk, 0, st, A, D, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

template <int D, int NY, class V>
static hipError_t l64_dary_v(Mode mode, const Plain64Args &a, hipStream_t st) {
  const uint64_t waves = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)((waves + 3) / 4)), blk(256);
  const V *A = (const V *)a.A;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_plain64_dary<D, NY, MODE_MAX, V>), grid, blk, 0, st, A, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_plain64_dary<D, NY, MODE_FIRST, V>), grid, blk, 0, st, A, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

template <int D, int NY>
static hipError_t l64_dary(Mode mode, const Plain64Args &a, hipStream_t st) {
  return a.fp64 ? l64_dary_v<D, NY, double>(mode, a, st) : l64_dary_v<D, NY, uint64_t>(mode, a, st);
}

int plain64_ny_pad(Kind kind, int d, int ny) {
  int mx;
  if (kind == KIND_L1 || kind == KIND_LM) mx = 64;
  else if (kind == KIND_L2) mx = 32;
  else mx = d == 3 ? 24 : (d == 4 ? 16 : 8);
  for (int w = 8; w <= mx; w += 8) {
    if ((kind == KIND_L1 || kind == KIND_LM) && w == 56) continue;   // not instantiated (spills); 64 covers it
    if (w >= ny) return w;
  }
  return 0;
}

template <int D>
static hipError_t l64_dary_w(Mode mode, const Plain64Args &a, hipStream_t st) {
  switch (a.ny_pad) {
    case 8: return l64_dary<D, 8>(mode, a, st);
    case 16:
      if constexpr (D <= 4) return l64_dary<D, 16>(mode, a, st);
      break;
    case 24:
      if constexpr (D <= 3) return l64_dary<D, 24>(mode, a, st);
      break;
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_plain64(Kind kind, Mode mode, const Plain64Args &a, hipStream_t st) {
  if (!a.A || (kind != KIND_LD && !a.D) || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (plain64_ny_pad(kind, a.d, a.ny_pad) != a.ny_pad) return hipErrorInvalidValue;
  if (a.fp64) {   // the double-precision binary walk (A and D hold doubles)
    switch (kind) {
      case KIND_L1:
        switch (a.ny_pad) {
#define X(n) case n: return lf64_bin<KIND_L1, n>(mode, a, st);
          X(8) X(16) X(24) X(32) X(40) X(48) X(64)
#undef X
        }
        break;
      case KIND_LM:
        switch (a.ny_pad) {
#define X(n) case n: return lf64_bin<KIND_LM, n>(mode, a, st);
          X(8) X(16) X(24) X(32) X(40) X(48) X(64)
#undef X
        }
        break;
      case KIND_L2:
        switch (a.ny_pad) {
#define X(n) case n: return lf64_bin<KIND_L2, n>(mode, a, st);
          X(8) X(16) X(24) X(32)
#undef X
        }
        break;
      default: break;   // d >= 3: the d-ary walk below, in doubles
    }
    if (kind != KIND_LD) return hipErrorInvalidValue;
  }
  switch (kind) {
    case KIND_L1:
      switch (a.ny_pad) {
#define X(n) case n: return l64_bin<KIND_L1, n>(mode, a, st);
        X(8) X(16) X(24) X(32) X(40) X(48) X(64)
#undef X
      }
      break;
    case KIND_LM:
      switch (a.ny_pad) {
#define X(n) case n: return l64_bin<KIND_LM, n>(mode, a, st);
        X(8) X(16) X(24) X(32) X(40) X(48) X(64)
#undef X
      }
      break;
    case KIND_L2:
      switch (a.ny_pad) {
#define X(n) case n: return l64_bin<KIND_L2, n>(mode, a, st);
        X(8) X(16) X(24) X(32)
#undef X
      }
      break;
    default:
      switch (a.d) {
        case 3: return l64_dary_w<3>(mode, a, st);
        case 4: return l64_dary_w<4>(mode, a, st);
        case 5: return l64_dary_w<5>(mode, a, st);
        case 6: return l64_dary_w<6>(mode, a, st);
        case 7: return l64_dary_w<7>(mode, a, st);
        case 8: return l64_dary_w<8>(mode, a, st);
        case 9: return l64_dary_w<9>(mode, a, st);
        case 10: return l64_dary_w<10>(mode, a, st);
        default: break;
      }
  }
  return hipErrorInvalidValue;
}

template <int KIND, int NY, int W>
static hipError_t lwf64(Mode mode, const Plain64Args &a, hipStream_t st) {
  const uint64_t groups = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)groups), blk(64 * W);
  const double *A = (const double *)a.A, *D = (const double *)a.D;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_widef64_bin<KIND, NY, W, MODE_MAX>), grid, blk, 0, st, A, D, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_widef64_bin<KIND, NY, W, MODE_FIRST>), grid, blk, 0, st, A, D, a.R, a.q, a.seg0,
                       a.nseg, a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

template <int D, int NY, int W>
static hipError_t lwf64d(Mode mode, const Plain64Args &a, hipStream_t st) {
  const uint64_t groups = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)groups), blk(64 * W);
  const double *A = (const double *)a.A;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_widef64_dary<D, NY, W, MODE_MAX>), grid, blk, 0, st, A, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_widef64_dary<D, NY, W, MODE_FIRST>), grid, blk, 0, st, A, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

// doubles per wave of the d-ary W-wave walk (d * NY <= 80, as the one-wave walk)
constexpr int widef64_dary_nw(int d) { return d == 3 ? 24 : d == 4 ? 16 : 8; }

template <int D>
static hipError_t lwf64d_w(Mode mode, const Plain64Args &a, hipStream_t st) {
  constexpr int NY = widef64_dary_nw(D);
  if (a.ny_pad == 2 * NY) return lwf64d<D, NY, 2>(mode, a, st);
  if (a.ny_pad == 4 * NY) return lwf64d<D, NY, 4>(mode, a, st);
  if (a.ny_pad == 8 * NY) return lwf64d<D, NY, 8>(mode, a, st);
  return hipErrorInvalidValue;
}

// Widths of the W-wave walks in doubles: L1 / LM 128, 256 or 512 columns (4
// or 8 waves of 32 doubles, 8 of 64), L2 128 or 256 (4 or 8 waves of 32
// pairs); d = 3..10 2, 4 or 8 waves of 24 (d = 3), 16 (d = 4) or 8 columns.
int widef64_ny_pad(Kind kind, int d, int ny) {
  if (kind == KIND_LD) {
    if (d < 3 || d > 10) return 0;
    for (int w : {2, 4, 8})
      if (w * widef64_dary_nw(d) >= ny) return w * widef64_dary_nw(d);
    return 0;
  }
  for (int w : {128, 256, 512})
    if (w >= ny && (kind != KIND_L2 || w <= 256)) return w;
  return 0;
}

hipError_t launch_widef64(Kind kind, Mode mode, const Plain64Args &a, hipStream_t st) {
  if (!a.A || (kind != KIND_LD && !a.D) || !a.unit_best || !a.gbest || !a.gfirst || !a.fp64)
    return hipErrorInvalidValue;
  if (widef64_ny_pad(kind, a.d, a.ny_pad) != a.ny_pad) return hipErrorInvalidValue;
  switch (kind) {
    case KIND_LD:
      switch (a.d) {
        case 3: return lwf64d_w<3>(mode, a, st);
        case 4: return lwf64d_w<4>(mode, a, st);
        case 5: return lwf64d_w<5>(mode, a, st);
        case 6: return lwf64d_w<6>(mode, a, st);
        case 7: return lwf64d_w<7>(mode, a, st);
        case 8: return lwf64d_w<8>(mode, a, st);
        case 9: return lwf64d_w<9>(mode, a, st);
        case 10: return lwf64d_w<10>(mode, a, st);
        default: return hipErrorInvalidValue;
      }
    case KIND_L1:
      if (a.ny_pad == 128) return lwf64<KIND_L1, 32, 4>(mode, a, st);
      if (a.ny_pad == 256) return lwf64<KIND_L1, 32, 8>(mode, a, st);
      return lwf64<KIND_L1, 64, 8>(mode, a, st);
    case KIND_LM:
      if (a.ny_pad == 128) return lwf64<KIND_LM, 32, 4>(mode, a, st);
      if (a.ny_pad == 256) return lwf64<KIND_LM, 32, 8>(mode, a, st);
      return lwf64<KIND_LM, 64, 8>(mode, a, st);
    case KIND_L2:
      if (a.ny_pad == 128) return lwf64<KIND_L2, 32, 4>(mode, a, st);
      return lwf64<KIND_L2, 32, 8>(mode, a, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_general64(Kind kind, Mode mode, const Plain64Args &a, int ny, hipStream_t st) {
  if (!a.A || !a.state || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  uint64_t seglen = 1;
  const int base = a.d < 2 ? 2 : a.d;
  for (int i = 0; i < a.q; ++i) seglen *= (uint64_t)base;
  const uint64_t grid = a.nseg < kGeneralSlices ? a.nseg : kGeneralSlices;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL(k_general64<MODE_MAX>, dim3((unsigned)grid), dim3(64), 0, st, a.A, a.R, ny, (int)kind, a.d,
                       seglen, a.seg0, a.nseg, a.state, a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL(k_general64<MODE_FIRST>, dim3((unsigned)grid), dim3(64), 0, st, a.A, a.R, ny, (int)kind, a.d,
                       seglen, a.seg0, a.nseg, a.state, a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

}  // namespace lnorm

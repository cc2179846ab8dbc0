// lnorm_blocked.hip -- the blocked L_1 / L_M enumeration on gfx950.
//
// The enumerated rows are split into kl low rows and nh high rows (row R-1
// keeps its fixed +1).  For every one of the 2^kl sign patterns p of the low
// rows, Y_L[:,p] = M_lo^T x_p is precomputed once by an int8 MFMA GEMM
// (k_table_mfma: exact int32 accumulation, entries beyond +-127 split into
// hi/lo int8 planes).  A wave then owns 2^s consecutive ranks of the Gray
// code over the high rows; the high partial sums y_H are the same for all its
// lanes, so they live in SGPRs and each high step is a handful of scalar adds.
// Lane L holds P low patterns (p = r*64 + L) in VGPRs.  With both halves
// biased, one v_sad_u16 evaluates |Y_L[l,p] + y_H[l]| for two columns and
// accumulates:
//   table word  = (Y_L[2c+1] + 2^15) << 16 | (Y_L[2c] + 2^15)
//   scalar word = (2^15 - y_H[2c+1]) << 16 | (2^15 - y_H[2c])
//   |table.lo - scalar.lo| = |Y_L[2c] + y_H[2c]| = |y[2c]|.
// The host guarantees every column's absolute sum is <= 2^15 - 1, so all
// halves stay in [0, 2^16) and a 32-bit add of a packed delta updates both
// halves exactly.  For L_M, column 0 is carried as one extra int32 word in
// both the table and the scalar state and enters the sum without |.|.
//
// Every wave evaluates ranks N = N_high * 2^kl + N_low for all N_low, i.e.
// the contiguous range [w 2^(s+kl), (w+1) 2^(s+kl)): the BRGC of N restricted
// to the high rows is the BRGC of N_high.  The wave's maximum goes to
// one 64-bit atomic maximum per launch (blocked_key: the value above the
// complement of the wave index), which leaves the launch's maximum and the
// first wave attaining it; the smallest maximising rank is recovered
// afterwards by the plain kernel in MODE_FIRST on that wave's range.
#include "lnorm_device.h"

namespace lnorm {

template <int NP, int P, int MARG>
__global__ __launch_bounds__(256) void k_blocked16(
    const uint32_t *__restrict__ table, const uint32_t *__restrict__ dpk,
    const uint32_t *__restrict__ seed, const uint32_t *__restrict__ dfull, int nh, int s,
    uint64_t wave0, uint64_t nwaves, unsigned long long *__restrict__ gkey) {
  constexpr int NW = NP + MARG;
  const int lane = wave_lane();
  const uint64_t wl = global_wave();
  if (wl >= nwaves) return;
  const uint64_t w = wave0 + wl;

  uint32_t tab[P][NW];
#pragma unroll
  for (int r = 0; r < P; ++r)
#pragma unroll
    for (int c = 0; c < NW; ++c) tab[r][c] = table[((size_t)r * NW + c) * 64 + lane];

  // scalar state at high rank h0 = w * 2^s: start from all high rows at -1
  // and add the flip of every row whose Gray bit is set.
  const uint64_t h0 = w << s;
  const uint64_t g0 = h0 ^ (h0 >> 1);
  uint32_t v = 0;
  if (lane < NW) {
    v = seed[lane];
    for (int m = 0; m < nh; ++m)
      if ((g0 >> m) & 1) v += dfull[(size_t)m * NW + lane];
  }
  uint32_t S[NW];
#pragma unroll
  for (int c = 0; c < NW; ++c) S[c] = (uint32_t)__builtin_amdgcn_readlane((int)v, c);

  // one running maximum over the lane's P patterns: two patterns' sums and
  // the maximum meet in one v_max3_i32, half an instruction per sign vector
  int32_t best = INT32_MIN;
  auto eval = [&]() {
    int32_t a[P];
#pragma unroll
    for (int r = 0; r < P; ++r) {
      uint32_t acc = MARG ? tab[r][NP] + S[NP] : 0u;
#pragma unroll
      for (int c = 0; c < NP; ++c) acc = sad16(tab[r][c], S[c], acc);
      a[r] = (int32_t)acc;
    }
#pragma unroll
    for (int r = 0; r < P; r += 2) {
      const int32_t y = (r + 1 < P) ? a[r + 1] : a[r];
      const int32_t xy = a[r] > y ? a[r] : y;
      best = xy > best ? xy : best;
    }
  };

  const uint32_t nsteps = 1u << s;
  const uint32_t nu = (uint32_t)(w & 1) << s;   // bit s of N_high
  // the packed delta of step t: row j = ctz t flips to 1 ^ (bit j+1 of N_high)
  auto delta = [&](uint32_t t) {
    const int j = __builtin_ctz(t);
    const uint32_t up = (((nu | t) >> (j + 1)) & 1u) ^ 1u;
    return dpk + ((size_t)up * (uint32_t)nh + (uint32_t)j) * NW;
  };
  // each step's delta address comes from the step before, and a scheduling
  // barrier keeps its scalar loads at the top of the iteration: their latency
  // hides behind eval() instead of stalling the wave just before the adds
  const uint32_t *dp = delta(1);
  for (uint32_t t = 1; t < nsteps; ++t) {
    uint32_t dl[NW];
#pragma unroll
    for (int c = 0; c < NW; ++c) dl[c] = dp[c];
    dp = delta(t + 1 < nsteps ? t + 1 : 1);
    __builtin_amdgcn_sched_barrier(0);
    eval();                      // state of rank t-1
#pragma unroll
    for (int c = 0; c < NW; ++c) S[c] += dl[c];
  }
  eval();                        // state of rank nsteps-1

  const int32_t m = wave_max_i32(best);
  // the key only grows, so one that is not above the current key cannot
  // change it: most waves skip the atomic (each one would write to memory)
  const unsigned long long key = blocked_key(m, wl);
  if (lane == 0 && key > *(volatile unsigned long long *)gkey) atomicMax(gkey, key);
}

// ---------------------------------------------------------------------------
// Low-pattern table by int8 MFMA: D = Alo^T (M = output columns) x X (N =
// patterns), K = low rows (kl <= 16, zero padded to 32).  One wave per
// 32 x 32 output tile, v_mfma_i32_32x32x32_i8.  Output element (m, n) sits in
// lane n + 32*g, accumulator r with m = 8*(r/4) + 4*g + (r%4), so each lane
// owns column pairs (m, m+1) of one pattern and packs them locally.
// Column order for L_M: m < np2 are the |.| columns 1..ny-1 (zero padded up to
// np2 = 2*np), m = np2 is the marginal column 0.
// ---------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(64) void k_table_mfma(const int8_t *__restrict__ lo, int planes, int kl,
                                                   int ny, int marginal, int np, int P, int wide,
                                                   uint32_t *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 5, i32 = lane & 31;
  const int mtile = blockIdx.y, ptile = blockIdx.x;
  const int m = mtile * 32 + i32;        // output column slot of this lane's A row
  const int n = ptile * 32 + i32;        // pattern of this lane's B column
  // |.| slots: 2 per packed word, or 1 per word in the 32-bit (wide) layout
  const int np2 = wide ? np : 2 * np;
  // matrix column feeding output slot m
  int col = -1;
  if (marginal) {
    if (m < np2) col = (m + 1 < ny) ? m + 1 : -1;
    else if (m == np2) col = 0;
  } else {
    col = (m < ny && m < np2) ? m : -1;
  }
  auto plane = [&](int pl) -> v16i {
    int8_t a8[16], b8[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const int k = g * 16 + b;
      a8[b] = (k < kl && col >= 0) ? lo[((size_t)pl * kl + k) * ny + col] : (int8_t)0;
      b8[b] = (k < kl) ? (int8_t)(((n >> k) & 1) ? 1 : -1) : (int8_t)0;
    }
    v4i av, bv;
    __builtin_memcpy(&av, a8, 16);
    __builtin_memcpy(&bv, b8, 16);
    v16i z = {0};
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, z, 0, 0, 0);
  };
  // planes == 2: plane 0 holds the high bytes, plane 1 the low bytes
  const v16i acc0 = plane(0);
  v16i acc1 = {0};
  if (planes == 2) acc1 = plane(1);
  int32_t d[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) d[r] = planes == 2 ? (int32_t)((uint32_t)acc0[r] << 8) + acc1[r] : acc0[r];

  const int NW = np + (marginal ? 1 : 0);
  const int rr = n >> 6, L = n & 63;     // pattern n -> (register block, lane)
  if (rr >= P) return;
  if (wide) {
    // one word per column: Y + 2^31 for |.| columns, Y for the marginal column
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int mm = mtile * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
      if (mm < np2) out[((size_t)rr * NW + mm) * 64 + L] = (uint32_t)d[r] + kBias32;
      else if (marginal && mm == np2) out[((size_t)rr * NW + np) * 64 + L] = (uint32_t)d[r];
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < 16; r += 2) {
    const int mm = mtile * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
    if (mm < np2) {
      const uint32_t w = ((uint32_t)(d[r + 1] + 32768) << 16) | ((uint32_t)(d[r] + 32768) & 0xFFFFu);
      out[((size_t)rr * NW + (mm >> 1)) * 64 + L] = w;
    } else if (marginal && mm == np2) {
      out[((size_t)rr * NW + np) * 64 + L] = (uint32_t)d[r];
    }
  }
}

hipError_t launch_table(const TableArgs &a, hipStream_t st) {
  if (!a.lo || !a.out) return hipErrorInvalidValue;
  const int patterns = 64 * a.P;
  const int mslots = (a.wide ? a.np : 2 * a.np) + (a.marginal ? 1 : 0);
  const dim3 grid((unsigned)(patterns / 32), (unsigned)((mslots + 31) / 32)), blk(64);
  hipLaunchKernelGGL(k_table_mfma, grid, blk, 0, st, a.lo, a.planes, a.kl, a.ny, a.marginal, a.np,
                     a.P, a.wide, a.out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k_blocked32: the same decomposition with 32-bit words, for matrices whose
// column sums exceed 2^15 (e.g. int16 entries).  table word = Y_L + 2^31,
// scalar word = 2^31 - y_H, and one v_sad_u32 gives |y| + acc per column.
// A workgroup of W waves shares the unit (high-rank range) and the low
// patterns; wave w owns |.| words [w*NC, (w+1)*NC) in its SGPRs, so up to
// 4 x 32 = 128 columns fit.  The waves' per-pattern partial sums meet in LDS
// once every B steps, and each wave folds B/W of the steps into its maxima.
// ---------------------------------------------------------------------------
// Three waves per SIMD (<= 168 VGPRs) keep the VALU fed while a wave sits
// at the combine barrier.
template <int NC, int P, int W, int MARG>
__global__ __launch_bounds__(64 * W, 3) void k_blocked32(
    const uint32_t *__restrict__ table, const uint32_t *__restrict__ dpk,
    const uint32_t *__restrict__ seed, const uint32_t *__restrict__ dfull, int nh, int s,
    uint64_t unit0, uint64_t nunits, unsigned long long *__restrict__ gkey, uint32_t B32) {
  constexpr int WNC = W * NC;            // |.| words of the whole group
  constexpr int NWT = WNC + MARG;        // words per pattern (table) and per state
  constexpr int B = 8;                   // steps per LDS combine
  static_assert(B % W == 0, "steps per combine must split evenly over the waves");
  __shared__ uint32_t part[W > 1 ? W * B * 64 * P : 1];
  __shared__ int32_t wbest[W];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t gl = blockIdx.x;
  if (gl >= nunits) return;              // uniform over the whole workgroup
  const uint64_t g = unit0 + gl;
  const int wbase = wv * NC;
  const bool mw = MARG && wv == 0;       // the wave carrying the marginal column

  uint32_t tab[P][NC];
  uint32_t tabm[P];
#pragma unroll
  for (int r = 0; r < P; ++r) {
#pragma unroll
    for (int c = 0; c < NC; ++c) tab[r][c] = table[((size_t)r * NWT + wbase + c) * 64 + lane];
    tabm[r] = mw ? table[((size_t)r * NWT + WNC) * 64 + lane] : 0u;
  }

  const uint64_t h0 = g << s;
  const uint64_t g0 = h0 ^ (h0 >> 1);
  const int myword = lane < NC ? wbase + lane : ((mw && lane == NC) ? WNC : -1);
  uint32_t v = 0;
  if (myword >= 0) {
    v = seed[myword];
    for (int m = 0; m < nh; ++m)
      if ((g0 >> m) & 1) v += dfull[(size_t)m * NWT + myword];
  }
  uint32_t S[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) S[c] = (uint32_t)__builtin_amdgcn_readlane((int)v, c);
  uint32_t Sm = MARG ? (uint32_t)__builtin_amdgcn_readlane((int)v, NC) : 0u;

  int32_t best[P];
#pragma unroll
  for (int r = 0; r < P; ++r) best[r] = INT32_MIN;

  const uint32_t nsteps = 1u << s;       // host guarantees s >= 3 (nsteps % B == 0)
  const uint32_t nu = (uint32_t)(g & 1) << s;
  for (uint32_t t0 = 0; t0 < nsteps; t0 += B) {
    // not unrolled: each step's NC scalar deltas must not stay live across
    // steps, or the state plus several steps' deltas overflow the SGPRs
#pragma unroll 1
    for (int b = 0; b < B; ++b) {
      const uint32_t t = t0 + (uint32_t)b;
      if (t > 0) {
        const int j = __builtin_ctz(t);
        const uint32_t up = (((nu | t) >> (j + 1)) & 1u) ^ 1u;
        const uint32_t *dp = dpk + ((size_t)up * (uint32_t)nh + (uint32_t)j) * NWT;
#pragma unroll
        for (int c = 0; c < NC; ++c) S[c] += dp[wbase + c];
        if (mw) Sm += dp[WNC];
      }
#pragma unroll
      for (int r = 0; r < P; ++r) {
        uint32_t acc = mw ? tabm[r] + Sm : 0u;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc = sad32(tab[r][c], S[c], acc);
        if constexpr (W == 1) best[r] = (int32_t)acc > best[r] ? (int32_t)acc : best[r];
        else part[(((size_t)wv * B + b) * 64 + lane) * P + r] = acc;
      }
    }
    if constexpr (W > 1) {
      __syncthreads();
#pragma unroll
      for (int bb = 0; bb < B / W; ++bb) {
        const int b = wv * (B / W) + bb;
#pragma unroll
        for (int r = 0; r < P; ++r) {
          uint32_t tot = 0;
#pragma unroll
          for (int w2 = 0; w2 < W; ++w2) tot += part[(((size_t)w2 * B + b) * 64 + lane) * P + r];
          best[r] = (int32_t)tot > best[r] ? (int32_t)tot : best[r];
        }
      }
      __syncthreads();
    }
  }

  int32_t bm = best[0];
#pragma unroll
  for (int r = 1; r < P; ++r) bm = best[r] > bm ? best[r] : bm;
  bm = wave_max_i32(bm);
  if constexpr (W > 1) {
    if (lane == 0) wbest[wv] = bm;
    __syncthreads();
    if (wv == 0 && lane == 0) {
      int32_t m = wbest[0];
      for (int w2 = 1; w2 < W; ++w2) m = wbest[w2] > m ? wbest[w2] : m;
      const unsigned long long key = blocked_key(m, gl);
      if (key > *(volatile unsigned long long *)gkey) atomicMax(gkey, key);
    }
  } else if (lane == 0) {
    const unsigned long long key = blocked_key(bm, gl);
    if (key > *(volatile unsigned long long *)gkey) atomicMax(gkey, key);
  }
}

bool blocked32_supported(int nc, int P, int W, int marginal) {
  if (!(nc == 8 || nc == 16 || nc == 24 || nc == 32)) return false;
  if (!(P == 1 || P == 2 || P == 4) || !(W == 1 || W == 2 || W == 4)) return false;
  if (marginal != 0 && marginal != 1) return false;
  return P * nc <= 128;
}

template <int NC, int P, int W, int M>
static hipError_t launch_b32(const BlockedArgs &a, hipStream_t st) {
  hipLaunchKernelGGL((k_blocked32<NC, P, W, M>), dim3((unsigned)a.nwaves), dim3(64 * W), 0, st, a.table, a.dpk,
                     a.seed, a.dfull, a.nh, a.s, a.wave0, a.nwaves, a.gkey, kBias32);
  return hipGetLastError();
}

template <int NC, int P, int M>
static hipError_t launch_b32w(int W, const BlockedArgs &a, hipStream_t st) {
  switch (W) {
    case 1: return launch_b32<NC, P, 1, M>(a, st);
    case 2: return launch_b32<NC, P, 2, M>(a, st);
    case 4: return launch_b32<NC, P, 4, M>(a, st);
  }
  return hipErrorInvalidValue;
}

template <int NC, int M>
static hipError_t launch_b32p(int P, int W, const BlockedArgs &a, hipStream_t st) {
  switch (P) {
    case 1: return launch_b32w<NC, 1, M>(W, a, st);
    case 2: return launch_b32w<NC, 2, M>(W, a, st);
    case 4:
      if constexpr (4 * NC <= 128) return launch_b32w<NC, 4, M>(W, a, st);
      break;
  }
  return hipErrorInvalidValue;
}

template <int M>
static hipError_t launch_b32n(int nc, int P, int W, const BlockedArgs &a, hipStream_t st) {
  switch (nc) {
    case 8: return launch_b32p<8, M>(P, W, a, st);
    case 16: return launch_b32p<16, M>(P, W, a, st);
    case 24: return launch_b32p<24, M>(P, W, a, st);
    case 32: return launch_b32p<32, M>(P, W, a, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_blocked32(int nc, int P, int W, int marginal, const BlockedArgs &a, hipStream_t st) {
  if (!a.table || !a.dpk || !a.seed || !a.dfull || !a.gkey || a.nwaves >= 0xFFFFFFFFull) return hipErrorInvalidValue;
  if (!blocked32_supported(nc, P, W, marginal)) return hipErrorInvalidValue;
  return marginal ? launch_b32n<1>(nc, P, W, a, st) : launch_b32n<0>(nc, P, W, a, st);
}

// ---------------------------------------------------------------------------
// instantiations: NP words of packed columns (<= 64 columns), P patterns per lane
// ---------------------------------------------------------------------------
#define LN_NP(X, P, M) X(4, P, M) X(6, P, M) X(8, P, M) X(10, P, M) X(12, P, M) X(14, P, M) \
  X(16, P, M) X(18, P, M) X(20, P, M) X(22, P, M) X(24, P, M) X(26, P, M) X(28, P, M) X(30, P, M) X(32, P, M)

bool blocked_supported(int np, int P, int marginal) {
  if (!(P == 1 || P == 2 || P == 4 || P == 8)) return false;
  if (marginal != 0 && marginal != 1) return false;
  if (np < 4 || np > 32 || (np & 1)) return false;
  return P * (np + marginal) <= 192;
}

int blocked_np_pad(int np, int P, int marginal) {
  int p = np < 4 ? 4 : (np + 1) & ~1;
  return blocked_supported(p, P, marginal) ? p : 0;
}

template <int NP, int P, int M>
static hipError_t launch_b(const BlockedArgs &a, hipStream_t st) {
  const int block = 256;
  const dim3 grid((unsigned)((a.nwaves + 3) / 4)), blk(block);
  hipLaunchKernelGGL((k_blocked16<NP, P, M>), grid, blk, 0, st, a.table, a.dpk, a.seed, a.dfull, a.nh,
                     a.s, a.wave0, a.nwaves, a.gkey);
  return hipGetLastError();
}

template <int P, int M>
static hipError_t launch_bp(int np, const BlockedArgs &a, hipStream_t st) {
  switch (np) {
#define X(n, p, m) case n: if constexpr (p * (n + m) <= 192) return launch_b<n, p, m>(a, st); break;
    LN_NP(X, P, M)
#undef X
    default: break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_blocked(int np, int P, int marginal, const BlockedArgs &a, hipStream_t st) {
  if (!a.table || !a.dpk || !a.seed || !a.dfull || !a.gkey || a.nwaves >= 0xFFFFFFFFull) return hipErrorInvalidValue;
  if (!blocked_supported(np, P, marginal)) return hipErrorInvalidValue;
  if (marginal) {
    switch (P) {
      case 1: return launch_bp<1, 1>(np, a, st);
      case 2: return launch_bp<2, 1>(np, a, st);
      case 4: return launch_bp<4, 1>(np, a, st);
      case 8: return launch_bp<8, 1>(np, a, st);
    }
  } else {
    switch (P) {
      case 1: return launch_bp<1, 0>(np, a, st);
      case 2: return launch_bp<2, 0>(np, a, st);
      case 4: return launch_bp<4, 0>(np, a, st);
      case 8: return launch_bp<8, 0>(np, a, st);
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace lnorm

// lnorm_wide_tab.h -- the tabulated d-ary walk (walk_blk16) over W waves,
// for widths beyond one wave's registers, in both word formats (WB = 16:
// packed halves; 32: one column per word), instantiated by
// lnorm_wide_tab.hip (16-bit) and lnorm_wide_tab32.hip (32-bit).
//
// A workgroup of W waves walks the 64 lane segments of one group (as the
// other W-wave walks); wave wv keeps words [wv*NW, (wv+1)*NW) of every label
// sum.  After a high step each wave computes the two changed labels' subset
// tables over its words (2 * 2^M * NW v_sad_u16, as walk_blk16) and stores
// these partial tables in LDS; after a barrier every wave sums the W
// partials, so each wave holds the full tables, and a second barrier frees
// the memory for the next step.  Above 4 waves and at d = 2 (wt16_two_step)
// each wave instead totals its share of the entries over the W partials into
// a table of totals, and after the second barrier every wave reads all the
// totals (1.2-1.6x at 8 waves, d = 2, 3); the next step writes the totals
// only after its first barrier.
// The block's d^M low
// labelings are split over the waves (wave wv takes l = wv, wv + W, ...), and
// the per-wave maxima (MODE_MAX) or first matching ranks (MODE_FIRST) are
// combined once at the end.  Labels, ranks and the pattern order are those of
// walk_blk16.
#pragma once

#include "lnorm_walks.h"

namespace lnorm {

// Words per wave of the W-wave tabulated walk (the widest single-wave
// table shape up to 32 words), its widest wave count (8 for d <= 6, 4
// beyond: at 8 waves of 512 threads a wave has 256 registers and no AGPRs)
// and its table rows: the single-wave rule, with 2^M <= 64 / W so the W
// waves' partial tables take 32 KiB of LDS, and M <= 2 at 8 waves for
// d = 5, 6 (register room).
constexpr int wt16_nw(int d) { return d <= 7 ? 32 : (d == 8 || d == 9) ? 24 : d == 10 ? 16 : 0; }
constexpr int wt16_wmax(int d) { return d <= 6 ? 8 : 4; }
// the table rows of the W-wave walk: the single wave's rows at its width,
// capped by LDS and registers.  LDS: the W waves' partial tables of two
// labels (W * 2 * 2^M * 256 B; with the two-step combine one table more)
// fill up to 64 KiB for d = 3 and 32 KiB for
// other d (at d = 2 both labels change every step, so a row more does not
// cut the table work per labeling: it ran 1.7x slower).  Registers: above 4
// waves a workgroup puts 2 waves on a SIMD (256 VGPRs each), where 4 rows fit
// at d = 3, 3 at d = 2 and 4 and 2 beyond.  d = 3 over 4 waves: 13.8x the
// packed W-wave walk at 256 columns, 9x with one row less.
constexpr int wt16_m(int d, int w) {
  const int m = blk16_low_rows(d, wt16_nw(d));
  const int lds = d == 3 ? 128 : 64;   // 2^M * W at most
  int capl = 0;
  while ((2 << capl) * w <= lds) ++capl;
  const int capr = w > 4 ? (d == 3 ? 4 : d >= 5 ? 2 : 3) : 6;
  const int cap = capl < capr ? capl : capr;
  return m < cap ? m : cap;
}

// wave counts of the W-wave tabulated walk: at most wt16_wmax(d), and not 5
// at d = 6 (its 2 table rows spilled there)
constexpr int kWtWaves[] = {2, 3, 4, 5, 6, 8};
constexpr bool wt16_has(int d, int w) { return w <= wt16_wmax(d) && !(d == 6 && w == 5); }
// The W-wave tabulated walk's combine in two steps (a table of totals) above
// 4 waves and at d = 2; elsewhere at 2-4 waves it measured 0.95-1.03x of the
// single step (d = 2 over 4 waves: 1.21x).
constexpr bool wt16_two_step(int d, int w) { return w > 4 || d == 2; }
// LDS words of the W-wave tabulated walk: W partial tables of two labels
// (2^(M+1) x 64 words each), and with the two-step combine one more for the
// totals
constexpr int wt16_lds_words(int d, int w, int m) { return (w + (wt16_two_step(d, w) ? 1 : 0)) * (2 << m) * 64; }

// This wave's partial tables of labels C and C + 1 into its LDS slice
// (p[e][S][lane], e = 0 for C, 1 for C + 1).
template <int C, int D, int NW, int NWT, int M, int WB>
__device__ __forceinline__ void wt_partial_pair(const uint32_t (&u)[D][NW], const cuint32_t *neg,
                                                uint32_t *p, int lane) {
  const uint32_t z = w_zero<WB>();
#pragma unroll
  for (int S = 0; S < (1 << M); ++S) {
    uint32_t a0 = z, a1 = z;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t n = neg[(size_t)S * NWT + w];
      a0 = w_sad<WB>(u[C][w], n, a0);
      a1 = w_sad<WB>(u[C + 1][w], n, a1);
    }
    a0 = w_done<WB>(a0);
    a1 = w_done<WB>(a1);
    p[S * 64 + lane] = a0;
    p[((1 << M) + S) * 64 + lane] = a1;
  }
}

// The full tables of labels C and C + 1 from the W waves' partials, every
// wave summing every entry (T[c][0] = |u_c|; d <= 3: F_c(S); d >= 4:
// F_c(S) - |u_c|).
template <int C, int D, int W, int M>
__device__ __forceinline__ void wt_sum_pair(uint32_t (&T)[D][1 << M], const uint32_t *part, int lane) {
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    uint32_t f[1 << M];
#pragma unroll
    for (int S = 0; S < (1 << M); ++S) {
      uint32_t acc = 0;
#pragma unroll
      for (int w2 = 0; w2 < W; ++w2) acc += part[(((size_t)w2 * 2 + e) * (1 << M) + S) * 64 + lane];
      f[S] = acc;
    }
#pragma unroll
    for (int S = 0; S < (1 << M); ++S) T[C + e][S] = (D <= 3 || S == 0) ? f[S] : f[S] - f[0];
  }
}

// The two-step combine of the W waves' partial tables (wt16_two_step), with
// a barrier between the steps.
// wt_scatter: wave wv totals the entries i = wv, wv + W, ... (i = e * 2^M + S
// of label C + e, 2^(M+1) entries) over the W partials into tot[i].  Each
// lane reads 2^(M+1) / W entries of W partials, where every wave summing
// every entry read W * 2^(M+1).
template <int W, int M>
__device__ __forceinline__ void wt_scatter(const uint32_t *part, uint32_t *tot, int wv, int lane) {
  constexpr int NE = 2 << M;
#pragma unroll
  for (int i = 0; i < (NE + W - 1) / W; ++i) {
    const int idx = wv + W * i;
    if (idx < NE) {   // wave-uniform
      uint32_t acc = 0;
#pragma unroll
      for (int w2 = 0; w2 < W; ++w2) acc += part[((size_t)w2 * NE + idx) * 64 + lane];
      tot[(size_t)idx * 64 + lane] = acc;
    }
  }
}

// wt_gather_pair: the full tables of labels C and C + 1 from the totals
// (T[c][0] = |u_c|; d <= 3: F_c(S); d >= 4: F_c(S) - |u_c|).
template <int C, int D, int M>
__device__ __forceinline__ void wt_gather_pair(uint32_t (&T)[D][1 << M], const uint32_t *tot, int lane) {
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    uint32_t f[1 << M];
#pragma unroll
    for (int S = 0; S < (1 << M); ++S) {
      const int idx = e * (1 << M) + S;
      f[S] = tot[(size_t)idx * 64 + lane];
    }
#pragma unroll
    for (int S = 0; S < (1 << M); ++S) T[C + e][S] = (D <= 3 || S == 0) ? f[S] : f[S] - f[0];
  }
}

template <int C, int D, int NW, int NWT, int W, int M, int WB>
struct WtPair {
  __device__ __forceinline__ static void partial(int lo, const uint32_t (&u)[D][NW], const cuint32_t *neg,
                                                 uint32_t *p, int lane) {
    if constexpr (C + 1 < D) {
      if (lo == C) wt_partial_pair<C, D, NW, NWT, M, WB>(u, neg, p, lane);
      else WtPair<C + 1, D, NW, NWT, W, M, WB>::partial(lo, u, neg, p, lane);
    }
  }
  __device__ __forceinline__ static void sum(int lo, uint32_t (&T)[D][1 << M], const uint32_t *part, int lane) {
    if constexpr (C + 1 < D) {
      if (lo == C) wt_sum_pair<C, D, W, M>(T, part, lane);
      else WtPair<C + 1, D, NW, NWT, W, M, WB>::sum(lo, T, part, lane);
    }
  }
  __device__ __forceinline__ static void gather(int lo, uint32_t (&T)[D][1 << M], const uint32_t *tot, int lane) {
    if constexpr (C + 1 < D) {
      if (lo == C) wt_gather_pair<C, D, M>(T, tot, lane);
      else WtPair<C + 1, D, NW, NWT, W, M, WB>::gather(lo, T, tot, lane);
    }
  }
};

// Wave WV's share of the block's labelings: l = WV, WV + W, ..., in pairs
// (one v_max3 per pair in MODE_MAX).
template <int D, int M, bool ODD, int MODE, int W, int L>
__device__ __forceinline__ void wt_eval_pair(const uint32_t (&T)[D][1 << M], uint32_t totU, int32_t &best,
                                             uint32_t &fb, int32_t target) {
  constexpr int NPAT = LowCode<D, M>::NPAT;
  constexpr int L2 = L + W < NPAT ? L + W : L;
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

template <int D, int M, bool ODD, int MODE, int W, int WV, int... Is>
__device__ __forceinline__ void wt_eval(const uint32_t (&T)[D][1 << M], uint32_t totU, int32_t &best, uint32_t &fb,
                                        int32_t target, std::integer_sequence<int, Is...>) {
  (wt_eval_pair<D, M, ODD, MODE, W, WV + 2 * W * Is>(T, totU, best, fb, target), ...);
}

template <int D, int M, bool ODD, int MODE, int W, int WV>
struct WtEval {
  __device__ __forceinline__ static void run(int wv, const uint32_t (&T)[D][1 << M], uint32_t totU,
                                             int32_t &best, uint32_t &fb, int32_t target) {
    if constexpr (WV < W) {
      if (wv == WV) {
        constexpr int NPAT = LowCode<D, M>::NPAT;
        constexpr int cnt = WV < NPAT ? (NPAT - WV + W - 1) / W : 0;
        if constexpr (cnt > 0)
          wt_eval<D, M, ODD, MODE, W, WV>(T, totU, best, fb, target, std::make_integer_sequence<int, (cnt + 1) / 2>{});
      } else {
        WtEval<D, M, ODD, MODE, W, WV + 1>::run(wv, T, totU, best, fb, target);
      }
    }
  }
};

template <int D, int NW, int W, int M, int MODE, int WB>
__global__ __launch_bounds__(64 * W) void k_wide_t16(
    const uint32_t *__restrict__ Ap_, const uint32_t *__restrict__ neg_, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst) {
  constexpr int NWT = NW * W;
  constexpr int NPAT = LowCode<D, M>::NPAT;
  constexpr int PW = 2 * (1 << M) * 64;   // one wave's partial tables (words)
  static_assert(wt16_lds_words(D, W, M) * 4 <= 81920, "two workgroups per CU");
  // dynamic LDS of wt16_lds_words(D, W, M) words: the W partial tables, and
  // with the two-step combine the totals
  extern __shared__ uint32_t part[];
  uint32_t *tot = part + (size_t)W * PW;
  // per-wave results, once the last partial tables are summed: every use of
  // the partial tables ends with a barrier
  uint32_t *fin = part;
  const cuint32_t *Ap = as_const(Ap_);
  const int lane = wave_lane();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t g = blockIdx.x;
  if (((g >> 1) << 7) >= nseg) return;   // uniform over the workgroup
  const uint64_t local = lane_segment(g, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ g) & 1));
  const int qh = q - M;
  const int wbase = wv * NW;
  uint32_t *mypart = part + (size_t)wv * PW;

  uint32_t par_mask;
  if (D & 1) par_mask = pseg ? ((1u << qh) - 1u) : 0u;
  else par_mask = (qh > 0 && pseg) ? (1u << (qh - 1)) : 0u;
  const uint32_t hpar0 = ((D & 1) || qh == 0) ? pseg : 0u;

  uint32_t u[D][NW];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int l = 0; l < NW; ++l)
      u[c][l] = w_add<WB>(w_bias<WB>(), c == 0 ? Ap[(size_t)(R - 1) * NWT + wbase + l] : 0u);
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
    seed_add16<D, NW, WB>(u, Ap + (size_t)(M + i) * NWT + wbase, lab);
  }

  // full tables of every label, two labels at a time (odd d: the last pair
  // is (d-2, d-1) again)
  uint32_t T[D][1 << M];
#pragma unroll 1
  for (int c0 = 0; c0 < D; c0 += 2) {
    const int lo = c0 + 1 < D ? c0 : D - 2;
    const cuint32_t *negk = as_const(neg_);
    asm volatile("" : "+s"(negk));
    WtPair<0, D, NW, NWT, W, M, WB>::partial(lo, u, negk + wbase, mypart, lane);
    __syncthreads();
    if constexpr (wt16_two_step(D, W)) {   // constant: the barriers stay uniform
      // the totals are read before the next step's first barrier, and written
      // after it
      wt_scatter<W, M>(part, tot, wv, lane);
      __syncthreads();
      WtPair<0, D, NW, NWT, W, M, WB>::gather(lo, T, tot, lane);
    } else {
      WtPair<0, D, NW, NWT, W, M, WB>::sum(lo, T, part, lane);
      __syncthreads();
    }
  }
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
  for (uint32_t k = 0; k < nblk; ++k) {   // nblk, j, from and up are workgroup-uniform
    if (k > 0) {
      const int j = __builtin_ctzll(~cnt) >> 2;
      const uint64_t low = (1ull << (4 * j)) - 1ull;
      cnt = (cnt + 1ull) | (biaspat & low);
      const int v = (int)((cnt >> (4 * j)) & 15ull) - (16 - D);
      const int p = (pm >> j) & 1;
      const int from = p ? D - v : v - 1;
      const int up = p ? 0 : 1;
      pm ^= (1u << j) - 1u;
      Move16<0, D, NW, WB>::run(u, Ap + (size_t)(M + j) * NWT + wbase, from, up);
      const int lo = up ? from : from - 1;
      const cuint32_t *negk = as_const(neg_);
      asm volatile("" : "+s"(negk));
      WtPair<0, D, NW, NWT, W, M, WB>::partial(lo, u, negk + wbase, mypart, lane);
      __syncthreads();
      if constexpr (wt16_two_step(D, W)) {   // constant: the barriers stay uniform
        // the totals are read before the next step's first barrier, and written
        // after it
        wt_scatter<W, M>(part, tot, wv, lane);
        __syncthreads();
        WtPair<0, D, NW, NWT, W, M, WB>::gather(lo, T, tot, lane);
      } else {
        WtPair<0, D, NW, NWT, W, M, WB>::sum(lo, T, part, lane);
        __syncthreads();
      }
      totU = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) totU += T[c][0];
    }
    if constexpr (MODE == MODE_MAX) {
      uint32_t unused = 0;
      WtEval<D, M, false, MODE, W, 0>::run(wv, T, totU, best, unused, target);
    } else {
      uint32_t fb = 0xFFFFFFFFu;
      if ((hpar0 ^ k) & 1u) {
#pragma unroll
        for (int c = 0; c < D; ++c)
#pragma unroll
          for (int S = 0; S < (1 << M); ++S) asm volatile("" : "+v"(T[c][S]));
        asm volatile("" : "+v"(totU));
        WtEval<D, M, true, MODE, W, 0>::run(wv, T, totU, best, fb, target);
      } else {
        WtEval<D, M, false, MODE, W, 0>::run(wv, T, totU, best, fb, target);
      }
      const uint32_t c = fb != 0xFFFFFFFFu ? k * (uint32_t)NPAT + fb : 0xFFFFFFFFu;
      first = c < first ? c : first;
    }
  }
  const uint32_t nsteps = nblk * (uint32_t)NPAT;

  if constexpr (MODE == MODE_MAX) fin[wv * 64 + lane] = (uint32_t)best;
  else fin[wv * 64 + lane] = first;
  __syncthreads();
  if (wv == 0) {
    if constexpr (MODE == MODE_MAX) {
      int32_t m = (int32_t)fin[lane];
#pragma unroll
      for (int w2 = 1; w2 < W; ++w2) m = (int32_t)fin[w2 * 64 + lane] > m ? (int32_t)fin[w2 * 64 + lane] : m;
      if (active) unit_best[local] = m;
      const int32_t wm = wave_max_i32(active ? m : INT32_MIN);
      if (wm > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, wm);
    } else {
      uint32_t f = fin[lane];
#pragma unroll
      for (int w2 = 1; w2 < W; ++w2) f = fin[w2 * 64 + lane] < f ? fin[w2 * 64 + lane] : f;
      const uint64_t r = (active && f != 0xFFFFFFFFu) ? (seg * (uint64_t)nsteps + f) : ~0ull;
      const uint64_t m = wave_min_u64(r);
      if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
    }
  }
}

// ---------------------------------------------------------------------------
// The tabulated L1 / LM walk (walk_tab16_bin) over W waves: wave wv keeps
// words [wv*NW, (wv+1)*NW) of the high sum u (LM: column 0, outside the
// words, in every wave; wave 0 adds it).  Per block each wave computes the
// 2^M patterns' partial sums over its words and stores them in LDS; after a
// barrier wave wv totals the W partials of its patterns (l = wv, wv + W, ...)
// and keeps their maximum (MODE_MAX) or first match (MODE_FIRST).
// Shapes (W, NW): (8, 32), (10, 32), (12, 32), (16, 32), (16, 64) for 512, 640,
// 768, 1024, 2048 columns.
// ---------------------------------------------------------------------------
constexpr int wtb_m(int w) { return w > 8 ? 4 : 5; }   // partials: W * 2^M * 256 B <= 64 KiB

// The 2^M patterns' partial sums over the wave's NW words.  The table words
// arrive by scalar loads, which may return out of order, so a wait for one
// load waits for all.  The wave reads its 2^M x NW table words in batches of
// KB (in pattern order); each batch's loads are issued right after the first
// operation of the batch before it (the one that waits for that batch), so a
// batch has KB - 1 operations of this wave to arrive in, not the few the
// compiler's own order leaves.  Scheduling barriers pin that order.
// KB divides NW: 40-word batches at 64 words per wave (batches straddling
// patterns) ran 50x slower, the batch loop no longer fully unrolled.
// NC: interleaved |.| chains; two measured faster only for 32-bit words at
// 64 words per wave (1024-word shape: 298 against 315 ms on 34 x 1000
// int16-range entries), one elsewhere (1 to 3 % ahead, one register fewer).
constexpr int wtb_batch(int nw) { return nw < 32 ? nw : 32; }
constexpr int wtb_chains(int nw, int wb) { return (wb == 32 && nw >= 64) ? 2 : 1; }

template <int NW, int NWT, int M, bool MARG, int WB, int KB = wtb_batch(NW), int NC = wtb_chains(NW, WB)>
__device__ __forceinline__ void wtb_partials(const uint32_t (&u)[NW], int32_t u0, const cuint32_t *neg,
                                             const cint32_t *pos0, bool wave0, uint32_t *p, int lane) {
  constexpr int NT = (1 << M) * NW;   // table words of this wave
  constexpr int NB = (NT + KB - 1) / KB;
  static_assert(KB <= NW && NW % KB == 0, "batches tile each pattern's words");
  uint32_t buf[2][KB];
  int32_t pn[2] = {0, 0};   // LM: pos0 of the pattern starting in the batch
#pragma unroll
  for (int i = 0; i < KB; ++i) buf[0][i] = neg[(size_t)(i / NW) * NWT + i % NW];
  if (MARG) pn[0] = pos0[0];
  uint32_t acc[NC] = {};   // NC interleaved chains (words w = c mod NC)
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      const int f = b * KB + i;
      if (f < NT) {
        const int S = f / NW, w = f % NW, c = w % NC;
        const uint32_t a0 = w >= NC ? acc[c] : (w == 0 && MARG && wave0) ? (uint32_t)(u0 + pn[b & 1]) : w_zero<WB>();
        acc[c] = w_sad<WB>(u[w], buf[b & 1][i], a0);
        if (w == NW - 1) {
          uint32_t t = w_done<WB>(acc[0]);
#pragma unroll
          for (int c2 = 1; c2 < NC; ++c2) t += w_done<WB>(acc[c2]);
          p[S * 64 + lane] = t;
        }
      }
      if (i == 0) {
        __builtin_amdgcn_sched_barrier(0);
        if (b + 1 < NB) {
#pragma unroll
          for (int j = 0; j < KB; ++j) {
            const int f1 = (b + 1) * KB + j;
            if (f1 < NT) buf[(b + 1) & 1][j] = neg[(size_t)(f1 / NW) * NWT + f1 % NW];
          }
          const int S1 = ((b + 1) * KB + NW - 1) / NW;   // the pattern starting in batch b + 1
          if (MARG && S1 < (1 << M) && S1 * NW < (b + 2) * KB) pn[(b + 1) & 1] = pos0[S1];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int W, int M, int MODE, bool ODD, int WV, int... Is>
__device__ __forceinline__ void wtb_eval(const uint32_t *part, int lane, int32_t &best, uint32_t &fb, int32_t target,
                                         std::integer_sequence<int, Is...>) {
  auto one = [&](auto II) {
    constexpr int l = WV + W * decltype(II)::value;
    // MODE_MAX: any order (a block holds all 2^M patterns); MODE_FIRST: the
    // pattern of low index l
    constexpr int S = MODE == MODE_MAX ? l : ((l ^ (l >> 1)) ^ (ODD ? (1 << (M - 1)) : 0));
    uint32_t tot = 0;
#pragma unroll
    for (int w2 = 0; w2 < W; ++w2) tot += part[((size_t)w2 * (1 << M) + S) * 64 + lane];
    if constexpr (MODE == MODE_MAX) {
      best = (int32_t)tot > best ? (int32_t)tot : best;
    } else {
      const uint32_t c = ((int32_t)tot == target) ? (uint32_t)l : 0xFFFFFFFFu;
      fb = c < fb ? c : fb;
    }
  };
  (one(std::integral_constant<int, Is>{}), ...);
}

template <int W, int M, int MODE, bool ODD, int WV>
struct WtbEval {
  __device__ __forceinline__ static void run(int wv, const uint32_t *part, int lane, int32_t &best, uint32_t &fb,
                                             int32_t target) {
    if constexpr (WV < W) {
      if (wv == WV) {
        constexpr int cnt = WV < (1 << M) ? ((1 << M) - WV + W - 1) / W : 0;
        if constexpr (cnt > 0)
          wtb_eval<W, M, MODE, ODD, WV>(part, lane, best, fb, target, std::make_integer_sequence<int, cnt>{});
      } else {
        WtbEval<W, M, MODE, ODD, WV + 1>::run(wv, part, lane, best, fb, target);
      }
    }
  }
};

// Waves per SIMD the compiler must leave room for (at most 512 / that many
// VGPRs): the workgroups per CU that LDS admits (W * 2^M * 256 B each, 160 KiB
// per CU), W waves each over 4 SIMDs, where the wave's registers can fit: 16
// waves of 32 words (64 KiB, 2 per CU: 8 waves, 64 VGPRs; unbounded, the
// compiler took 71 and one workgroup per CU ran 1.2x slower), 10 of 32
// (40 KiB, 3 per CU: 8 waves); otherwise what the registers give.
constexpr int wtb_waves_per_eu(int w, int nw) { return nw == 32 && (w == 16 || w == 10) ? 8 : 1; }

// M: the table rows (wtb_m(W) in both word formats).
template <int W, int NW, bool MARG, int MODE, int WB, int M>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(wtb_waves_per_eu(W, NW))))
void k_wide_tb16(
    const uint32_t *__restrict__ Ap_, const uint32_t *__restrict__ tab_, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst) {
  constexpr int NWT = NW * W;
  static_assert(W * (1 << M) * 64 * 4 <= 65536, "partials exceed 64 KiB");
  __shared__ uint32_t part[W * (1 << M) * 64];
  uint32_t *fin = part;   // per-wave results, once the walk's last barrier has passed
  const cuint32_t *Ap = as_const(Ap_);
  const int lane = wave_lane();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t g = blockIdx.x;
  if (((g >> 1) << 7) >= nseg) return;   // uniform over the workgroup
  const uint64_t local = lane_segment(g, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ g) & 1));
  const int qh = q - M;
  const int wbase = wv * NW;
  const cuint32_t *A2 = Ap + (size_t)R * NWT;                          // packed 2 A
  const cint32_t *A0 = (const cint32_t *)(Ap + (size_t)2 * R * NWT);   // LM: column 0
  const cint32_t *pos0 = (const cint32_t *)(as_const(tab_) + ((size_t)NWT << M));
  uint32_t *mypart = part + (size_t)wv * (1 << M) * 64;

  const uint64_t h0 = seg << qh;
  const uint64_t g0 = h0 ^ (h0 >> 1);
  uint32_t u[NW];
  int32_t u0 = MARG ? A0[R - 1] : 0;
#pragma unroll
  for (int l = 0; l < NW; ++l) u[l] = w_add<WB>(w_bias<WB>(), Ap[(size_t)(R - 1) * NWT + wbase + l]);
  for (int j = 0; j < R - 1 - M; ++j) {
    const uint32_t pmask = ((g0 >> j) & 1) ? 0xFFFFFFFFu : 0u;   // +row where the bit is 1, -row elsewhere
    const cuint32_t *row = Ap + (size_t)(M + j) * NWT + wbase;
#pragma unroll
    for (int l = 0; l < NW; ++l) {
      const uint32_t v = row[l];
      u[l] = w_add<WB>(u[l], v & pmask);
      u[l] = w_sub<WB>(u[l], v & ~pmask);
    }
    if (MARG) u0 += pmask ? A0[M + j] : -A0[M + j];
  }

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  const uint32_t nblk = 1u << qh;
  for (uint32_t k = 0; k < nblk; ++k) {   // workgroup-uniform
    if (k > 0) {
      const int j = __builtin_ctz(k);
      const uint32_t hb = (j + 1 < qh) ? ((k >> (j + 1)) & 1u) : pseg;
      const cuint32_t *row2 = A2 + (size_t)(M + j) * NWT + wbase;
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
    const cuint32_t *negk = as_const(tab_);
    asm volatile("" : "+s"(negk));
    if constexpr (MODE == MODE_MAX) {
      wtb_partials<NW, NWT, M, MARG, WB>(u, u0, negk + wbase, pos0, wv == 0, mypart, lane);
    } else {
      // MODE_FIRST re-walks a single unit: the patterns' partial sums as a
      // loop over the wave-uniform pattern index (half the unrolled code)
#pragma unroll 1
      for (int S = 0; S < (1 << M); ++S) {
        const cuint32_t *row = negk + wbase + (size_t)S * NWT;
        uint32_t acc = (MARG && wv == 0) ? (uint32_t)(u0 + pos0[S]) : w_zero<WB>();
#pragma unroll
        for (int w = 0; w < NW; ++w) acc = w_sad<WB>(u[w], row[w], acc);
        mypart[S * 64 + lane] = acc;
      }
    }
    __syncthreads();
    if constexpr (MODE == MODE_MAX) {
      uint32_t unused = 0;
      WtbEval<W, M, MODE, false, 0>::run(wv, part, lane, best, unused, target);
    } else {
      uint32_t fb = 0xFFFFFFFFu;
      const uint32_t hp = qh > 0 ? (k & 1u) : pseg;
      if (hp) WtbEval<W, M, MODE, true, 0>::run(wv, part, lane, best, fb, target);
      else WtbEval<W, M, MODE, false, 0>::run(wv, part, lane, best, fb, target);
      const uint32_t c = fb != 0xFFFFFFFFu ? (k << M) + fb : 0xFFFFFFFFu;
      first = c < first ? c : first;
    }
    __syncthreads();
  }
  const uint32_t nsteps = nblk << M;

  if constexpr (MODE == MODE_MAX) fin[wv * 64 + lane] = (uint32_t)best;
  else fin[wv * 64 + lane] = first;
  __syncthreads();
  if (wv == 0) {
    if constexpr (MODE == MODE_MAX) {
      int32_t m = (int32_t)fin[lane];
#pragma unroll
      for (int w2 = 1; w2 < W; ++w2) m = (int32_t)fin[w2 * 64 + lane] > m ? (int32_t)fin[w2 * 64 + lane] : m;
      if (active) unit_best[local] = m;
      const int32_t wm = wave_max_i32(active ? m : INT32_MIN);
      if (wm > __builtin_amdgcn_readfirstlane(*(volatile int32_t *)gbest)) atomicMax(gbest, wm);
    } else {
      uint32_t f = fin[lane];
#pragma unroll
      for (int w2 = 1; w2 < W; ++w2) f = fin[w2 * 64 + lane] < f ? fin[w2 * 64 + lane] : f;
      const uint64_t r = (active && f != 0xFFFFFFFFu) ? (seg * (uint64_t)nsteps + f) : ~0ull;
      const uint64_t m = wave_min_u64(r);
      if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
    }
  }
}

// Launchers of both word formats; the width in columns is a.ny_pad (16-bit:
// 2 columns per word).
template <int W, int NW, bool MARG, int WB, int M>
static hipError_t lwtb(Mode mode, const PlainArgs &a, hipStream_t st) {
  const uint64_t groups = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)groups), blk(64 * W);
  const uint32_t *Ap = (const uint32_t *)a.A;
  const uint32_t *tab = (const uint32_t *)a.D;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_wide_tb16<W, NW, MARG, MODE_MAX, WB, M>), grid, blk, 0, st, Ap, tab, a.R, a.q, a.seg0,
                       a.nseg, a.target, a.unit_best, a.gbest, a.gfirst);
  else
    hipLaunchKernelGGL((k_wide_tb16<W, NW, MARG, MODE_FIRST, WB, M>), grid, blk, 0, st, Ap, tab, a.R, a.q, a.seg0,
                       a.nseg, a.target, a.unit_best, a.gbest, a.gfirst);
  return hipGetLastError();
}

// Shapes (W, NW) of the W-wave L1 / LM walk by total words, in both word
// formats: 256, 320, 384, 512 or 1024 words (8, 10, 12 or 16 waves of 32
// words, 16 of 64) with wtb_m(W) table rows.  Each wave reads its slice of
// every table row by scalar loads, so the scalar-cache working set of a CU
// is the whole table: in 32-bit words the 1024-word shape misses in the
// scalar cache on about half its loads, where one wave reading the whole
// table misses on 2%; the batched loads of wtb_partials hide most of that
// latency.  Fewer rows (3 at 1024 words: 0.93x with the batched loads) or
// 128-word waves (2 to 8 waves, 5 rows: 0.77x) measured slower.
template <bool MARG, int WB>
static hipError_t lwtb_m(Mode mode, const PlainArgs &a, hipStream_t st) {
  switch (WB == 16 ? a.ny_pad / 2 : a.ny_pad) {
    case 256: return lwtb<8, 32, MARG, WB, wtb_m(8)>(mode, a, st);
    case 320: return lwtb<10, 32, MARG, WB, wtb_m(10)>(mode, a, st);
    case 384: return lwtb<12, 32, MARG, WB, wtb_m(12)>(mode, a, st);
    case 512: return lwtb<16, 32, MARG, WB, wtb_m(16)>(mode, a, st);
    case 1024: return lwtb<16, 64, MARG, WB, wtb_m(16)>(mode, a, st);
    default: return hipErrorInvalidValue;
  }
}

// the W-wave L1 / LM walk's total words for `need` words (0: none)
inline int wtb_words_for(int need) {
  for (int w : {256, 320, 384, 512, 1024})
    if (need <= w) return w;
  return 0;
}

template <int D, int W, int WB>
static hipError_t lwt(Mode mode, const PlainArgs &a, hipStream_t st) {
  constexpr int NW = wt16_nw(D), M = wt16_m(D, W);
  const uint64_t groups = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)groups), blk(64 * W);
  const size_t lds = (size_t)wt16_lds_words(D, W, M) * 4;
  const uint32_t *Ap = (const uint32_t *)a.A;
  const uint32_t *neg = (const uint32_t *)a.D;
  auto launch = [&](auto kernel) {
    if (lds > (64u << 10)) {   // beyond the default limit of dynamic LDS
      const hipError_t e =
          hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, grid, blk, lds, st, Ap, neg, a.R, a.q, a.seg0, a.nseg, a.target, a.unit_best, a.gbest,
                       a.gfirst);
    return hipGetLastError();
  };
  if (mode == MODE_MAX) return launch(k_wide_t16<D, NW, W, M, MODE_MAX, WB>);
  return launch(k_wide_t16<D, NW, W, M, MODE_FIRST, WB>);
}

template <int D, int WB>
static hipError_t lwt_d(Mode mode, const PlainArgs &a, hipStream_t st) {
  const int words = WB == 16 ? a.ny_pad / 2 : a.ny_pad;
  if (words == 2 * wt16_nw(D)) return lwt<D, 2, WB>(mode, a, st);
  if (words == 3 * wt16_nw(D)) return lwt<D, 3, WB>(mode, a, st);
  if (words == 4 * wt16_nw(D)) return lwt<D, 4, WB>(mode, a, st);
  if constexpr (wt16_has(D, 5))
    if (words == 5 * wt16_nw(D)) return lwt<D, 5, WB>(mode, a, st);
  if constexpr (wt16_has(D, 6))
    if (words == 6 * wt16_nw(D)) return lwt<D, 6, WB>(mode, a, st);
  if constexpr (wt16_has(D, 8))
    if (words == 8 * wt16_nw(D)) return lwt<D, 8, WB>(mode, a, st);
  return hipErrorInvalidValue;
}

// the W-wave d-ary walk's total words for `need` words (0: none)
inline int wt_words_for(int d, int need) {
  const int nw = (d >= 2 && d <= 10) ? wt16_nw(d) : 0;
  for (int w : kWtWaves)
    if (nw > 0 && wt16_has(d, w) && need <= nw * w) return nw * w;
  return 0;
}

inline int wt_rows_for(int d, int words) {
  const int nw = (d >= 2 && d <= 10) ? wt16_nw(d) : 0;
  return nw > 0 ? wt16_m(d, words / nw) : 0;
}

// lnorm_wide_tab32_hi.hip: the 32-bit W-wave d-ary walk for d = 5..10 (the
// d = 2..4 kernels and the launch entry are in lnorm_wide_tab32.hip, so the
// two halves compile in parallel)
hipError_t launch_wide_t32_hi(Mode mode, const PlainArgs &a, hipStream_t st);

template <int WB>
static hipError_t lwt_any(Mode mode, const PlainArgs &a, hipStream_t st) {
  switch (a.d) {
    case 2: return lwt_d<2, WB>(mode, a, st);
    case 3: return lwt_d<3, WB>(mode, a, st);
    case 4: return lwt_d<4, WB>(mode, a, st);
    case 5: return lwt_d<5, WB>(mode, a, st);
    case 6: return lwt_d<6, WB>(mode, a, st);
    case 7: return lwt_d<7, WB>(mode, a, st);
    case 8: return lwt_d<8, WB>(mode, a, st);
    case 9: return lwt_d<9, WB>(mode, a, st);
    case 10: return lwt_d<10, WB>(mode, a, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace lnorm

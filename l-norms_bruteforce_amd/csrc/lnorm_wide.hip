// lnorm_wide.hip -- the walks for matrices wider than one wave's registers
// hold: the binary walk (L1 / LM / L2, more than 128 partial sums) and the
// packed 16-bit d-ary walk (wider than walk_p16's widths).
//
// A workgroup of W waves walks the 64 lane segments lane_segment(g, .) of
// workgroup g, exactly like one wave of the plain kernel, but wave wv keeps
// only the columns [wv*NY, (wv+1)*NY) of the partial sums.  The flipped row
// and its direction are the same for the whole workgroup, so every wave
// still reads its slice of the row delta through scalar loads.  Each wave's
// |.| sum over its slice is staged in LDS for 8 steps; after a barrier each
// wave totals 8/W of those steps over the W slices and keeps the running
// maximum (MODE_MAX) or the first step equal to the target (MODE_FIRST).
// Same objective, units and two-phase contract as k_plain_bin.
#include "lnorm_walks.h"

namespace lnorm {

template <int KIND, int NY, int W, int MODE>
__global__ __launch_bounds__(64 * W) void k_wide_bin(
    const int32_t *__restrict__ A_, const int32_t *__restrict__ D_, int R, int q, uint64_t seg0, uint64_t nseg,
    int32_t target, int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest,
    unsigned long long *__restrict__ gfirst, uint32_t B) {
  constexpr int NYT = NY * W;            // padded width of the whole workgroup
  constexpr int K = 8;                   // steps per LDS combine
  static_assert(K % W == 0, "steps per combine must split evenly over the waves");
  constexpr bool TWO = (KIND == KIND_L2);
  __shared__ uint32_t part[W * K * 64];
  __shared__ uint32_t fin[W * 64];
  const cint32_t *A = as_const(A_), *Dt = as_const(D_);
  const int lane = wave_lane();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t g = blockIdx.x;
  if (((g >> 1) << 7) >= nseg) return;  // uniform over the workgroup
  const uint64_t local = lane_segment(g, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t par = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ g) & 1));
  const int wbase = wv * NY;
  const bool mw = KIND == KIND_LM && wv == 0;   // the wave holding the marginal column 0

  uint32_t y[NY];
  uint32_t z[TWO ? NY : 1];
  const uint64_t N0 = seg << q;
  const uint64_t g0 = N0 ^ (N0 >> 1);
#pragma unroll
  for (int l = 0; l < NY; ++l) y[l] = (uint32_t)A[(size_t)(R - 1) * NYT + wbase + l] + kBias32;
  if constexpr (TWO) {
#pragma unroll
    for (int l = 0; l < NY; ++l) z[l] = kBias32;
  }
  for (int i = 0; i < R - 1; ++i) {
    const bool gi = (g0 >> i) & 1;
    const cint32_t *row = A + (size_t)i * NYT + wbase;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const uint32_t a = (uint32_t)row[l];
      if constexpr (TWO) {
        z[l] += gi ? a : 0u;
        y[l] += gi ? 0u : a;
      } else {
        y[l] += gi ? a : 0u - a;
      }
    }
  }

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  const uint32_t zero = opaque_zero();
  const uint32_t nsteps = 1u << q;       // the host guarantees q >= 3 (nsteps % K == 0)
  const uint32_t nu = par << q;
  for (uint32_t t0 = 0; t0 < nsteps; t0 += K) {
#pragma unroll 1
    for (int b = 0; b < K; ++b) {
      const uint32_t t = t0 + (uint32_t)b;
      if (t > 0) {
        const int j = __builtin_ctz(t);
        const uint32_t up = (((nu | t) >> (j + 1)) & 1u) ^ 1u;
        const cint32_t *dp = Dt + (size_t)(up * (uint32_t)(R - 1) + (uint32_t)j) * NYT + wbase;
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
      }
      uint32_t acc = zero;
      if constexpr (KIND == KIND_LM) {
        // column 0 enters without |.| (L_OpenMP.c:40-41); it is wave 0's first
        acc = mw ? (y[0] ^ kBias32) : sad32(y[0], B, zero);
#pragma unroll
        for (int l = 1; l < NY; ++l) acc = sad32(y[l], B, acc);
      } else {
#pragma unroll
        for (int l = 0; l < NY; ++l) {
          acc = sad32(y[l], B, acc);
          if constexpr (TWO) acc = sad32(z[l], B, acc);
        }
      }
      part[((size_t)wv * K + b) * 64 + lane] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int bb = 0; bb < K / W; ++bb) {
      const int b = wv * (K / W) + bb;
      uint32_t tot = 0;
#pragma unroll
      for (int w2 = 0; w2 < W; ++w2) tot += part[((size_t)w2 * K + b) * 64 + lane];
      if constexpr (MODE == MODE_MAX) {
        best = (int32_t)tot > best ? (int32_t)tot : best;
      } else {
        const uint32_t c = ((int32_t)tot == target) ? t0 + (uint32_t)b : 0xFFFFFFFFu;
        first = c < first ? c : first;
      }
    }
    __syncthreads();
  }

  // combine the waves' results per lane segment
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
      const uint64_t r = (active && f != 0xFFFFFFFFu) ? ((seg << q) + f) : ~0ull;
      const uint64_t m = wave_min_u64(r);
      if (m != ~0ull) atomicMin(gfirst, (unsigned long long)m);
    }
  }
}

// ---------------------------------------------------------------------------
// Packed 16-bit d-ary walk (d = 3..10) over W waves: wave wv keeps words
// [wv*NW, (wv+1)*NW) of every label vector (two columns per word, as in
// walk_p16).  The digit counter and the move dispatch are workgroup-uniform;
// each wave moves and re-sums its slice, and the per-step sums combine
// through LDS in groups of 8 steps (the last group of a segment may be
// shorter: d^q is no multiple of 8 for odd d).
// ---------------------------------------------------------------------------
template <int D, int NW, int W, int MODE>
__global__ __launch_bounds__(64 * W) void k_wide_p16(
    const uint32_t *__restrict__ Ap_, int R, int q, uint64_t seg0, uint64_t nseg, int32_t target,
    int32_t *__restrict__ unit_best, int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst,
    uint32_t B) {
  constexpr int NWT = NW * W;
  constexpr int K = 8;
  static_assert(K % W == 0, "steps per combine must split evenly over the waves");
  __shared__ uint32_t part[W * K * 64];
  __shared__ uint32_t fin[W * 64];
  const cuint32_t *Ap = as_const(Ap_);
  const int lane = wave_lane();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t g = blockIdx.x;
  if (((g >> 1) << 7) >= nseg) return;  // uniform over the workgroup
  const uint64_t local = lane_segment(g, lane);
  const bool active = local < nseg;
  const uint64_t seg = seg0 + local;
  const uint32_t pseg = (uint32_t)__builtin_amdgcn_readfirstlane((int)((seg0 ^ g) & 1));
  const int wbase = wv * NW;

  uint32_t par_mask;
  if (D & 1) par_mask = pseg ? ((1u << q) - 1u) : 0u;
  else par_mask = (q > 0 && pseg) ? (1u << (q - 1)) : 0u;

  uint32_t t[D][NW];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int l = 0; l < NW; ++l)
      t[c][l] = pk_add(kBias16x2, c == 0 ? Ap[(size_t)(R - 1) * NWT + wbase + l] : 0u);
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
    seed_add16<D, NW>(t, Ap + (size_t)i * NWT + wbase, lab);
  }
  const uint32_t zero = opaque_zero();
  uint32_t S[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    uint32_t sc = zero;
#pragma unroll
    for (int l = 0; l < NW; ++l) sc = sad16(t[c][l], B, sc);
    S[c] = sc;
  }

  uint64_t biaspat = 0;
  for (int i = 0; i < q; ++i) biaspat |= (uint64_t)(16 - D) << (4 * i);
  uint64_t cnt = biaspat;
  uint32_t pm = par_mask;
  uint32_t nsteps = 1;
  for (int i = 0; i < q; ++i) nsteps *= (uint32_t)D;

  int32_t best = INT32_MIN;
  uint32_t first = 0xFFFFFFFFu;
  for (uint32_t t0 = 0; t0 < nsteps; t0 += K) {
    const uint32_t kb = nsteps - t0 < (uint32_t)K ? nsteps - t0 : (uint32_t)K;   // uniform
#pragma unroll 1
    for (uint32_t b = 0; b < kb; ++b) {
      const uint32_t tt = t0 + b;
      if (tt > 0) {
        const int j = __builtin_ctzll(~cnt) >> 2;
        const uint64_t low = (1ull << (4 * j)) - 1ull;
        cnt = (cnt + 1ull) | (biaspat & low);
        const int v = (int)((cnt >> (4 * j)) & 15ull) - (16 - D);
        const int p = (pm >> j) & 1;
        const int from = p ? D - v : v - 1;
        pm ^= (1u << j) - 1u;
        Mover16<0, D, NW>::run(t, S, Ap + (size_t)j * NWT + wbase, from, p ? 0 : 1, B);
      }
      uint32_t acc = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) acc += S[c];
      part[((size_t)wv * K + b) * 64 + lane] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int bb = 0; bb < K / W; ++bb) {
      const uint32_t b = (uint32_t)(wv * (K / W) + bb);
      if (b < kb) {
        uint32_t tot = 0;
#pragma unroll
        for (int w2 = 0; w2 < W; ++w2) tot += part[((size_t)w2 * K + b) * 64 + lane];
        if constexpr (MODE == MODE_MAX) {
          best = (int32_t)tot > best ? (int32_t)tot : best;
        } else {
          const uint32_t c = ((int32_t)tot == target) ? t0 + b : 0xFFFFFFFFu;
          first = c < first ? c : first;
        }
      }
    }
    __syncthreads();
  }

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

// Packed W-wave shapes: NW words per wave (d * NW state registers, sized
// so the walk state stays in the 256 architectural VGPRs) and W = 2, 4 or 8
// waves.
template <int D>
constexpr int wide16_nw() {
  return D == 3 ? 40 : D == 4 ? 28 : D == 5 ? 24 : D == 6 ? 20 : D == 7 ? 16 : D == 8 ? 14 : 12;
}
static int wide16_nw_rt(int d) {
  switch (d) {
    case 3: return wide16_nw<3>();
    case 4: return wide16_nw<4>();
    case 5: return wide16_nw<5>();
    case 6: return wide16_nw<6>();
    case 7: return wide16_nw<7>();
    case 8: return wide16_nw<8>();
    case 9: return wide16_nw<9>();
    case 10: return wide16_nw<10>();
    default: return 0;
  }
}

int wide16_words(int d, int ny) {
  const int nw = wide16_nw_rt(d);
  const int need = (ny + 1) / 2;
  for (int w = 2; nw > 0 && w <= 8; w *= 2)
    if (need <= nw * w) return nw * w;
  return 0;
}

template <int D, int W>
static hipError_t lw16(Mode mode, const PlainArgs &a, hipStream_t st) {
  constexpr int NW = wide16_nw<D>();
  const uint64_t groups = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)groups), blk(64 * W);
  const uint32_t *Ap = (const uint32_t *)a.A;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_wide_p16<D, NW, W, MODE_MAX>), grid, blk, 0, st, Ap, a.R, a.q, a.seg0, a.nseg, a.target,
                       a.unit_best, a.gbest, a.gfirst, kBias16x2);
  else
    hipLaunchKernelGGL((k_wide_p16<D, NW, W, MODE_FIRST>), grid, blk, 0, st, Ap, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst, kBias16x2);
  return hipGetLastError();
}

template <int D>
static hipError_t lw16_d(Mode mode, const PlainArgs &a, hipStream_t st) {
  const int words = a.ny_pad / 2;
  if (words == 2 * wide16_nw<D>()) return lw16<D, 2>(mode, a, st);
  if (words == 4 * wide16_nw<D>()) return lw16<D, 4>(mode, a, st);
  if (words == 8 * wide16_nw<D>()) return lw16<D, 8>(mode, a, st);
  return hipErrorInvalidValue;
}

hipError_t launch_wide16(Mode mode, const PlainArgs &a, hipStream_t st) {
  if (!a.A || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (wide16_words(a.d, a.ny_pad) * 2 != a.ny_pad) return hipErrorInvalidValue;
  switch (a.d) {
    case 3: return lw16_d<3>(mode, a, st);
    case 4: return lw16_d<4>(mode, a, st);
    case 5: return lw16_d<5>(mode, a, st);
    case 6: return lw16_d<6>(mode, a, st);
    case 7: return lw16_d<7>(mode, a, st);
    case 8: return lw16_d<8>(mode, a, st);
    case 9: return lw16_d<9>(mode, a, st);
    case 10: return lw16_d<10>(mode, a, st);
    default: return hipErrorInvalidValue;
  }
}

// (NY, W) of each instantiated padded width: NY columns per wave, W waves.
// L2 keeps two vectors per column, so its waves hold at most 64 columns.
static bool wide_shape(Kind kind, int ny_pad, int &NY, int &W) {
  static const int shapes_bin[][2] = {{64, 4}, {96, 4}, {128, 4}, {96, 8}, {128, 8}};
  static const int shapes_l2[][2] = {{64, 4}, {64, 8}};
  const bool two = kind == KIND_L2;
  const int n = two ? 2 : 5;
  for (int i = 0; i < n; ++i) {
    const int *sh = two ? shapes_l2[i] : shapes_bin[i];
    if (sh[0] * sh[1] == ny_pad) {
      NY = sh[0];
      W = sh[1];
      return true;
    }
  }
  return false;
}

// Smallest instantiated width >= ny, 0 if none.
int wide_ny_pad(Kind kind, int ny) {
  static const int bin[] = {256, 384, 512, 768, 1024};
  static const int l2[] = {256, 512};
  if (kind == KIND_L2) {
    for (int w : l2) if (w >= ny) return w;
    return 0;
  }
  if (kind != KIND_L1 && kind != KIND_LM) return 0;
  for (int w : bin) if (w >= ny) return w;
  return 0;
}

template <int KIND, int NY, int W>
static hipError_t lw(Mode mode, const PlainArgs &a, hipStream_t st) {
  const uint64_t groups = 2 * ((a.nseg + 127) / 128);
  const dim3 grid((unsigned)groups), blk(64 * W);
  if (mode == MODE_MAX)
    hipLaunchKernelGGL((k_wide_bin<KIND, NY, W, MODE_MAX>), grid, blk, 0, st, a.A, a.D, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
  else
    hipLaunchKernelGGL((k_wide_bin<KIND, NY, W, MODE_FIRST>), grid, blk, 0, st, a.A, a.D, a.R, a.q, a.seg0, a.nseg,
                       a.target, a.unit_best, a.gbest, a.gfirst, kBias32);
  return hipGetLastError();
}

template <int KIND>
static hipError_t lw_kind(Mode mode, const PlainArgs &a, hipStream_t st, int NY, int W) {
#define X(ny, w) if (NY == ny && W == w) return lw<KIND, ny, w>(mode, a, st);
  if constexpr (KIND == KIND_L2) {
    X(64, 4) X(64, 8)
  } else {
    X(64, 4) X(96, 4) X(128, 4) X(96, 8) X(128, 8)
  }
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_wide(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st) {
  if (!a.A || !a.D || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  if (a.q < 3) return hipErrorInvalidValue;   // steps combine in groups of 8
  int NY = 0, W = 0;
  if (wide_ny_pad(kind, a.ny_pad) != a.ny_pad || !wide_shape(kind, a.ny_pad, NY, W)) return hipErrorInvalidValue;
  switch (kind) {
    case KIND_L1: return lw_kind<KIND_L1>(mode, a, st, NY, W);
    case KIND_LM: return lw_kind<KIND_LM>(mode, a, st, NY, W);
    case KIND_L2: return lw_kind<KIND_L2>(mode, a, st, NY, W);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace lnorm

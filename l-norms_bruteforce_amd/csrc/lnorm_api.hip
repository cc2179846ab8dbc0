// lnorm_api.hip -- the C ABI (include/lnorm_hip.h) over the gfx950 kernels.
//
// One call = the reference's calc_Lnorm (L_CUDA.cu:281-359): upload the
// matrix, enumerate, reduce, return the norm and the strategy.  The search
// runs in two phases:
//   1. MODE_MAX over all ranks: every unit (a contiguous range of Gray ranks:
//      a lane's segment for the plain kernels, a wave's range for the blocked
//      kernel) records its best value; the global maximum is an atomic max.
//   2. The first unit whose best equals the maximum is re-walked in
//      MODE_FIRST, which returns the smallest rank attaining it.  That is the
//      strategy the reference prints: within a worker it keeps the first
//      strict improvement (L_OpenMP.c:79) and across workers the lowest
//      index wins (L_OpenMP.c:291-294), and workers own ascending contiguous
//      segments (calc_jmin_jmax, L_OpenMP.c:6-13).
// The strategy is the label vector of that rank: BRGC bits for d <= 2
// (L_OpenMP.c:63), reflected d-ary digits P[(N / d^i) mod 2d] for d >= 3
// (L_OpenMP.c:140,204; functions.h:350-359).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/lnorm_hip.h"
#include "lnorm_colwalk.h"
#include "lnorm_internal.h"

using namespace lnorm;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(LNORM_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

constexpr int kMaxLabels = 10;      // RANK_OF_NORM, functions.h:3
constexpr int kMaxCols = 4096;      // length, functions.h:2
constexpr uint64_t kUnitsPerChunk = 1ull << 22;   // units per launch (16 MiB of unit maxima)
// Streams besides the main one over which the kernel groups of a batch run
// concurrently (each stream at the default priority).
constexpr int kSideStreams = 3;

struct Buf {
  void *p = nullptr;
  size_t cap = 0;
  bool fixed = false;   // a carve-out of the no-allocation scratch: never reallocated
};

struct DevCtx {
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  Buf A, D, units, scal, table, dpk, seed, dfull, lo, state;
  // batch buffers, the side streams the groups of a batch spread over, and
  // the captured launch sequence
  Buf bA, bD, bUnits, bScal, bTasks, bCnt;
  // the int_type = long long variant
  Buf A64, D64, units64, state64;
  hipStream_t side[kSideStreams] = {};
  hipEvent_t bFork[2] = {}, bJoin[2][kSideStreams] = {};
  hipGraphExec_t bGraph = nullptr;
  std::vector<uint64_t> bGraphSig;
  int cus = 0;   // compute units
  bool init = false;
  // units per enumeration launch (the capacity of `units`) and the general
  // kernel's concurrent waves (its per-wave state slices)
  uint64_t unit_chunk = kUnitsPerChunk;
  uint64_t general_slices = kGeneralSlices;
  bool noalloc = false;   // every buffer is a fixed carve-out of `scratch`
};

DevCtx g_ctx[64];

int ensure(Buf &b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.cap >= bytes) return 0;
  if (b.fixed)
    return fail(LNORM_E_SIZE, "no-allocation call: %zu bytes of device scratch needed where %zu are reserved",
                bytes, b.cap);
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) return fail(LNORM_E_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  b.cap = bytes;
  return 0;
}

// The per-unit maxima (one chunk of units) and the scalar results that every
// enumeration writes.  Each run_* function makes sure of them itself: the
// autotuner's probes run before a call's own uploads.
int ensure_result_bufs(DevCtx &c) {
  if (int r = ensure(c.units, c.unit_chunk * 4)) return r;
  return ensure(c.scal, 256);
}

int ctx_init(DevCtx &c) {
  if (c.init) return 0;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  for (auto &e : c.ev) HIPCHK(hipEventCreate(&e));
  for (auto &st : c.side) HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  for (int ph = 0; ph < 2; ++ph) {
    HIPCHK(hipEventCreateWithFlags(&c.bFork[ph], hipEventDisableTiming));
    for (auto &e : c.bJoin[ph]) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  c.init = true;
  return 0;
}

// ---- device memory of the no-allocation entry point ----------------------------
// lnorm_hip_calc_noalloc takes every device buffer of a single-problem call
// from this one array, which each device that loads the library holds until
// the process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: the contexts below keep each device's own address, looked
// up once per device by hipGetSymbolAddress, and only that address reaches
// kernels and copies.  The layout (bytes, in this order):
//   unit maxima     4 MiB   2^20 units per launch (launches are chunked)
//   results         4 KiB   the maximum and the first rank/unit
//   matrix          1 MiB   R x C int32 (general kernel: 63 x 4096 at most)
//   deltas/tables   1 MiB   [2][R-1][1024] deltas of the widest binary walk, or
//                   + 4 KiB the column walk's 2 A rows (63 x 4096) and column 0
//   blocked       468 KiB   int8 planes, low-pattern table, seed, high deltas
//   general state   rest    per-wave label vectors: as many concurrent waves
//                           as whole slices fit (at most kGeneralSlices)
constexpr size_t kScratchBytes = LNORM_NOALLOC_SCRATCH_BYTES;
constexpr uint64_t kNoallocUnits = 1ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

DevCtx g_ctx_noalloc[64];

// Creates the context's stream and events and carves its buffers out of this
// device's copy of `scratch` (the calling thread's current device).
int noalloc_init(DevCtx &c) {
  if (c.noalloc) return 0;
  if (int r = ctx_init(c)) return r;
  void *base = nullptr;
  HIPCHK(hipGetSymbolAddress(&base, HIP_SYMBOL(scratch)));
  if (!base) return fail(LNORM_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
  size_t off = 0;
  auto carve = [&](Buf &b, size_t bytes) {
    b.p = (char *)base + off;
    b.cap = bytes;
    b.fixed = true;
    off += (bytes + 255) & ~(size_t)255;
  };
  carve(c.units, kNoallocUnits * 4);
  carve(c.scal, 4096);
  carve(c.A, (size_t)1 << 20);
  carve(c.D, ((size_t)1 << 20) + (4u << 10));
  carve(c.lo, (size_t)16 << 10);
  carve(c.table, (size_t)256 << 10);
  carve(c.seed, (size_t)4 << 10);
  carve(c.dfull, (size_t)64 << 10);
  carve(c.dpk, (size_t)128 << 10);
  if (off >= kScratchBytes) return fail(LNORM_E_HIP, "scratch layout exceeds %zu bytes", kScratchBytes);
  carve(c.state, kScratchBytes - off);
  c.unit_chunk = kNoallocUnits;
  c.noalloc = true;
  return 0;
}

// ---- problem description ----------------------------------------------------
struct Problem {
  const int32_t *A;
  int R, C, n, marginal;
  Kind kind;
  int base;          // 2 for d <= 2, else d
  uint64_t total;    // ranks the reference enumerates: min(total_num_to_calc, full)
  uint64_t full = 0; // base^(R-1): every labeling
};

uint64_t ipow(uint64_t b, int e) {
  uint64_t r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

int max_rows(int n) {  // calc_Parameters, L_OpenMP.c:354-366
  if (n <= 2) return 63;
  return (int)(std::floor(64.0 / std::log2((double)n)) + 1);
}

// The ranks the reference enumerates.  total_num_to_calc (L_OpenMP.c:352-367)
// is 2^(k-1) for n <= 2 and (unsigned long long) pow(n, k-1) beyond, which is
// inexact from 2^53 on: a count below n^(k-1) leaves the highest ranks out;
// one above it walks on into ranks whose digits below k-1 repeat labelings
// already seen (the changed digit is searched below k-1 only), which never
// pass the strict comparison.  Either way the reference's result is that of
// the ranks [0, min(count, n^(k-1))).  n^(k-1) itself must fit 64 bits: only
// n = 4 at 33 rows reaches 2^64, where the reference's count wraps to 0.
int rank_count(Problem &p) {
  if (p.n <= 2) {
    p.total = p.full = 1ull << (p.R - 1);
    return 0;
  }
  unsigned __int128 f = 1;
  for (int i = 0; i < p.R - 1; ++i) f *= (unsigned)p.n;
  if (f > (unsigned __int128)UINT64_MAX)
    return fail(LNORM_E_SIZE, "%d^%d labelings do not fit 64 bits (the reference's count wraps to 0)", p.n, p.R - 1);
  p.full = (uint64_t)f;
  const double dt = std::pow((double)p.n, (double)(p.R - 1));
  const uint64_t count = dt < 18446744073709551616.0 ? (uint64_t)dt : 0;
  if (count == 0) return fail(LNORM_E_SIZE, "the reference's count of %d^%d labelings is 0", p.n, p.R - 1);
  p.total = std::min(count, p.full);
  return 0;
}

int validate(Problem &p) {
  if (!p.A) return fail(LNORM_E_ARG, "mtx_as_vec is NULL");
  if (p.R < 1 || p.C < 1) return fail(LNORM_E_ARG, "matrix must be at least 1 x 1 (got %d x %d)", p.R, p.C);
  if (p.n < 1 || p.n > kMaxLabels) return fail(LNORM_E_SIZE, "order %d outside 1..%d", p.n, kMaxLabels);
  if (p.marginal != 0 && p.marginal != 1) return fail(LNORM_E_ARG, "marginal must be 0 or 1");
  if (p.marginal && p.n != 1) return fail(LNORM_E_ARG, "marginals are defined for n = 1 only");
  if (p.R > max_rows(p.n))
    return fail(LNORM_E_SIZE, "%d rows exceed maxRows %d for n = %d", p.R, max_rows(p.n), p.n);
  if (p.C > kMaxCols) return fail(LNORM_E_SIZE, "%d columns exceed length %d", p.C, kMaxCols);
  int64_t s = 0;
  for (size_t i = 0; i < (size_t)p.R * p.C; ++i) s += p.A[i] < 0 ? -(int64_t)p.A[i] : (int64_t)p.A[i];
  if (s > INT32_MAX)
    return fail(LNORM_E_BOUND, "sum |A| = %lld exceeds INT32_MAX; the reference's int arithmetic would overflow",
                (long long)s);
  p.kind = p.n == 1 ? (p.marginal ? KIND_LM : KIND_L1) : (p.n == 2 ? KIND_L2 : KIND_LD);
  p.base = p.n <= 2 ? 2 : p.n;
  return rank_count(p);
}

// The ranks [rb, re) a single-problem call enumerates (opts->rank_begin,
// rank_end; rank_end = 0: the reference's count).
int rank_range(const Problem &p, const lnorm_hip_opts &o, uint64_t &rb, uint64_t &re) {
  rb = o.rank_begin;
  re = o.rank_end ? o.rank_end : p.total;
  if (rb >= re || re > p.full)
    return fail(LNORM_E_ARG, "rank range [%llu, %llu) is empty or beyond the %llu ranks", (unsigned long long)rb,
                (unsigned long long)re, (unsigned long long)p.full);
  return 0;
}

// A piece of a rank range: the units [u0, u0 + n) of base^e ranks each.
struct RankRun {
  int e;
  uint64_t u0, n;
};

// Splits [rb, re) into runs of aligned units in rank order: at each rank the
// largest unit base^j (j <= e_main) that starts there and ends within the
// range, consecutive units of one size merged into one run.  The whole
// units of base^e_main form one run; below them at most base - 1 units of
// each smaller size sit on either side of it.
std::vector<RankRun> split_ranks(int base, int e_main, uint64_t rb, uint64_t re) {
  std::vector<RankRun> runs;
  uint64_t a = rb;
  while (a < re) {
    int j = 0;
    uint64_t L = 1;
    while (j < e_main && (a / L) % (uint64_t)base == 0 && L * (uint64_t)base <= re - a) {
      L *= (uint64_t)base;
      ++j;
    }
    const uint64_t n = j == e_main ? (re - a) / L : 1;
    if (!runs.empty() && runs.back().e == j && (runs.back().u0 + runs.back().n) * L == a)
      runs.back().n += n;
    else
      runs.push_back(RankRun{j, a / L, n});
    a += n * L;
  }
  return runs;
}

void labels_of_rank(uint64_t N, int R, int n, int32_t *out) {
  for (int i = 0; i < R - 1; ++i) {
    if (n <= 2) {
      out[i] = (int32_t)(((N ^ (N >> 1)) >> i) & 1);
    } else {
      uint64_t x = N;
      for (int k = 0; k < i; ++k) x /= (uint64_t)n;
      const uint64_t r = x % (uint64_t)(2 * n);
      out[i] = (int32_t)(r < (uint64_t)n ? r : 2 * n - 1 - r);
    }
  }
}

// ---- plain / general launch geometry --------------------------------------------
struct Geometry {
  int q;           // segment = base^q ranks
  uint64_t seglen;
  uint64_t nseg;   // segments covering the range
};

// The lane- or wave-per-segment walks.
enum Walk { WALK_PLAIN = 0, WALK_PLAIN16 = 1, WALK_GENERAL = 2, WALK_WIDE = 3, WALK_WIDE16 = 4, WALK_TAB16 = 5,
            WALK_TAB16W = 6, WALK_TAB32 = 7, WALK_TAB32W = 8, WALK_COLS = 9, WALK_COLSD = 10 };
// WALK_COLS: the column walk's L1/LM form (binary counter, segments up to
// 2^30); WALK_COLSD: its d-ary form, L2 included (nibble counter, at most
// 15 digits per segment)

// The tabulated walks in 32-bit words: one column per word (ny_pad = words).
bool is_tab32(Walk w) { return w == WALK_TAB32 || w == WALK_TAB32W; }

// Segments for a range of base^e ranks: aim at about 2^target_log lanes.
// The step counter of the d-ary and packed walks holds 15 base-d digits.
// qmin: the walk's smallest segment (the tabulated walk's table rows).
Geometry geometry_for(int base, int e, int target_log, int forced, Walk walk, int qmin = 0) {
  Geometry g;
  const double lb = std::log2((double)base);
  int qmax;
  if (walk == WALK_GENERAL) qmax = (int)(62.0 / lb);
  else if ((walk == WALK_PLAIN || walk == WALK_WIDE || walk == WALK_COLS) && base == 2) qmax = 30;
  else qmax = std::min(15, (int)(31.0 / lb));
  if (walk == WALK_COLS || walk == WALK_COLSD) target_log = std::min(target_log, 16);   // one wave per segment
  int q = forced > 0 ? forced : std::max(0, e - (int)std::floor(target_log / lb));
  q = std::min(q, std::min(qmax, e));
  if (walk == WALK_WIDE) q = std::max(q, std::min(3, e));   // steps combine in groups of 8
  if (walk == WALK_WIDE16) q = std::max(q, std::min(2, e));
  q = std::max(q, std::min(qmin, e));
  g.q = q;
  g.seglen = ipow((uint64_t)base, q);
  g.nseg = ipow((uint64_t)base, e - q);
  return g;
}

struct Prepared {
  Walk walk = WALK_PLAIN;
  int ny_pad = 0;      // kernel width in columns (packed walks: 2 x words; 32-bit tabulated: words)
  int qmin = 0;        // smallest segment exponent (tabulated walk: its table rows)
  std::vector<int32_t> A, D;
};

int64_t max_col_abs_sum(const Problem &p, int first_col) {
  int64_t mx = 0;
  for (int l = first_col; l < p.C; ++l) {
    int64_t cs = 0;
    for (int i = 0; i < p.R; ++i) cs += std::llabs((long long)p.A[(size_t)i * p.C + l]);
    mx = std::max(mx, cs);
  }
  return mx;
}

// Tables of the tabulated walks in 32-bit words (walk_blk16 / walk_tab16_bin
// with WB = 32).  NW = ny_pad words, one column each, padding columns 0.
//   n = 1: A = R x NW |.| columns, R x NW of 2 A, (LM) the R column-0
//          entries; D = 2^M x NW words 2^31 - A_S (A_S = sum_{i<M} s_i A_i,
//          s_i = 2 S_i - 1), (LM) the 2^M column-0 sums A_S,0.
//   n >= 2: A = R x NW columns; D = 2^M x NW words 2^31 - sum_{i in S} A_i.
void prepare_tab32(const Problem &p, Walk walk, int NW, Prepared &pp) {
  const int marg = p.kind == KIND_LM ? 1 : 0;
  int M;
  if (p.n == 1) M = walk == WALK_TAB32W ? wtb32_rows(NW) : tab32_bin_rows(NW);
  else M = walk == WALK_TAB32W ? wt32_rows(p.n, NW) : blk32_rows(p.n, NW);
  pp.qmin = M;
  const int C = p.C - marg;   // the columns held in words
  auto at = [&](int i, int c) { return p.A[(size_t)i * p.C + marg + c]; };
  // signed (n = 1) or subset (n >= 2) sums of the M fastest rows, each from
  // the pattern without its lowest row: one add per column.  Every sum is
  // bounded by sum |A| <= INT32_MAX; the adds wrap (unsigned) on the way.
  const int VC = p.n == 1 ? p.C : C;   // n = 1 keeps column 0 too (LM's marginal column)
  std::vector<uint32_t> val((size_t)VC << M, 0u);
  if (p.n == 1)
    for (int l = 0; l < p.C; ++l) {
      uint32_t sum = 0;
      for (int i = 0; i < M; ++i) sum -= (uint32_t)p.A[(size_t)i * p.C + l];
      val[l] = sum;
    }
  for (int S = 1; S < (1 << M); ++S) {
    const uint32_t *prev = val.data() + (size_t)(S & (S - 1)) * VC;
    uint32_t *cur = val.data() + (size_t)S * VC;
    const int i = __builtin_ctz(S);
    for (int l = 0; l < VC; ++l) {
      const uint32_t a = (uint32_t)(p.n == 1 ? p.A[(size_t)i * p.C + l] : at(i, l));
      cur[l] = prev[l] + (p.n == 1 ? 2u * a : a);
    }
  }
  if (p.n == 1) {
    pp.A.assign((size_t)2 * p.R * NW + (marg ? p.R : 0), 0);
    for (int i = 0; i < p.R; ++i) {
      for (int c = 0; c < C; ++c) {
        pp.A[(size_t)i * NW + c] = at(i, c);
        pp.A[(size_t)(p.R + i) * NW + c] = (int32_t)(2u * (uint32_t)at(i, c));
      }
      if (marg) pp.A[(size_t)2 * p.R * NW + i] = p.A[(size_t)i * p.C];
    }
    pp.D.assign(((size_t)NW << M) + (marg ? (1u << M) : 0), (int32_t)kBias32);
    for (int S = 0; S < (1 << M); ++S) {
      for (int c = 0; c < C; ++c) pp.D[(size_t)S * NW + c] = (int32_t)(kBias32 - val[(size_t)S * VC + marg + c]);
      if (marg) pp.D[((size_t)NW << M) + S] = (int32_t)val[(size_t)S * VC];
    }
  } else {
    pp.A.assign((size_t)p.R * NW, 0);
    for (int i = 0; i < p.R; ++i)
      for (int c = 0; c < C; ++c) pp.A[(size_t)i * NW + c] = at(i, c);
    pp.D.assign((size_t)NW << M, (int32_t)kBias32);
    for (int S = 0; S < (1 << M); ++S)
      for (int c = 0; c < C; ++c) pp.D[(size_t)S * NW + c] = (int32_t)(kBias32 - val[(size_t)S * VC + c]);
  }
}

void prepare_walk(const Problem &p, Walk walk, int ny_pad, Prepared &pp) {
  pp.walk = walk;
  pp.ny_pad = walk == WALK_GENERAL ? 0 : ny_pad;
  pp.qmin = 0;
  pp.D.clear();
  if (is_tab32(walk)) {
    prepare_tab32(p, walk, ny_pad, pp);
    return;
  }
  if ((walk == WALK_TAB16 || walk == WALK_TAB16W) && p.n == 1) {
    // L1 / LM sign-pattern table (walk_tab16_bin, k_wide_tb16): A = R x NW
    // packed |.| columns, R x NW packed 2A, (LM) R column-0 entries; D = 2^M x
    // NW words -A_S + 2^15, (LM) 2^M column-0 sums A_S,0
    const int marg = p.kind == KIND_LM ? 1 : 0;
    const int NW = ny_pad / 2, M = walk == WALK_TAB16W ? wtb16_rows(NW) : tab_bin_rows(NW);
    auto pack = [&](int32_t *dst, int i, int f) {
      for (int l = marg; l < p.C; ++l) {
        const uint32_t v = (uint32_t)(f * p.A[(size_t)i * p.C + l]) & 0xFFFFu;
        dst[(l - marg) / 2] |= (int32_t)(v << (16 * ((l - marg) & 1)));
      }
    };
    pp.A.assign((size_t)2 * p.R * NW + (marg ? p.R : 0), 0);
    for (int i = 0; i < p.R; ++i) {
      pack(pp.A.data() + (size_t)i * NW, i, 1);
      pack(pp.A.data() + (size_t)(p.R + i) * NW, i, 2);
      if (marg) pp.A[(size_t)2 * p.R * NW + i] = p.A[(size_t)i * p.C];
    }
    pp.D.assign(((size_t)NW << M) + (marg ? (1u << M) : 0), 0);
    // signed sums of the M fastest rows, each pattern from the one without its
    // lowest set bit: val[S] = val[S & (S - 1)] + 2 A[ctz S]
    std::vector<int32_t> val((size_t)p.C << M);
    for (int l = 0; l < p.C; ++l) {
      int32_t sum = 0;
      for (int i = 0; i < M; ++i) sum -= p.A[(size_t)i * p.C + l];
      val[l] = sum;
    }
    for (int S = 1; S < (1 << M); ++S) {
      const int32_t *prev = val.data() + (size_t)(S & (S - 1)) * p.C, *row = p.A + (size_t)__builtin_ctz(S) * p.C;
      for (int l = 0; l < p.C; ++l) val[(size_t)S * p.C + l] = prev[l] + 2 * row[l];
    }
    for (int S = 0; S < (1 << M); ++S) {
      for (int l = 0; l < p.C; ++l) {
        const int32_t sum = val[(size_t)S * p.C + l];
        if (marg && l == 0) {
          pp.D[((size_t)NW << M) + S] = sum;
        } else {
          const int c = l - marg;
          const uint32_t v = (uint32_t)(kBias16 - sum) & 0xFFFFu;
          pp.D[(size_t)S * NW + c / 2] |= (int32_t)(v << (16 * (c & 1)));
        }
      }
      for (int c = p.C - marg; c < 2 * NW; ++c)   // padding columns: |0 + 0|
        pp.D[(size_t)S * NW + c / 2] |= (int32_t)(kBias16 << (16 * (c & 1)));
    }
    pp.qmin = M;
    return;
  }
  if (walk == WALK_PLAIN16 || walk == WALK_WIDE16 || walk == WALK_TAB16 || walk == WALK_TAB16W) {
    // two columns per word, low half = even column
    const int NW = ny_pad / 2;
    pp.A.assign((size_t)p.R * NW, 0);
    for (int i = 0; i < p.R; ++i)
      for (int l = 0; l < p.C; ++l) {
        const uint32_t v = (uint32_t)p.A[(size_t)i * p.C + l] & 0xFFFFu;
        pp.A[(size_t)i * NW + l / 2] |= (int32_t)(v << (16 * (l & 1)));
      }
    if (walk == WALK_TAB16 || walk == WALK_TAB16W) {
      // subset table of the M fastest rows: word w of subset S holds
      // -sum_{i in S} A[i][col] + 2^15 for its two columns
      const int M = walk == WALK_TAB16W ? wt16_rows(p.n, NW) : blk16_rows(p.n, NW);
      pp.qmin = M;
      pp.D.assign((size_t)NW << M, 0);
      // subset sums, each from the subset without its lowest row:
      // sum[S] = sum[S & (S - 1)] + A[ctz S]; padding columns stay 0
      std::vector<int32_t> sum((size_t)(2 * NW) << M, 0);
      for (int S = 1; S < (1 << M); ++S) {
        const int32_t *prev = sum.data() + (size_t)(S & (S - 1)) * 2 * NW, *row = p.A + (size_t)__builtin_ctz(S) * p.C;
        int32_t *cur = sum.data() + (size_t)S * 2 * NW;
        for (int l = 0; l < p.C; ++l) cur[l] = prev[l] + row[l];
      }
      for (int S = 0; S < (1 << M); ++S)
        for (int l = 0; l < 2 * NW; ++l) {
          const uint32_t v = (uint32_t)(kBias16 - sum[(size_t)S * 2 * NW + l]) & 0xFFFFu;
          pp.D[(size_t)S * NW + l / 2] |= (int32_t)(v << (16 * (l & 1)));
        }
    }
    return;
  }
  if (walk == WALK_COLS) {
    // the column walk, L1 / LM: A = R x Cp rows of the |.| columns, D = R x Cp
    // rows of 2 A, then the R entries of column 0 (LM's marginal)
    const int marg = p.kind == KIND_LM ? 1 : 0;
    pp.A.assign((size_t)p.R * ny_pad, 0);
    pp.D.assign((size_t)p.R * ny_pad + p.R, 0);
    for (int i = 0; i < p.R; ++i) {
      for (int l = marg; l < p.C; ++l) {
        const int32_t a = p.A[(size_t)i * p.C + l];
        pp.A[(size_t)i * ny_pad + l - marg] = a;
        pp.D[(size_t)i * ny_pad + l - marg] = (int32_t)(2u * (uint32_t)a);
      }
      pp.D[(size_t)p.R * ny_pad + i] = marg ? p.A[(size_t)i * p.C] : 0;
    }
    return;
  }
  const int W = (walk == WALK_PLAIN || walk == WALK_WIDE || walk == WALK_COLSD) ? ny_pad : p.C;
  pp.A.assign((size_t)p.R * W, 0);
  for (int i = 0; i < p.R; ++i)
    for (int l = 0; l < p.C; ++l) pp.A[(size_t)i * W + l] = p.A[(size_t)i * p.C + l];
  if ((walk == WALK_PLAIN || walk == WALK_WIDE) && p.n <= 2 && p.R > 1) {
    // D[1][j] = +f A[j], D[0][j] = -f A[j]; f = 2 for L1/LM (y += 2A), 1 for L2
    const uint32_t f = p.n == 1 ? 2u : 1u;
    pp.D.assign((size_t)2 * (p.R - 1) * W, 0);
    for (int j = 0; j < p.R - 1; ++j)
      for (int l = 0; l < W; ++l) {
        const uint32_t a = (uint32_t)pp.A[(size_t)j * W + l] * f;
        pp.D[((size_t)(p.R - 1) + j) * W + l] = (int32_t)a;
        pp.D[(size_t)j * W + l] = (int32_t)(0u - a);
      }
  }
}

// The shortest segment exponent a walk runs: its table rows, or the W-wave
// walks' shortest segments (base^3 ranks binary, base^2 packed).
int walk_qmin(const Prepared &pp) {
  if (pp.walk == WALK_WIDE) return 3;
  if (pp.walk == WALK_WIDE16) return 2;
  return pp.qmin;
}

hipError_t launch_walk(const Problem &p, const Prepared &pp, Mode mode, const PlainArgs &a, hipStream_t st,
                       int block) {
  switch (pp.walk) {
    case WALK_PLAIN: return launch_plain(p.kind, mode, a, st, block);
    case WALK_PLAIN16: return launch_plain16(mode, a, st, block);
    case WALK_WIDE: return launch_wide(p.kind, mode, a, st);
    case WALK_WIDE16: return launch_wide16(mode, a, st);
    case WALK_TAB16: return p.n == 1 ? launch_tab16_bin(p.kind, mode, a, st, block) : launch_blk16(mode, a, st, block);
    case WALK_TAB16W: return p.n == 1 ? launch_wide_tb16(p.kind, mode, a, st) : launch_wide_t16(mode, a, st);
    case WALK_TAB32: return p.n == 1 ? launch_tab32_bin(p.kind, mode, a, st, block) : launch_blk32(mode, a, st, block);
    case WALK_TAB32W: return p.n == 1 ? launch_wide_tb32(p.kind, mode, a, st) : launch_wide_t32(mode, a, st);
    case WALK_COLS:
    case WALK_COLSD: return launch_colwalk(p.kind, mode, a, st);
    default: return launch_general(p.kind, mode, a, p.C, st);
  }
}

// Runs MODE_MAX over [0, total) with the plain or general kernel, chunked.
// Returns the maximum and the first unit attaining it.
struct Phase1Out {
  int32_t best = INT32_MIN;
  uint64_t first_unit = 0;
  uint64_t units = 0;
  double ms_kernel = 0;   // HIP-event time of the enumeration kernel launches alone
};

// max_units > 0 limits the walk to the first max_units units (the
// autotuner's probes); the result is then a partial maximum.  unit0: the
// first unit's index (the walk covers units [unit0, unit0 + g.nseg)); the
// first unit returned is absolute.
int run_walk_max(DevCtx &cx, const Problem &p, const Prepared &pp, const Geometry &g, hipStream_t st,
                 int block, Phase1Out &out, uint64_t max_units = 0, uint64_t chunk_units = 0,
                 uint64_t unit0 = 0) {
  if (int r = ensure_result_bufs(cx)) return r;
  const uint64_t chunk = chunk_units ? std::min(chunk_units, cx.unit_chunk) : cx.unit_chunk;
  const uint64_t nunits = max_units ? std::min(max_units, g.nseg) : g.nseg;
  out.units = nunits;
  out.best = INT32_MIN;
  out.first_unit = ~0ull;
  int32_t *ub = (int32_t *)cx.units.p;
  int32_t *gbest = (int32_t *)cx.scal.p;
  unsigned long long *gfirst = (unsigned long long *)((char *)cx.scal.p + 64);
  for (uint64_t c0 = 0; c0 < nunits; c0 += chunk) {
    const uint64_t n = std::min(chunk, nunits - c0);
    HIPCHK(launch_fill_i32(gbest, INT32_MIN, 1, st));
    HIPCHK(launch_fill_u64(gfirst, ~0ull, 1, st));
    PlainArgs a{};
    a.A = (const int32_t *)cx.A.p;
    a.D = (const int32_t *)cx.D.p;
    a.R = p.R;
    a.ny_pad = pp.ny_pad;
    a.d = p.n;
    a.q = g.q;
    a.seg0 = unit0 + c0;
    a.nseg = n;
    a.unit_best = ub;
    a.gbest = gbest;
    a.gfirst = gfirst;
    a.state = (uint32_t *)cx.state.p;
    a.slices = cx.general_slices;
    HIPCHK(hipEventRecord(cx.ev[4], st));
    hipError_t e = launch_walk(p, pp, MODE_MAX, a, st, block);
    HIPCHK(hipEventRecord(cx.ev[5], st));
    if (e != hipSuccess) return fail(LNORM_E_HIP, "enumeration launch failed: %s", hipGetErrorString(e));
    HIPCHK(launch_first_unit(ub, n, gbest, gfirst, st));
    int32_t hb;
    unsigned long long hf;
    HIPCHK(hipMemcpyAsync(&hb, gbest, sizeof hb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&hf, gfirst, sizeof hf, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float kms = 0;
    HIPCHK(hipEventElapsedTime(&kms, cx.ev[4], cx.ev[5]));
    out.ms_kernel += kms;
    if (hb > out.best || out.first_unit == ~0ull) {
      out.best = hb;
      out.first_unit = unit0 + c0 + hf;
    }
  }
  return 0;
}

// MODE_FIRST over the rank range [u * base^e, (u+1) * base^e).
int run_walk_first(DevCtx &cx, const Problem &p, const Prepared &pp, uint64_t u, int e, int32_t target,
                   hipStream_t st, int block, uint64_t &rank) {
  if (int r = ensure_result_bufs(cx)) return r;
  const Geometry g = geometry_for(p.base, e, 16, 0, pp.walk, pp.qmin);
  unsigned long long *gfirst = (unsigned long long *)((char *)cx.scal.p + 64);
  HIPCHK(launch_fill_u64(gfirst, ~0ull, 1, st));
  PlainArgs a{};
  a.A = (const int32_t *)cx.A.p;
  a.D = (const int32_t *)cx.D.p;
  a.R = p.R;
  a.ny_pad = pp.ny_pad;
  a.d = p.n;
  a.q = g.q;
  a.seg0 = u * g.nseg;
  a.nseg = g.nseg;
  a.target = target;
  a.unit_best = (int32_t *)cx.units.p;
  a.gbest = (int32_t *)cx.scal.p;
  a.gfirst = gfirst;
  a.state = (uint32_t *)cx.state.p;
  a.slices = cx.general_slices;
  hipError_t err = launch_walk(p, pp, MODE_FIRST, a, st, block);
  if (err != hipSuccess) return fail(LNORM_E_HIP, "strategy launch failed: %s", hipGetErrorString(err));
  unsigned long long hf;
  HIPCHK(hipMemcpyAsync(&hf, gfirst, sizeof hf, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (hf == ~0ull) return fail(LNORM_E_HIP, "strategy recovery found no rank attaining %d", target);
  rank = hf;
  return 0;
}

// ---- blocked L1/LM -----------------------------------------------------------
// Two word formats: 16-bit (two biased halves per word, v_sad_u16; every |.|
// column's absolute sum <= 2^15 - 1, at most 64 |.| columns) and 32-bit
// (one word per column, v_sad_u32, up to 128 |.| columns over W waves).
struct BlockedPlan {
  bool wide = false;
  int P = 0, kl = 0, nh = 0, np = 0, NW = 0, s = 0, planes = 1;
  int W = 1, NC = 0;       // wide: waves per unit, words per wave (np = W * NC)
  uint64_t nunits = 0;     // waves (16-bit) or workgroups (32-bit)
};

// Signed-digit int8 planes of the low rows: v = 256 * hi + lo, lo in [-128, 127].
// Returns 1 or 2 planes, or 0 when an entry needs more than two.
int low_planes(const Problem &p, int kl) {
  int planes = 1;
  for (int i = 0; i < kl; ++i)
    for (int l = 0; l < p.C; ++l) {
      const int32_t v = p.A[(size_t)i * p.C + l];
      if (v < -128 || v > 127) planes = 2;
      if (v < -32768 + 128 || v > 32767 - 128) return 0;
    }
  return planes;
}

void split_planes(const Problem &p, int kl, int planes, std::vector<int8_t> &lo) {
  lo.assign((size_t)planes * kl * p.C, 0);
  for (int i = 0; i < kl; ++i)
    for (int l = 0; l < p.C; ++l) {
      const int32_t v = p.A[(size_t)i * p.C + l];
      if (planes == 1) {
        lo[(size_t)i * p.C + l] = (int8_t)v;
      } else {
        const int32_t lo8 = ((v + 128) & 255) - 128;
        lo[(size_t)i * p.C + l] = (int8_t)((v - lo8) / 256);
        lo[((size_t)kl + i) * p.C + l] = (int8_t)lo8;
      }
    }
}

int log2i(int x) {
  int k = 0;
  while ((1 << (k + 1)) <= x) ++k;
  return k;
}

bool blocked_plan(const Problem &p, const lnorm_hip_opts &o, BlockedPlan &bp, std::string &why) {
  if (p.kind != KIND_L1 && p.kind != KIND_LM) { why = "blocked kernels handle n = 1 only"; return false; }
  const int marg = p.kind == KIND_LM ? 1 : 0;
  const int nabs = p.C - marg;
  int64_t maxcs = 0;
  for (int l = marg; l < p.C; ++l) {
    int64_t cs = 0;
    for (int i = 0; i < p.R; ++i) cs += std::llabs((long long)p.A[(size_t)i * p.C + l]);
    maxcs = std::max(maxcs, cs);
  }
  const int bits = o.word_bits;
  if (bits != 0 && bits != 16 && bits != 32) { why = "word_bits must be 0, 16 or 32"; return false; }
  bool wide;
  if (bits == 16) {
    if (maxcs > 32767) { why = "a column's absolute sum exceeds 32767 (16-bit halves)"; return false; }
    if (nabs > 64) { why = "more than 64 |.| columns for 16-bit words"; return false; }
    wide = false;
  } else if (bits == 32) {
    wide = true;
  } else {
    wide = maxcs > 32767 || nabs > 64;
  }
  int P = o.patterns_per_lane;
  if (!wide) {
    const int np = (nabs + 1) / 2;
    if (P == 0) {
      P = 4;
      const int npp = blocked_np_pad(np, P, marg);
      if (!npp || P * (npp + marg) > 132) P = 2;
    }
    const int npp = blocked_np_pad(np, P, marg);
    if (!npp) { why = "no 16-bit blocked instantiation for this width / P"; return false; }
    bp.np = npp;
    bp.NW = npp + marg;
  } else {
    if (nabs > 128) { why = "more than 128 |.| columns"; return false; }
    const int W = nabs <= 32 ? 1 : (nabs <= 64 ? 2 : 4);
    const int NC = ((nabs + W - 1) / W + 7) / 8 * 8;
    // measured on 128 columns (W = 4): P = 2 and P = 4 within 1%, P = 1 30% slower
    if (P == 0) P = W == 4 ? 2 : (4 * NC <= 128 ? 4 : 2);
    if (!blocked32_supported(NC, P, W, marg)) { why = "no 32-bit blocked instantiation for this width / P"; return false; }
    bp.W = W;
    bp.NC = NC;
    bp.np = W * NC;
    bp.NW = bp.np + marg;
  }
  bp.wide = wide;
  if (!(P == 1 || P == 2 || P == 4 || P == 8)) { why = "patterns per lane must be 1, 2, 4 or 8"; return false; }
  const int kl = 6 + log2i(P);
  const int min_nh = wide ? 3 : 1;
  if (p.R - 1 < kl + min_nh) { why = "too few rows for the blocked split"; return false; }
  bp.planes = low_planes(p, kl);
  if (!bp.planes) { why = "low-row entries beyond the two int8 planes of int16"; return false; }
  bp.P = P;
  bp.kl = kl;
  bp.nh = p.R - 1 - kl;
  // about 2^19 waves of 2^12-ish steps: many rounds of the 1024 SIMDs keep
  // the tail of the last round short (measured: 2^12 steps per wave run 1.3%
  // faster than 2^16 on a 2^39 enumeration)
  const int wl = wide ? log2i(bp.W) : 0;
  int s = o.wave_log2 > 0 ? o.wave_log2 : std::max(0, bp.nh - 19 + wl);
  s = std::min(std::min(s, bp.nh), 30);
  if (wide) s = std::max(s, 3);
  bp.s = s;
  bp.nunits = 1ull << (bp.nh - s);
  return true;
}

// The blocked walk over waves [wave_begin, wave_begin + min(max_units, all)).
int run_blocked(DevCtx &cx, const Problem &p, const BlockedPlan &bp, hipStream_t st, Phase1Out &out,
                uint64_t max_units = 0, uint64_t wave_begin = 0) {
  if (int r = ensure_result_bufs(cx)) return r;
  const int marg = p.kind == KIND_LM ? 1 : 0;
  const int NW = bp.NW, kl = bp.kl, nh = bp.nh;
  auto Aat = [&](int i, int l) -> int64_t { return l < 0 ? 0 : (int64_t)p.A[(size_t)i * p.C + l]; };
  std::vector<int8_t> lo;
  split_planes(p, kl, bp.planes, lo);
  // scalar state: all high rows at -1, and the flip deltas
  std::vector<uint32_t> seed(NW), dfull((size_t)std::max(nh, 1) * NW), dpk((size_t)2 * std::max(nh, 1) * NW);
  auto yminus = [&](int l) -> int64_t {
    if (l < 0) return 0;
    int64_t y = Aat(p.R - 1, l);
    for (int m = 0; m < nh; ++m) y -= Aat(kl + m, l);
    return y;
  };
  if (!bp.wide) {
    // word c packs |.| columns (marg + 2c, marg + 2c + 1); -1 = zero padding
    auto colof = [&](int c, int h) -> int {
      const int l = marg + 2 * c + h;
      return l < p.C ? l : -1;
    };
    for (int c = 0; c < bp.np; ++c) {
      const int64_t ylo = yminus(colof(c, 0)), yhi = yminus(colof(c, 1));
      seed[c] = ((uint32_t)(32768 - yhi) << 16) | ((uint32_t)(32768 - ylo) & 0xFFFFu);
      for (int m = 0; m < nh; ++m) {
        const int64_t dlo = 2 * Aat(kl + m, colof(c, 0)), dhi = 2 * Aat(kl + m, colof(c, 1));
        dfull[(size_t)m * NW + c] = (uint32_t)(int64_t)(-(dhi * 65536 + dlo));
      }
    }
  } else {
    // word c holds |.| column marg + c: 2^31 - y_H
    for (int c = 0; c < bp.np; ++c) {
      const int l = marg + c < p.C ? marg + c : -1;
      seed[c] = kBias32 - (uint32_t)(int64_t)yminus(l);
      for (int m = 0; m < nh; ++m) dfull[(size_t)m * NW + c] = (uint32_t)(int64_t)(-2 * Aat(kl + m, l));
    }
  }
  if (marg) {
    seed[bp.np] = (uint32_t)(int32_t)yminus(0);
    for (int m = 0; m < nh; ++m) dfull[(size_t)m * NW + bp.np] = (uint32_t)(2 * Aat(kl + m, 0));
  }
  for (int j = 0; j < nh; ++j)
    for (int c = 0; c < NW; ++c) {
      dpk[((size_t)nh + j) * NW + c] = dfull[(size_t)j * NW + c];
      dpk[(size_t)j * NW + c] = 0u - dfull[(size_t)j * NW + c];
    }
  if (int r = ensure(cx.lo, lo.size())) return r;
  if (int r = ensure(cx.table, (size_t)bp.P * NW * 64 * 4)) return r;
  if (int r = ensure(cx.seed, seed.size() * 4)) return r;
  if (int r = ensure(cx.dfull, dfull.size() * 4)) return r;
  if (int r = ensure(cx.dpk, dpk.size() * 4)) return r;
  HIPCHK(hipMemcpyAsync(cx.lo.p, lo.data(), lo.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(cx.seed.p, seed.data(), seed.size() * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(cx.dfull.p, dfull.data(), dfull.size() * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(cx.dpk.p, dpk.data(), dpk.size() * 4, hipMemcpyHostToDevice, st));
  TableArgs ta{};
  ta.lo = (const int8_t *)cx.lo.p;
  ta.planes = bp.planes;
  ta.kl = kl;
  ta.ny = p.C;
  ta.marginal = marg;
  ta.np = bp.np;
  ta.P = bp.P;
  ta.out = (uint32_t *)cx.table.p;
  ta.wide = bp.wide ? 1 : 0;
  HIPCHK(launch_table(ta, st));

  // one 64-bit key per launch: its maximum and first wave (blocked_key)
  unsigned long long *gkey = (unsigned long long *)((char *)cx.scal.p + 128);
  const uint64_t nunits = max_units ? std::min(max_units, bp.nunits) : bp.nunits;
  out.units = nunits;
  out.best = INT32_MIN;
  out.first_unit = ~0ull;
  for (uint64_t c0 = 0; c0 < nunits; c0 += cx.unit_chunk) {
    const uint64_t n = std::min(cx.unit_chunk, nunits - c0);
    HIPCHK(launch_fill_u64(gkey, 0ull, 1, st));
    BlockedArgs a{};
    a.table = (const uint32_t *)cx.table.p;
    a.dpk = (const uint32_t *)cx.dpk.p;
    a.seed = (const uint32_t *)cx.seed.p;
    a.dfull = (const uint32_t *)cx.dfull.p;
    a.nh = nh;
    a.s = bp.s;
    a.wave0 = wave_begin + c0;
    a.nwaves = n;
    a.gkey = gkey;
    HIPCHK(hipEventRecord(cx.ev[4], st));
    hipError_t e = bp.wide ? launch_blocked32(bp.NC, bp.P, bp.W, marg, a, st)
                           : launch_blocked(bp.np, bp.P, marg, a, st);
    HIPCHK(hipEventRecord(cx.ev[5], st));
    if (e != hipSuccess) return fail(LNORM_E_HIP, "blocked launch failed: %s", hipGetErrorString(e));
    unsigned long long hk = 0;
    HIPCHK(hipMemcpyAsync(&hk, gkey, sizeof hk, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (hk == 0) return fail(LNORM_E_HIP, "blocked launch left no result");
    const int32_t hb = blocked_key_value(hk);
    const uint64_t hf = blocked_key_wave(hk);
    float kms = 0;
    HIPCHK(hipEventElapsedTime(&kms, cx.ev[4], cx.ev[5]));
    out.ms_kernel += kms;
    if (hb > out.best || out.first_unit == ~0ull) {
      out.best = hb;
      out.first_unit = wave_begin + c0 + hf;
    }
  }
  return 0;
}

int select_device(const lnorm_hip_opts &o, int &dev) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count < 1)
    return fail(LNORM_E_HIP, "no HIP device available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
  if (o.device >= 0) {
    if (o.device >= count || o.device >= 64) return fail(LNORM_E_ARG, "device %d out of range", o.device);
    HIPCHK(hipSetDevice(o.device));
    dev = o.device;
  } else {
    HIPCHK(hipGetDevice(&dev));
  }
  return 0;
}

// ---- kernel choice ------------------------------------------------------------
struct Choice {
  int kernel = LNORM_KERNEL_AUTO;
  BlockedPlan bp;
  int ny_plain = 0, w16 = 0, ny_wide = 0, w16_wide = 0, wtab = 0, wtabw = 0;
  int wtab32 = 0, wtab32w = 0;   // the 32-bit tabulated walks: one wave or W waves (words = columns)
  int cols = 0;                  // the column walk: 64 x columns per lane (0: none)
};

int choose_kernel(const Problem &p, const lnorm_hip_opts &o, Choice &c) {
  std::string why;
  int kernel = o.kernel;
  const bool can_block =
      (kernel == LNORM_KERNEL_AUTO || kernel == LNORM_KERNEL_BLOCKED) && blocked_plan(p, o, c.bp, why);
  c.ny_plain = plain_ny_pad(p.kind, p.n, p.C);
  const bool fits16 = p.n >= 2 && max_col_abs_sum(p, 0) <= 32767;
  c.w16 = fits16 ? plain16_words(p.n, p.C) : 0;
  // the W-wave walk combines steps in groups of 8: at least 3 enumerated rows
  c.ny_wide = (p.R - 1 >= 3) ? wide_ny_pad(p.kind, p.C) : 0;
  c.w16_wide = (fits16 && p.n >= 3) ? wide16_words(p.n, p.C) : 0;
  if (p.n == 1) {
    const int marg = p.kind == KIND_LM ? 1 : 0;
    c.wtab = max_col_abs_sum(p, marg) <= 32767 ? tab16_bin_words(p.C - marg) : 0;
    if (c.wtab && p.R - 1 < tab_bin_rows(c.wtab)) c.wtab = 0;
    // beyond one wave's table: the W-wave tabulated walk
    if (max_col_abs_sum(p, marg) <= 32767 && !tab16_bin_words(p.C - marg)) {
      c.wtabw = wtb16_words(p.C - marg);
      if (c.wtabw && p.R - 1 < wtb16_rows(c.wtabw)) c.wtabw = 0;
    }
  } else {
    c.wtab = (fits16 && p.n >= 2) ? blk16_words(p.n, p.C) : 0;
    if (c.wtab && p.R - 1 < blk16_rows(p.n, c.wtab)) c.wtab = 0;   // fewer enumerated rows than table rows
    // beyond one wave's tables: the W-wave tabulated walk
    if (fits16 && p.n >= 2 && !blk16_words(p.n, p.C)) {
      c.wtabw = wt16_words(p.n, p.C);
      if (c.wtabw && p.R - 1 < wt16_rows(p.n, c.wtabw)) c.wtabw = 0;
    }
  }
  // the 32-bit tabulated walks: any int32 input, at widths of up to 192 |.|
  // columns in one wave and 1024 over W waves (n = 1), up to 96 (d = 2) ..
  // 16 (d = 10) columns in one wave and 256 (d <= 6) .. 64 (d = 10) over W
  // waves (n >= 2), with more enumerated rows than table rows
  {
    const int marg = p.kind == KIND_LM ? 1 : 0;
    if (p.n == 1) {
      c.wtab32 = tab32_bin_words(p.C - marg);
      if (c.wtab32 && p.R - 1 < tab32_bin_rows(c.wtab32)) c.wtab32 = 0;
      if (!tab32_bin_words(p.C - marg)) {
        c.wtab32w = wtb32_words(p.C - marg);
        if (c.wtab32w && p.R - 1 < wtb32_rows(c.wtab32w)) c.wtab32w = 0;
      }
    } else {
      c.wtab32 = blk32_words(p.n, p.C);
      if (c.wtab32 && p.R - 1 < blk32_rows(p.n, c.wtab32)) c.wtab32 = 0;
      if (!blk32_words(p.n, p.C)) {
        c.wtab32w = wt32_words(p.n, p.C);
        if (c.wtab32w && p.R - 1 < wt32_rows(p.n, c.wtab32w)) c.wtab32w = 0;
      }
    }
  }
  // the column walk: any int32 input whose width has an instantiation and
  // whose rows fit the LDS of one workgroup
  {
    const int ncols = p.C - (p.kind == KIND_LM ? 1 : 0);
    const int cpl = ncols > 0 ? colwalk_cpl(p.n, ncols) : 0;
    if (cpl && colwalk_lds_bytes(p.R, 64 * cpl) <= kColwalkLdsMax) c.cols = 64 * cpl;
  }
  if (kernel == LNORM_KERNEL_BLOCKED && !can_block)
    return fail(LNORM_E_UNSUPPORTED, "blocked kernel not applicable: %s", why.c_str());
  if (kernel == LNORM_KERNEL_COLUMNS && !c.cols)
    return fail(LNORM_E_UNSUPPORTED, "column walk needs an instantiated width (%d columns at n = %d) whose rows "
                "fit one workgroup's LDS", p.C, p.n);
  if (kernel == LNORM_KERNEL_TABLE32 && !c.wtab32 && !c.wtab32w)
    return fail(LNORM_E_UNSUPPORTED, "32-bit tabulated walk needs an instantiated width and more enumerated rows "
                "than table rows (%d x %d at n = %d)", p.R, p.C, p.n);
  if (kernel == LNORM_KERNEL_PLAIN && !c.ny_plain)
    return fail(LNORM_E_UNSUPPORTED, "no plain instantiation for %d columns at n = %d", p.C, p.n);
  if (kernel == LNORM_KERNEL_WIDE && !c.ny_wide && !c.w16_wide)
    return fail(LNORM_E_UNSUPPORTED, "W-wave walk needs an instantiated width (%d columns at n = %d) and, for "
                "n <= 2, at least 3 enumerated rows, for n >= 3, column sums <= 32767", p.C, p.n);
  if (kernel == LNORM_KERNEL_TABLE16 && !c.wtab && !c.wtabw)
    return fail(LNORM_E_UNSUPPORTED, "tabulated packed walk needs |.| column sums <= 32767, an "
                "instantiated width and more enumerated rows than table rows (%d x %d at n = %d)", p.R, p.C, p.n);
  if (kernel == LNORM_KERNEL_PLAIN16 && !c.w16)
    return fail(LNORM_E_UNSUPPORTED, "packed 16-bit walk needs n >= 2, column sums <= 32767 and an instantiated "
                "width (%d columns at n = %d)", p.C, p.n);
  if (kernel == LNORM_KERNEL_AUTO) {
    // n = 1: the 16-bit blocked kernel from 2^33 sign vectors at up to 20
    // words (40 |.| columns; measured 1-2% ahead of the tabulated walk there),
    // else the tabulated walk (1-5% ahead from 22 words at 2^35-2^39, 6-10% at
    // 2^31; below 2^33 the blocked kernel's waves get too few high steps: 2^29
    // ran 27x slower), else the 32-bit blocked kernel
    // Where the 16-bit walks do not apply (column sums beyond 32767, or too
    // few rows for their tables), the 32-bit tabulated walk.
    if (can_block && !c.bp.wide && ((p.total >= (1ull << 33) && c.wtab <= 20) || !c.wtab))
      kernel = LNORM_KERNEL_BLOCKED;
    else if (c.wtab || c.wtabw)
      kernel = LNORM_KERNEL_TABLE16;
    else if (c.wtab32 || c.wtab32w)
      kernel = LNORM_KERNEL_TABLE32;
    else if (can_block)
      kernel = LNORM_KERNEL_BLOCKED;
    else if (c.w16)
      kernel = LNORM_KERNEL_PLAIN16;
    else if (c.ny_plain)
      kernel = LNORM_KERNEL_PLAIN;
    else if (c.ny_wide || c.w16_wide)
      kernel = LNORM_KERNEL_WIDE;
    else
      kernel = c.cols ? LNORM_KERNEL_COLUMNS : LNORM_KERNEL_GENERAL;
  }
  if (kernel < LNORM_KERNEL_PLAIN || kernel > LNORM_KERNEL_COLUMNS)
    return fail(LNORM_E_ARG, "unknown kernel %d", o.kernel);
  c.kernel = kernel;
  return 0;
}

void prepare_for(const Problem &p, const Choice &c, Prepared &pp) {
  if (c.kernel == LNORM_KERNEL_PLAIN16) prepare_walk(p, WALK_PLAIN16, 2 * c.w16, pp);
  else if (c.kernel == LNORM_KERNEL_TABLE16 && c.wtab) prepare_walk(p, WALK_TAB16, 2 * c.wtab, pp);
  else if (c.kernel == LNORM_KERNEL_TABLE16) prepare_walk(p, WALK_TAB16W, 2 * c.wtabw, pp);
  else if (c.kernel == LNORM_KERNEL_TABLE32 && c.wtab32) prepare_walk(p, WALK_TAB32, c.wtab32, pp);
  else if (c.kernel == LNORM_KERNEL_TABLE32) prepare_walk(p, WALK_TAB32W, c.wtab32w, pp);
  else if (c.kernel == LNORM_KERNEL_COLUMNS) prepare_walk(p, p.n == 1 ? WALK_COLS : WALK_COLSD, c.cols, pp);
  else if (c.kernel == LNORM_KERNEL_WIDE && c.w16_wide) prepare_walk(p, WALK_WIDE16, 2 * c.w16_wide, pp);
  else if (c.kernel == LNORM_KERNEL_WIDE) prepare_walk(p, WALK_WIDE, c.ny_wide, pp);
  else if (c.kernel == LNORM_KERNEL_BLOCKED && !c.ny_plain && c.ny_wide) prepare_walk(p, WALK_WIDE, c.ny_wide, pp);
  else if (c.kernel == LNORM_KERNEL_GENERAL || !c.ny_plain) prepare_walk(p, WALK_GENERAL, 0, pp);
  else prepare_walk(p, WALK_PLAIN, c.ny_plain, pp);
}

int upload_walk(DevCtx &cx, const Problem &p, const Prepared &pp, hipStream_t st) {
  const int nvec = (p.kind == KIND_L1 || p.kind == KIND_LM) ? 1 : p.n;
  if (int r = ensure(cx.A, pp.A.size() * 4)) return r;
  if (int r = ensure(cx.D, pp.D.size() * 4)) return r;
  if (int r = ensure(cx.units, cx.unit_chunk * 4)) return r;
  if (int r = ensure(cx.scal, 256)) return r;
  if (pp.walk == WALK_GENERAL) {
    // one state slice (nvec x C words) per concurrent wave of the general
    // kernel; a fixed scratch runs as many waves as its slices hold
    const size_t slice = (size_t)nvec * p.C * 4;
    cx.general_slices = cx.state.fixed ? std::min<uint64_t>(kGeneralSlices, cx.state.cap / slice) : kGeneralSlices;
    if (cx.general_slices == 0) return ensure(cx.state, slice);   // the size error
    if (int r = ensure(cx.state, (size_t)cx.general_slices * slice)) return r;
  }
  HIPCHK(hipMemcpyAsync(cx.A.p, pp.A.data(), pp.A.size() * 4, hipMemcpyHostToDevice, st));
  if (!pp.D.empty()) HIPCHK(hipMemcpyAsync(cx.D.p, pp.D.data(), pp.D.size() * 4, hipMemcpyHostToDevice, st));
  return 0;
}

// ---- autotuner ------------------------------------------------------------------
// Launch shapes are timed on probes: the real kernel over a prefix of the
// problem's own units, sized to a few full rounds of the chip (about 2^35
// labelings), so per-unit work and occupancy are those of the full run.
//   blocked: patterns per lane P (1, 2, 4, 8) x high ranks per wave 2^s
//            (s = 10..14), wherever an instantiation exists;
//   plain, plain16: threads per workgroup (64, 128, 256).
// The fastest shape is cached per device and kernel shape (kernel, kind, d,
// padded width, marginal, word size); calls with opts->autotune use it.
struct TuneKey {
  int dev, kernel, kind, n, width, marg, wide;
  bool operator<(const TuneKey &o) const {
    return std::tie(dev, kernel, kind, n, width, marg, wide) <
           std::tie(o.dev, o.kernel, o.kind, o.n, o.width, o.marg, o.wide);
  }
};
struct TuneVal {
  int P = 0, s = 0, block = 0;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

TuneKey tune_key(int dev, const Problem &p, const Choice &c) {
  TuneKey k{dev, c.kernel, (int)p.kind, p.n, 0, p.kind == KIND_LM ? 1 : 0, 0};
  if (c.kernel == LNORM_KERNEL_BLOCKED) {
    k.width = c.bp.NW;
    k.wide = c.bp.wide ? 1 : 0;
  } else if (c.kernel == LNORM_KERNEL_PLAIN16) {
    k.width = c.w16;
  } else if (c.kernel == LNORM_KERNEL_TABLE16) {
    k.width = c.wtab ? c.wtab : 1000 + c.wtabw;
  } else if (c.kernel == LNORM_KERNEL_TABLE32) {
    k.width = c.wtab32 ? c.wtab32 : 1000 + c.wtab32w;
  } else {
    k.width = c.ny_plain;
  }
  return k;
}

// Times the candidates on the locked context; fills `best` (the tuned
// fields of o) and appends one entry per candidate to `log` when given.
int tune_single(DevCtx &cx, const Problem &p, const Choice &c, const lnorm_hip_opts &o, hipStream_t st,
                TuneVal &best, std::vector<lnorm_hip_tune_entry> *log) {
  double best_rate = -1, default_rate = -1;
  auto note = [&](const lnorm_hip_tune_entry &e) {
    if (log) log->push_back(e);
    if (e.labelings_per_s > best_rate) {
      best_rate = e.labelings_per_s;
      best.P = e.patterns_per_lane;
      best.s = e.wave_log2;
      best.block = e.block;
    }
  };
  if (c.kernel == LNORM_KERNEL_BLOCKED) {
    for (int P : {1, 2, 4, 8}) {
      for (int s = 10; s <= 14; ++s) {
        lnorm_hip_opts oc = o;
        oc.patterns_per_lane = P;
        oc.wave_log2 = s;
        BlockedPlan bp;
        std::string why;
        if (!blocked_plan(p, oc, bp, why) || bp.s != s) continue;
        const uint64_t per_wave = (uint64_t)64 * P << s;
        const uint64_t probe = std::max<uint64_t>(1ull << 14, std::min<uint64_t>(1ull << 17, (1ull << 35) / per_wave));
        Phase1Out out;
        if (int r = run_blocked(cx, p, bp, st, out, probe)) return r;
        lnorm_hip_tune_entry e{};
        e.kernel = LNORM_KERNEL_BLOCKED;
        e.patterns_per_lane = P;
        e.wave_log2 = s;
        e.ms = out.ms_kernel;
        e.labelings_per_s = (double)(out.units * per_wave) / (out.ms_kernel / 1e3);
        note(e);
        // the default shape (P and s chosen by blocked_plan) stays unless a
        // candidate beats it by more than 2%: shapes within the probes' noise
        // would otherwise change from run to run
        lnorm_hip_opts od = o;
        od.patterns_per_lane = 0;
        od.wave_log2 = 0;
        BlockedPlan bd;
        if (blocked_plan(p, od, bd, why) && bd.P == P && bd.s == s) default_rate = e.labelings_per_s;
      }
    }
    if (default_rate > 0 && default_rate * 1.02 >= best_rate) {
      best.P = 0;   // 0: blocked_plan's own choice
      best.s = 0;
    }
  } else if (c.kernel == LNORM_KERNEL_PLAIN || c.kernel == LNORM_KERNEL_PLAIN16 ||
             (c.kernel == LNORM_KERNEL_TABLE16 && c.wtab) ||
             (c.kernel == LNORM_KERNEL_TABLE32 && c.wtab32)) {   // the W-wave walks' workgroup is fixed
    Prepared pp;
    prepare_for(p, c, pp);
    if (int r = upload_walk(cx, p, pp, st)) return r;
    const double lb = std::log2((double)p.base);
    const Geometry full = geometry_for(p.base, p.R - 1, 20, o.seg_log, pp.walk, pp.qmin);
    // segments of at most ~2^14 ranks, 2^20 of them: the probe's shape
    const int qp = std::min(full.q, (int)std::floor(14.0 / lb));
    const Geometry g = geometry_for(p.base, p.R - 1, 63, std::max(qp, 0), pp.walk, pp.qmin);
    for (int block : {64, 128, 256}) {
      Phase1Out out;
      if (qp <= 0) break;
      if (int r = run_walk_max(cx, p, pp, g, st, block, out, 1ull << 20)) return r;
      lnorm_hip_tune_entry e{};
      e.kernel = c.kernel;
      e.block = block;
      e.ms = out.ms_kernel;
      e.labelings_per_s = (double)(out.units * g.seglen) / (out.ms_kernel / 1e3);
      note(e);
    }
  }
  return 0;
}

double walk_cost(Walk walk, int d, int width);

// Segment length and launch size of a single problem's tabulated walk.  All
// lanes of a launch walk the same number of ranks, so a launch of W
// workgroups takes ceil(W / resident workgroups) rounds of equal length, and
// the last round of a launch may run a fraction of the chip.  Among segment
// exponents from 3 below the default (at least the table rows) to 1 above it,
// this picks the one with the least modelled time: rounds x (segment work +
// the lane's start-up: the seed and, for d >= 2, all d label tables) plus
// 30 us per launch, at launch sizes of whole rounds.  Other walks, and a
// caller's seg_log, keep the default (about 2^20 segments).
Geometry single_geometry(const DevCtx &cx, const Problem &p, const Prepared &pp, const lnorm_hip_opts &o,
                         uint64_t &chunk) {
  chunk = 0;
  const Geometry def = geometry_for(p.base, p.R - 1, 20, o.seg_log, pp.walk, pp.qmin);
  if (o.seg_log > 0 || (pp.walk != WALK_TAB16 && pp.walk != WALK_TAB32) || cx.cus <= 0) return def;
  const int block = (o.block > 0 && o.block % 64 == 0 && o.block <= 256) ? o.block : 256;
  const bool w32 = pp.walk == WALK_TAB32;
  const int words = w32 ? pp.ny_pad : pp.ny_pad / 2, wpb = block / 64;
  int nb;
  if (w32) nb = p.n == 1 ? tab32_bin_blocks_per_cu(p.kind, words, block) : blk32_blocks_per_cu(p.n, words, block);
  else nb = p.n == 1 ? tab16_bin_blocks_per_cu(p.kind, words, block) : blk16_blocks_per_cu(p.n, words, block);
  if (nb <= 0) return def;
  const uint64_t slots = (uint64_t)nb * cx.cus;              // resident workgroups
  const uint64_t per_round = slots * wpb * 64;                 // lane segments per round
  const uint64_t launch = per_round <= cx.unit_chunk ? cx.unit_chunk / per_round * per_round : cx.unit_chunk;
  const double c = walk_cost(pp.walk, p.n, pp.ny_pad), nw = words;
  const double start = (p.n == 1 ? 0.0 : (double)p.n * (1 << pp.qmin) * nw) + 2.0 * p.R * nw;
  // one round: every lane of nb * wpb / 4 waves per SIMD walks its segment,
  // 4 cycles per wave64 VALU op at about 2.4 GHz
  const double round_s = 4.0 * nb * wpb / 4.0 / 2.4e9;
  const int e = p.R - 1;
  Geometry best = def;
  double best_t = -1;
  for (int q = std::max(pp.qmin, def.q - 3); q <= std::min(e, def.q + 1); ++q) {
    const Geometry g = geometry_for(p.base, e, 63, q, pp.walk, pp.qmin);
    if (g.q != q) continue;
    // rounds of a launch of n lane segments; full launches, then the rest
    auto launch_rounds = [&](uint64_t n) {
      const uint64_t wg = (2 * ((n + 127) / 128) + wpb - 1) / wpb;
      return (double)((wg + slots - 1) / slots);
    };
    const uint64_t full = g.nseg / launch, rest = g.nseg % launch;
    const double rounds = (double)full * launch_rounds(launch) + (rest ? launch_rounds(rest) : 0.0);
    const double launches = (double)full + (rest ? 1.0 : 0.0);
    const double t = rounds * ((double)g.seglen * c + start) * round_s + launches * 30e-6;
    if (best_t < 0 || t < best_t) {
      best = g;
      best_t = t;
    }
  }
  chunk = launch;
  return best;
}

int calc_impl(const int32_t *A, int R, int C, int n, int marginal, const lnorm_hip_opts *optsp,
              int32_t *strategy, lnorm_hip_result *res, bool noalloc = false) {
  const auto t_start = std::chrono::steady_clock::now();
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  if (optsp) o = *optsp;
  g_err.clear();
  Problem p{A, R, C, n, marginal, KIND_L1, 2, 0};
  if (int r = validate(p)) return r;
  uint64_t rb = 0, re = 0;
  if (int r = rank_range(p, o, rb, re)) return r;
  if (o.want_strategy && R > 1 && !strategy) return fail(LNORM_E_ARG, "strategy output is NULL");
  int dev = 0;
  if (int r = select_device(o, dev)) return r;
  DevCtx &cx = noalloc ? g_ctx_noalloc[dev] : g_ctx[dev];
  std::lock_guard<std::mutex> lock(cx.mu);
  if (int r = noalloc ? noalloc_init(cx) : ctx_init(cx)) return r;
  hipStream_t st = o.stream ? (hipStream_t)o.stream : cx.stream;

  lnorm_hip_result rr;
  std::memset(&rr, 0, sizeof rr);

  Choice ch;
  if (int r = choose_kernel(p, o, ch)) return r;
  if (o.autotune && (ch.kernel == LNORM_KERNEL_BLOCKED || ch.kernel == LNORM_KERNEL_PLAIN ||
                     ch.kernel == LNORM_KERNEL_PLAIN16 || ch.kernel == LNORM_KERNEL_TABLE16 ||
                     ch.kernel == LNORM_KERNEL_TABLE32)) {
    // launch shapes from the cache (tuned on the first miss); explicit
    // options of the caller win
    const TuneKey key = tune_key(dev, p, ch);
    TuneVal tv;
    bool hit;
    {
      std::lock_guard<std::mutex> tl(g_tune_mu);
      auto it = g_tune.find(key);
      hit = it != g_tune.end();
      if (hit) tv = it->second;
    }
    if (!hit) {
      if (int r = tune_single(cx, p, ch, o, st, tv, nullptr)) return r;
      std::lock_guard<std::mutex> tl(g_tune_mu);
      g_tune[key] = tv;
    }
    const lnorm_hip_opts asked = o;
    if (!o.patterns_per_lane && tv.P) o.patterns_per_lane = tv.P;
    if (!o.wave_log2 && tv.s) o.wave_log2 = tv.s;
    if (!o.block && tv.block) o.block = tv.block;
    if (ch.kernel == LNORM_KERNEL_BLOCKED) {
      std::string why;
      if (!blocked_plan(p, o, ch.bp, why)) {
        // the cached shape was tuned on a problem of the same width with more
        // rows (the key has no row count): this one keeps the caller's shape
        o = asked;
        if (int r = choose_kernel(p, o, ch)) return r;
      }
    }
  }
  const int kernel = ch.kernel;
  const BlockedPlan &bp = ch.bp;
  rr.total = re - rb;

  // the walk that enumerates (non-blocked kernels) and recovers the strategy (all)
  Prepared pp;
  prepare_for(p, ch, pp);
  const bool general = pp.walk == WALK_GENERAL;
  if (int r = upload_walk(cx, p, pp, st)) return r;

  // The ranks in runs of aligned units (one run of whole units for the
  // whole range): whole units on the chosen kernel, shorter units at the
  // ends of a partial range on the walk `pp` where its segments may be that
  // short, else on the plain (or general) walk `pp2`.
  int unit_e;   // log_base of the ranks per unit
  uint64_t chunk = 0;
  Geometry gmain{};
  if (kernel == LNORM_KERNEL_BLOCKED) {
    unit_e = bp.s + bp.kl;
  } else {
    gmain = single_geometry(cx, p, pp, o, chunk);
    unit_e = gmain.q;
  }
  const std::vector<RankRun> runs = split_ranks(p.base, unit_e, rb, re);
  const int qlow = walk_qmin(pp);
  std::vector<char> on2(runs.size(), 0);
  bool need2 = false;
  for (size_t i = 0; i < runs.size(); ++i)
    if (runs[i].e < unit_e && runs[i].e < qlow) on2[i] = need2 = true;
  Prepared pp2;
  if (need2) prepare_walk(p, ch.ny_plain ? WALK_PLAIN : WALK_GENERAL, ch.ny_plain, pp2);

  std::vector<Phase1Out> outs(runs.size());
  Phase1Out p1;
  HIPCHK(hipEventRecord(cx.ev[0], st));
  for (int pass = 0; pass < (need2 ? 2 : 1); ++pass) {
    if (pass == 1)
      if (int r = upload_walk(cx, p, pp2, st)) return r;
    for (size_t i = 0; i < runs.size(); ++i) {
      const RankRun &run = runs[i];
      if (on2[i] != pass) continue;
      int r;
      if (run.e == unit_e && kernel == LNORM_KERNEL_BLOCKED) {
        r = run_blocked(cx, p, bp, st, outs[i], run.n, run.u0);
      } else if (run.e == unit_e) {
        Geometry g = gmain;
        g.nseg = run.n;
        r = run_walk_max(cx, p, pp, g, st, o.block, outs[i], 0, chunk, run.u0);
      } else {
        const Geometry g{run.e, ipow((uint64_t)p.base, run.e), run.n};
        r = run_walk_max(cx, p, pass ? pp2 : pp, g, st, o.block, outs[i], 0, 0, run.u0);
      }
      if (r) return r;
      if (run.e == unit_e) p1.units += run.n;
      p1.ms_kernel += outs[i].ms_kernel;
    }
  }
  // the maximum, and the first run attaining it (runs are in rank order)
  size_t win = 0;
  for (size_t i = 1; i < runs.size(); ++i)
    if (outs[i].best > outs[win].best) win = i;
  p1.best = outs[win].best;
  p1.first_unit = outs[win].first_unit;
  if (kernel == LNORM_KERNEL_BLOCKED) {
    rr.patterns_per_lane = bp.P;
    if (bp.wide)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "blocked32<nc=%d,P=%d,W=%d,marg=%d,s=%d>", bp.NC, bp.P, bp.W,
               p.kind == KIND_LM, bp.s);
    else
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "blocked16<np=%d,P=%d,marg=%d,s=%d>", bp.np, bp.P,
               p.kind == KIND_LM, bp.s);
  } else {
    const int q = gmain.q;
    if (general)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "general<kind=%d,d=%d,q=%d>", (int)p.kind, p.n, q);
    else if (pp.walk == WALK_PLAIN16)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "plain16<d=%d,words=%d,q=%d>", p.n, pp.ny_pad / 2, q);
    else if (pp.walk == WALK_WIDE)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "wide<kind=%d,ny=%d,q=%d>", (int)p.kind, pp.ny_pad, q);
    else if (pp.walk == WALK_WIDE16)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "wide16<d=%d,words=%d,q=%d>", p.n, pp.ny_pad / 2, q);
    else if (pp.walk == WALK_TAB16 && p.n == 1)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "table16<%s,words=%d,M=%d,q=%d>",
               p.kind == KIND_LM ? "LM" : "L1", pp.ny_pad / 2, pp.qmin, q);
    else if (pp.walk == WALK_TAB16)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "table16<d=%d,words=%d,M=%d,q=%d>", p.n, pp.ny_pad / 2,
               pp.qmin, q);
    else if (pp.walk == WALK_TAB16W && p.n == 1)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "table16w<%s,words=%d,M=%d,q=%d>",
               p.kind == KIND_LM ? "LM" : "L1", pp.ny_pad / 2, pp.qmin, q);
    else if (pp.walk == WALK_TAB16W)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "table16w<d=%d,words=%d,M=%d,q=%d>", p.n, pp.ny_pad / 2,
               pp.qmin, q);
    else if (is_tab32(pp.walk) && p.n == 1)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "table32%s<%s,words=%d,M=%d,q=%d>",
               pp.walk == WALK_TAB32W ? "w" : "", p.kind == KIND_LM ? "LM" : "L1", pp.ny_pad, pp.qmin, q);
    else if (is_tab32(pp.walk))
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "table32%s<d=%d,words=%d,M=%d,q=%d>",
               pp.walk == WALK_TAB32W ? "w" : "", p.n, pp.ny_pad, pp.qmin, q);
    else if (pp.walk == WALK_COLS || pp.walk == WALK_COLSD)
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "cols<kind=%d,d=%d,cpl=%d,q=%d>", (int)p.kind, p.n,
               pp.ny_pad / 64, q);
    else
      snprintf(rr.kernel_name, sizeof rr.kernel_name, "plain<kind=%d,d=%d,ny=%d,q=%d>", (int)p.kind, p.n,
               pp.ny_pad, q);
  }
  HIPCHK(hipEventRecord(cx.ev[1], st));
  rr.units = p1.units;
  rr.unit_log = unit_e;
  rr.kernel = kernel;
  rr.Lnorm = p1.best;

  if (o.want_strategy && R > 1) {
    // the smallest maximising rank: MODE_FIRST over the winning unit, on the
    // walk of its run (the walk on the device is pp2 after a second pass)
    if (need2 && !on2[win])
      if (int r = upload_walk(cx, p, pp, st)) return r;
    HIPCHK(hipEventRecord(cx.ev[2], st));
    uint64_t rank = 0;
    if (int r = run_walk_first(cx, p, on2[win] ? pp2 : pp, p1.first_unit, runs[win].e, p1.best, st, o.block, rank))
      return r;
    HIPCHK(hipEventRecord(cx.ev[3], st));
    HIPCHK(hipEventSynchronize(cx.ev[3]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, cx.ev[2], cx.ev[3]));
    rr.ms_strategy = ms;
    rr.best_rank = rank;
    labels_of_rank(rank, R, n, strategy);
  } else {
    rr.best_rank = R == 1 ? 0 : ~0ull;   // one row: rank 0 is the only labeling
  }
  HIPCHK(hipEventSynchronize(cx.ev[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, cx.ev[0], cx.ev[1]));
  rr.ms_enumerate = ms;
  rr.ms_kernel = p1.ms_kernel;
  rr.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (res) *res = rr;
  return LNORM_OK;
}

// ---- batches ------------------------------------------------------------------
// Items that fit a batched walk (L1/LM with at most 128 columns; d >= 2 with
// column sums <= 2^15 - 1, packed, and tabulated where the width
// and row count allow) run in one persistent launch per instantiation; the
// rest go through calc_impl one by one.
struct BItem {
  int idx;             // position in the caller's array
  Problem p;
  Prepared pp;
  Geometry g1, g2;     // phase-1 segments; phase-2 sub-segments of one unit
  double work;         // labelings x walk_cost: longest first within a group
  size_t a_off = 0, d_off = 0;
  uint64_t ub_off = 0;
};

struct BGroup {
  Walk walk;
  Kind kind;
  int d, width;
  std::vector<int> members;   // indices into the BItem vector, longest first
  uint64_t ntasks1 = 0, ntasks2 = 0;
  int first = 0;              // position of the group's first item in task order
};

uint64_t waves_for(uint64_t nseg) { return 2 * ((nseg + 127) / 128); }

// Items above 2^32 labelings fill the chip on their own; they take the
// single-problem path (blocked MFMA kernel for L1/LM, chunked launches).
constexpr uint64_t kBatchMaxTotal = 1ull << 32;

// Phase-1 segments of a batched item: at least the 128 lane segments of the
// two waves one task pair covers, and at most ~2^cap_log ranks per lane so a
// task stays short next to the others (balance over the persistent grid and
// a short tail at the end of each group).
constexpr int kBatchCapLog = 14;
constexpr int kBatchStreams = 4;
constexpr int kPrepThreads = 16;   // host threads preparing batch items (the box's CPU share)
Geometry batch_geometry(int base, int e, int forced, Walk walk, int cap_log, int qmin = 0, int qfloor = 0) {
  if (forced > 0) return geometry_for(base, e, 63, forced, walk, qmin);
  const double lb = std::log2((double)base);
  Geometry g = geometry_for(base, e, 7, 0, walk, qmin);
  const int qcap = std::max({0, qfloor, (int)std::floor(cap_log / lb)});
  if (g.q > qcap) g = geometry_for(base, e, 63, qcap, walk, qmin);
  return g;
}

// Relative VALU work per labeling of a batch walk, for balancing groups over
// streams: the binary walk 2 per column, the packed walk 4 per word plus d,
// the tabulated walk 1.5 plus its table sums 2 + 2^(M+1) per word over d^M.
double walk_cost(Walk walk, int d, int width) {
  if (walk == WALK_PLAIN) return 2.0 * width;
  if (walk == WALK_TAB32 && d == 1) return width + 1.5 + width / 32.0;
  if (walk == WALK_TAB32) {
    const int M = blk32_rows(d, width);
    return 1.5 + (2.0 + (double)(2 << M)) * width / std::pow((double)d, (double)M);
  }
  const int nw = width / 2;
  if (walk == WALK_TAB16 && d == 1) return nw + 1.5 + nw / 32.0;
  if (walk == WALK_TAB16) {
    const int M = blk16_rows(d, nw);
    return 1.5 + (2.0 + (double)(2 << M)) * nw / std::pow((double)d, (double)M);
  }
  return 4.0 * nw + d;
}

// Batch shape for the tuner's cache: FNV-1a over the sorted item sizes.
uint64_t batch_shape_key(const lnorm_hip_item *items, int count) {
  std::vector<uint64_t> v((size_t)count);
  for (int i = 0; i < count; ++i)
    v[i] = ((uint64_t)(uint32_t)items[i].iRows << 40) ^ ((uint64_t)(uint32_t)items[i].iCols << 16) ^
           ((uint64_t)(uint32_t)items[i].n << 4) ^ (uint64_t)(items[i].marginal & 1);
  std::sort(v.begin(), v.end());
  uint64_t h = 1469598103934665603ull;
  for (uint64_t x : v)
    for (int b = 0; b < 8; ++b) {
      h ^= (x >> (8 * b)) & 0xFF;
      h *= 1099511628211ull;
    }
  return h ^ (uint64_t)count;
}
std::map<uint64_t, std::pair<int, int>> g_tune_batch;   // shape -> (streams, cap_log), under g_tune_mu

int batch_impl(lnorm_hip_item *items, int count, const lnorm_hip_opts &o, lnorm_hip_batch_stats *stats);

// Times (streams, cap) candidates on a sample of the batch; stores the best.
int tune_batch(const lnorm_hip_item *items, int count, const lnorm_hip_opts &o, lnorm_hip_opts &best,
               std::vector<lnorm_hip_tune_entry> *log) {
  const int stride = count >= 64 ? 8 : 1;
  std::vector<lnorm_hip_item> sample;
  for (int i = 0; i < count; i += stride) {
    lnorm_hip_item it = items[i];
    it.strategy = nullptr;
    sample.push_back(it);
  }
  best = o;
  double best_ms = -1;
  int ntimed = 0;
  for (int streams : {1, 2, 4}) {
    for (int cap : {10, 12, 14, 16}) {
      lnorm_hip_opts oc = o;
      oc.autotune = 0;
      oc.batch_graph = 0;
      oc.batch_streams = streams;
      oc.batch_cap_log = cap;
      lnorm_hip_batch_stats st;
      std::vector<lnorm_hip_item> work = sample;
      const int r = batch_impl(work.data(), (int)work.size(), oc, &st);
      if (r == LNORM_E_HIP || r == LNORM_E_ARG) return r;
      if (st.batched == 0) return 0;          // nothing in this batch runs batched
      lnorm_hip_tune_entry e{};
      e.batch_streams = streams;
      e.batch_cap_log = cap;
      e.ms = st.ms_device;
      e.labelings_per_s = st.ms_device > 0 ? (double)st.labelings / (st.ms_device / 1e3) : 0;
      if (log) log->push_back(e);
      ++ntimed;
      if (best_ms < 0 || st.ms_device < best_ms) {
        best_ms = st.ms_device;
        best.batch_streams = streams;
        best.batch_cap_log = cap;
      }
    }
  }
  std::lock_guard<std::mutex> tl(g_tune_mu);
  g_tune_batch[batch_shape_key(items, count)] = {best.batch_streams, best.batch_cap_log};
  return ntimed;
}

// Classifies batch item i: validates it, picks its walk and prepares its
// tables and geometry in b.  Returns 0 (batched), 1 (computed on its own) or
// the item's error code.
int batch_item(const lnorm_hip_item &it, int i, const lnorm_hip_opts &o, int cap_log, BItem &b) {
  b.idx = i;
  b.p = Problem{it.mtx_as_vec, it.iRows, it.iCols, it.n, it.marginal, KIND_L1, 2, 0};
  if (int r = validate(b.p)) return r;
  Walk walk;
  int width;
  if (b.p.n >= 2) {
    const bool fits16 = max_col_abs_sum(b.p, 0) <= 32767;
    const int w16 = fits16 ? plain16_words(b.p.n, b.p.C) : 0;
    int wtab = fits16 ? batch_t16_words(b.p.n, b.p.C) : 0;   // 4..32 words in steps of 4
    if (wtab && b.p.R - 1 < blk16_rows(b.p.n, wtab)) wtab = 0;
    if (wtab) {
      walk = WALK_TAB16;
      width = 2 * wtab;
    } else {
      if (!w16) return 1;
      walk = WALK_PLAIN16;
      width = 2 * w16;
    }
  } else {
    const int marg = b.p.kind == KIND_LM ? 1 : 0;
    int wtb = max_col_abs_sum(b.p, marg) <= 32767 ? batch_tb16_words(b.p.C - marg) : 0;   // up to 64 words
    if (wtb && b.p.R - 1 < tab_bin_rows(wtb)) wtb = 0;
    if (wtb) {
      walk = WALK_TAB16;
      width = 2 * wtb;
    } else {
      width = plain_ny_pad(b.p.kind, 1, b.p.C);
      if (!width) return 1;
      walk = WALK_PLAIN;
    }
  }
  if (b.p.total > kBatchMaxTotal) return 1;
  prepare_walk(b.p, walk, width, b.pp);
  // a lane segment of the tabulated d-ary walk starts with all d label tables
  // (d 2^M NW v_sad_u16) and the seed (about 2 R NW): its segments hold at
  // least 32 times that work, whatever the cap (3^9 labelings at d = 3 and 16
  // words, against 3^7 at the default cap)
  int qfloor = 0;
  if (walk == WALK_TAB16 && b.p.n >= 2) {
    const double nw = width / 2, start = (double)b.p.n * (1 << b.pp.qmin) * nw + 2.0 * b.p.R * nw;
    const double c = walk_cost(walk, b.p.n, width);
    while (qfloor < b.p.R - 1 && std::pow((double)b.p.n, qfloor) * c < 32 * start) ++qfloor;
  }
  b.g1 = batch_geometry(b.p.base, b.p.R - 1, o.seg_log, walk, cap_log, b.pp.qmin, qfloor);
  b.g2 = geometry_for(b.p.base, b.g1.q, 8, 0, walk, b.pp.qmin);
  b.work = (double)b.p.total * walk_cost(walk, b.p.n, width);
  return 0;
}

int batch_impl(lnorm_hip_item *items, int count, const lnorm_hip_opts &o, lnorm_hip_batch_stats *stats) {
  const auto t_start = std::chrono::steady_clock::now();
  int first_err = 0;
  std::vector<BItem> bis;
  std::vector<int> fallback;
  if (o.rank_begin || o.rank_end)
    return fail(LNORM_E_ARG, "batches enumerate whole problems: rank_begin and rank_end must be 0");
  if (o.batch_streams < 0 || o.batch_streams > kSideStreams + 1)
    return fail(LNORM_E_ARG, "batch_streams must be 0..%d", kSideStreams + 1);
  if (o.batch_cap_log != 0 && (o.batch_cap_log < 8 || o.batch_cap_log > 20))
    return fail(LNORM_E_ARG, "batch_cap_log must be 0 or 8..20");
  int cap_log = o.batch_cap_log ? o.batch_cap_log : kBatchCapLog;
  int nlanes = o.batch_streams ? o.batch_streams : kBatchStreams;
  if (o.autotune && !o.batch_cap_log && !o.batch_streams && count > 0) {
    const uint64_t key = batch_shape_key(items, count);
    std::pair<int, int> tuned{0, 0};
    {
      std::lock_guard<std::mutex> tl(g_tune_mu);
      auto it = g_tune_batch.find(key);
      if (it != g_tune_batch.end()) tuned = it->second;
    }
    if (!tuned.first) {
      lnorm_hip_opts bestopt;
      if (int r = tune_batch(items, count, o, bestopt, nullptr); r < 0) return r;
      tuned = {bestopt.batch_streams, bestopt.batch_cap_log};
    }
    if (tuned.first) {
      nlanes = tuned.first;
      cap_log = tuned.second;
    }
  }
  // items are classified and prepared in parallel (tables and packing are
  // the host's share of a batch); results are gathered in order, so the
  // statuses, the first error code and the last error message are those of
  // a sequential pass
  std::vector<BItem> slot(count);
  std::vector<int> cls(count, 0);
  std::vector<std::string> msg(count);
  {
    std::atomic<int> next{0};
    auto worker = [&] {
      for (;;) {
        const int lo = next.fetch_add(32);
        if (lo >= count) break;
        for (int i = lo; i < std::min(count, lo + 32); ++i) {
          cls[i] = batch_item(items[i], i, o, cap_log, slot[i]);
          if (cls[i] < 0) msg[i] = g_err;
        }
      }
    };
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nthr = std::min({kPrepThreads, hw, (count + 127) / 128});
    std::vector<std::thread> pool;
    for (int t = 1; t < nthr; ++t) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
  }
  for (int i = 0; i < count; ++i) {
    lnorm_hip_item &it = items[i];
    it.Lnorm = 0;
    it.status = LNORM_E_HIP;   // until this item's result is in
    if (cls[i] < 0) {
      it.status = cls[i];
      if (!first_err) first_err = cls[i];
      g_err = msg[i];
    } else if (cls[i] == 1) {
      fallback.push_back(i);
    } else {
      bis.push_back(std::move(slot[i]));
    }
  }

  // groups by instantiation, longest item first within each
  std::vector<BGroup> groups;
  for (int k = 0; k < (int)bis.size(); ++k) {
    const BItem &b = bis[k];
    BGroup *g = nullptr;
    for (auto &gg : groups)
      if (gg.walk == b.pp.walk && gg.kind == b.p.kind && gg.d == b.p.n && gg.width == b.pp.ny_pad) g = &gg;
    if (!g) {
      groups.push_back(BGroup{b.pp.walk, b.p.kind, b.p.n, b.pp.ny_pad, {}, 0, 0, 0});
      g = &groups.back();
    }
    g->members.push_back(k);
  }
  std::vector<int> order;      // task order: group after group
  for (auto &g : groups) {
    std::stable_sort(g.members.begin(), g.members.end(), [&](int a, int b) { return bis[a].work > bis[b].work; });
    g.first = (int)order.size();
    for (int k : g.members) order.push_back(k);
  }
  const int nb = (int)order.size();

  double ms_dev = 0;
  bool replayed = false;
  if (nb > 0) {
    int dev = 0;
    if (int r = select_device(o, dev)) return r;
    DevCtx &cx = g_ctx[dev];
    std::lock_guard<std::mutex> lock(cx.mu);
    if (int r = ctx_init(cx)) return r;
    hipStream_t st = o.stream ? (hipStream_t)o.stream : cx.stream;
    // pools
    size_t aw = 0, dw = 0;
    uint64_t uw = 0;
    for (int k : order) {
      BItem &b = bis[k];
      b.a_off = aw;
      aw += b.pp.A.size();
      b.d_off = dw;
      dw += b.pp.D.size();
      b.ub_off = uw;
      uw += b.g1.nseg;
    }
    if (int r = ensure(cx.bA, aw * 4)) return r;
    if (int r = ensure(cx.bD, dw * 4)) return r;
    if (int r = ensure(cx.bUnits, uw * 4)) return r;
    if (int r = ensure(cx.bScal, (size_t)nb * 24)) return r;
    if (int r = ensure(cx.bTasks, (size_t)2 * nb * sizeof(BatchTask))) return r;
    if (int r = ensure(cx.bCnt, (size_t)2 * groups.size() * kCounterShards * kCounterStride * 4)) return r;
    int32_t *gbest = (int32_t *)cx.bScal.p;
    unsigned long long *gfirst = (unsigned long long *)((char *)cx.bScal.p + (size_t)nb * 8);
    unsigned long long *funit = gfirst + nb;
    BatchTask *t1 = (BatchTask *)cx.bTasks.p, *t2 = t1 + nb;
    unsigned int *cnt = (unsigned int *)cx.bCnt.p;
    std::vector<int32_t> hA(aw), hD(dw);
    std::vector<BatchTask> h1(nb), h2(nb);
    for (auto &g : groups) {
      uint64_t task1 = 0, task2 = 0;
      for (int m = 0; m < (int)g.members.size(); ++m) {
        const int pos = g.first + m;
        BItem &b = bis[order[pos]];
        std::copy(b.pp.A.begin(), b.pp.A.end(), hA.begin() + b.a_off);
        std::copy(b.pp.D.begin(), b.pp.D.end(), hD.begin() + b.d_off);
        BatchTask x{};
        x.A = (const int32_t *)cx.bA.p + b.a_off;
        x.D = (const int32_t *)cx.bD.p + b.d_off;
        x.R = b.p.R;
        x.q = b.g1.q;
        x.nseg = b.g1.nseg;
        x.seg_base = 0;
        x.task0 = task1;
        x.unit_best = (int32_t *)cx.bUnits.p + b.ub_off;
        x.gbest = gbest + pos;
        x.gfirst = gfirst + pos;
        x.unit = funit + pos;
        h1[pos] = x;
        task1 += waves_for(b.g1.nseg);
        x.q = b.g2.q;
        x.nseg = b.g2.nseg;
        x.seg_base = b.g2.nseg;
        x.task0 = task2;
        h2[pos] = x;
        task2 += waves_for(b.g2.nseg);
      }
      if (task1 >= (1ull << 31) || task2 >= (1ull << 31))   // t = k * 8 + shard stays in uint32
        return fail(LNORM_E_SIZE, "batch group of %llu tasks: split the batch", (unsigned long long)task1);
      g.ntasks1 = task1;
      g.ntasks2 = task2;
    }
    if (aw) HIPCHK(hipMemcpyAsync(cx.bA.p, hA.data(), aw * 4, hipMemcpyHostToDevice, st));
    if (dw) HIPCHK(hipMemcpyAsync(cx.bD.p, hD.data(), dw * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(t1, h1.data(), (size_t)nb * sizeof(BatchTask), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(t2, h2.data(), (size_t)nb * sizeof(BatchTask), hipMemcpyHostToDevice, st));

    // Groups spread over the main and the side streams, heaviest first onto
    // the least loaded stream, so that one group's last tasks overlap the
    // others' work instead of leaving the chip idle.
    const int kLanes = nlanes;
    std::vector<int> gorder(groups.size());
    for (int gi = 0; gi < (int)groups.size(); ++gi) gorder[gi] = gi;
    std::vector<double> gwork(groups.size(), 0.0);
    for (int gi = 0; gi < (int)groups.size(); ++gi)
      for (int k : groups[gi].members) gwork[gi] += bis[k].work;
    std::stable_sort(gorder.begin(), gorder.end(), [&](int a, int b) { return gwork[a] > gwork[b]; });
    std::vector<int> glane(groups.size(), 0);
    {
      double load[kSideStreams + 1] = {};
      for (int gi : gorder) {
        int best = 0;
        for (int l = 1; l < kLanes; ++l) if (load[l] < load[best]) best = l;
        glane[gi] = best;
        load[best] += gwork[gi];
      }
    }
    auto lane_stream = [&](int l) { return l == 0 ? st : cx.side[l - 1]; };

    // the launch sequence: no host synchronisation inside, so it can be
    // captured (the side streams join the capture through the fork event)
    auto phase = [&](int ph, Mode mode, BatchTask *tasks) -> int {
      HIPCHK(hipEventRecord(cx.bFork[ph], st));
      for (int l = 1; l < kLanes; ++l) HIPCHK(hipStreamWaitEvent(lane_stream(l), cx.bFork[ph], 0));
      for (int gi : gorder) {
        const BGroup &g = groups[gi];
        BatchLaunch bl{tasks + g.first, (int)g.members.size(), ph == 0 ? g.ntasks1 : g.ntasks2,
                       cnt + (size_t)(2 * gi + ph) * kCounterShards * kCounterStride,
                       g.kind, g.d, g.width, g.walk == WALK_PLAIN16 || g.walk == WALK_TAB16,
                       g.walk == WALK_TAB16};
        HIPCHK(launch_batch_walk(mode, bl, lane_stream(glane[gi])));
      }
      for (int l = 1; l < kLanes; ++l) {
        HIPCHK(hipEventRecord(cx.bJoin[ph][l - 1], lane_stream(l)));
        HIPCHK(hipStreamWaitEvent(st, cx.bJoin[ph][l - 1], 0));
      }
      return 0;
    };
    auto sequence = [&]() -> int {
      HIPCHK(launch_fill_i32(gbest, INT32_MIN, (uint64_t)nb, st));
      HIPCHK(launch_fill_u64(gfirst, ~0ull, 2 * (uint64_t)nb, st));
      HIPCHK(launch_fill_i32((int32_t *)cnt, 0, 2 * groups.size() * kCounterShards * kCounterStride, st));
      if (int r = phase(0, MODE_MAX, t1)) return r;
      HIPCHK(launch_batch_first_unit(t1, nb, funit, st));
      return phase(1, MODE_FIRST, t2);
    };
    HIPCHK(hipEventRecord(cx.ev[0], st));
    if (o.batch_graph) {
      // the graph bakes in buffer addresses and the group structure
      std::vector<uint64_t> sig = {(uint64_t)nb, (uint64_t)(uintptr_t)cx.bScal.p, (uint64_t)(uintptr_t)cx.bTasks.p,
                                   (uint64_t)(uintptr_t)cx.bCnt.p, (uint64_t)(uintptr_t)st};
      for (auto &g : groups) {
        sig.push_back(((uint64_t)g.walk << 48) | ((uint64_t)g.kind << 40) | ((uint64_t)g.d << 32) | (uint64_t)g.width);
        sig.push_back(g.members.size());
        sig.push_back(g.ntasks1);
        sig.push_back(g.ntasks2);
      }
      sig.push_back((uint64_t)nlanes);
      for (int l : glane) sig.push_back((uint64_t)l);
      if (!cx.bGraph || cx.bGraphSig != sig) {
        if (cx.bGraph) (void)hipGraphExecDestroy(cx.bGraph);
        cx.bGraph = nullptr;
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        graph = nullptr;
        const int rc = sequence();
        hipError_t ec = hipStreamEndCapture(st, &graph);
        if (rc) {
          if (ec == hipSuccess && graph) (void)hipGraphDestroy(graph);
          return rc;
        }
        if (ec != hipSuccess) return fail(LNORM_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        hipError_t ei = hipGraphInstantiate(&cx.bGraph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(LNORM_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        cx.bGraphSig = sig;
      } else {
        replayed = true;
      }
      HIPCHK(hipGraphLaunch(cx.bGraph, st));
    } else {
      if (int r = sequence()) return r;
    }
    HIPCHK(hipEventRecord(cx.ev[1], st));
    std::vector<int32_t> hb(nb);
    std::vector<unsigned long long> hf(nb);
    HIPCHK(hipMemcpyAsync(hb.data(), gbest, (size_t)nb * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hf.data(), gfirst, (size_t)nb * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, cx.ev[0], cx.ev[1]));
    ms_dev = ms;
    for (int pos = 0; pos < nb; ++pos) {
      const BItem &b = bis[order[pos]];
      lnorm_hip_item &it = items[b.idx];
      if (b.p.R > 1 && hf[pos] == ~0ull) {
        it.status = fail(LNORM_E_HIP, "batch item %d: no rank attains %d", b.idx, hb[pos]);
        if (!first_err) first_err = it.status;
        continue;
      }
      it.Lnorm = hb[pos];
      if (it.strategy && b.p.R > 1) labels_of_rank(hf[pos], b.p.R, b.p.n, it.strategy);
      it.status = LNORM_OK;
    }
  }
  for (int i : fallback) {
    lnorm_hip_item &it = items[i];
    lnorm_hip_opts oo = o;
    oo.want_strategy = it.strategy != nullptr;
    oo.kernel = LNORM_KERNEL_AUTO;
    lnorm_hip_result r;
    it.status = calc_impl(it.mtx_as_vec, it.iRows, it.iCols, it.n, it.marginal, &oo, it.strategy, &r);
    it.Lnorm = it.status == 0 ? r.Lnorm : 0;
    if (it.status && !first_err) first_err = it.status;
  }
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->items = count;
    stats->batched = nb;
    stats->fallback = (int32_t)fallback.size();
    stats->groups = (int32_t)groups.size();
    stats->graph_replayed = replayed ? 1 : 0;
    stats->ms_device = ms_dev;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    for (const BItem &b : bis) stats->labelings += b.p.total;
  }
  return first_err;
}


// ---- the int_type = long long variant (README.md:100-105) -------------------
// Problems inside the int32 contract run on the int32 kernels (identical
// results); the rest on the 64-bit walks of lnorm_int64.hip.
int run64_phase(DevCtx &cx, Kind kind, int d, int ny, bool general, bool wide, bool fp64, Mode mode,
                const Geometry &g, uint64_t seg0, uint64_t nseg, int R, int64_t target, hipStream_t st,
                double &ms_kernel, int64_t &best, uint64_t &first) {
  long long *gbest = (long long *)cx.scal.p;
  unsigned long long *gfirst = (unsigned long long *)((char *)cx.scal.p + 64);
  HIPCHK(launch_fill_i64(gbest, INT64_MIN, 1, st));
  HIPCHK(launch_fill_u64(gfirst, ~0ull, 1, st));
  Plain64Args a{};
  a.A = (const int64_t *)cx.A64.p;
  a.D = (const int64_t *)cx.D64.p;
  a.R = R;
  a.ny_pad = ny;
  a.d = d;
  a.q = g.q;
  a.seg0 = seg0;
  a.nseg = nseg;
  a.target = target;
  a.unit_best = (int64_t *)cx.units64.p;
  a.gbest = gbest;
  a.gfirst = gfirst;
  a.state = (uint64_t *)cx.state64.p;
  a.fp64 = fp64 ? 1 : 0;
  HIPCHK(hipEventRecord(cx.ev[4], st));
  hipError_t e = general ? launch_general64(kind, mode, a, ny, st)
                 : wide  ? launch_widef64(kind, mode, a, st)
                         : launch_plain64(kind, mode, a, st);
  HIPCHK(hipEventRecord(cx.ev[5], st));
  if (e != hipSuccess) return fail(LNORM_E_HIP, "64-bit launch failed: %s", hipGetErrorString(e));
  if (mode == MODE_MAX) HIPCHK(launch_first_unit64(a.unit_best, nseg, gbest, gfirst, st));
  long long hb = 0;
  unsigned long long hf = 0;
  HIPCHK(hipMemcpyAsync(&hb, gbest, sizeof hb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&hf, gfirst, sizeof hf, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0;
  HIPCHK(hipEventElapsedTime(&kms, cx.ev[4], cx.ev[5]));
  ms_kernel += kms;
  best = hb;
  first = hf;
  return 0;
}

int calc_impl64(const int64_t *A, int R, int C, int n, int marginal, const lnorm_hip_opts *optsp,
                int32_t *strategy, int64_t *Lnorm, lnorm_hip_result *res) {
  const auto t_start = std::chrono::steady_clock::now();
  g_err.clear();
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  if (optsp) o = *optsp;
  if (!A || !Lnorm) return fail(LNORM_E_ARG, "mtx_as_vec or Lnorm is NULL");
  if (R < 1 || C < 1) return fail(LNORM_E_ARG, "matrix must be at least 1 x 1 (got %d x %d)", R, C);
  // sum |A| <= INT64_MAX keeps every partial sum and objective of the
  // long long variant in range (the int32 contract, one level up)
  unsigned __int128 s = 0;
  bool fits32 = true;
  for (size_t i = 0; i < (size_t)R * C; ++i) {
    const int64_t v = A[i];
    const unsigned __int128 av = v < 0 ? (unsigned __int128)(-(v + 1)) + 1 : (unsigned __int128)v;
    s += av;
    if (v < INT32_MIN || v > INT32_MAX) fits32 = false;
  }
  if (s > (unsigned __int128)INT64_MAX)
    return fail(LNORM_E_BOUND, "sum |A| exceeds INT64_MAX; the long long variant would overflow");
  if (fits32 && s <= (unsigned __int128)INT32_MAX) {
    std::vector<int32_t> a32((size_t)R * C);
    for (size_t i = 0; i < a32.size(); ++i) a32[i] = (int32_t)A[i];
    lnorm_hip_result r;
    if (int rc = calc_impl(a32.data(), R, C, n, marginal, &o, strategy, &r)) return rc;
    *Lnorm = r.Lnorm;
    if (res) *res = r;
    return LNORM_OK;
  }
  // the checks of validate() on the shape
  std::vector<int32_t> dummy(1, 0);
  Problem p{dummy.data(), R, C, n, marginal, KIND_L1, 2, 0};
  if (p.n < 1 || p.n > kMaxLabels) return fail(LNORM_E_SIZE, "order %d outside 1..%d", p.n, kMaxLabels);
  if (p.marginal != 0 && p.marginal != 1) return fail(LNORM_E_ARG, "marginal must be 0 or 1");
  if (p.marginal && p.n != 1) return fail(LNORM_E_ARG, "marginals are defined for n = 1 only");
  if (p.R > max_rows(p.n)) return fail(LNORM_E_SIZE, "%d rows exceed maxRows %d for n = %d", p.R, max_rows(p.n), p.n);
  if (p.C > kMaxCols) return fail(LNORM_E_SIZE, "%d columns exceed length %d", p.C, kMaxCols);
  p.kind = p.n == 1 ? (p.marginal ? KIND_LM : KIND_L1) : (p.n == 2 ? KIND_L2 : KIND_LD);
  p.base = p.n <= 2 ? 2 : p.n;
  if (int r = rank_count(p)) return r;
  uint64_t rb = 0, re = 0;
  if (int r = rank_range(p, o, rb, re)) return r;
  if (o.want_strategy && R > 1 && !strategy) return fail(LNORM_E_ARG, "strategy output is NULL");
  // the plain walks in doubles while sum |A| < 2^52 (every value an exact
  // integer below 2^53); word_bits 64 keeps the 64-bit integer walks
  if (o.word_bits != 0 && o.word_bits != 64)
    return fail(LNORM_E_ARG, "word_bits must be 0 (auto) or 64 (integer walk) for the long long variant");
  const bool exact53 = o.word_bits != 64 && s < ((unsigned __int128)1 << 52);
  const int ny_plain = plain64_ny_pad(p.kind, p.n, C);
  // beyond one wave's width, the binary walk in doubles over W waves
  const int ny_wide = (exact53 && !ny_plain) ? widef64_ny_pad(p.kind, p.n, C) : 0;
  int kernel = o.kernel;
  if (kernel == LNORM_KERNEL_BLOCKED || kernel == LNORM_KERNEL_PLAIN16 || kernel == LNORM_KERNEL_TABLE16)
    return fail(LNORM_E_UNSUPPORTED, "the blocked and packed 16-bit kernels have no 64-bit counterpart");
  if (kernel == LNORM_KERNEL_PLAIN && !ny_plain)
    return fail(LNORM_E_UNSUPPORTED, "no 64-bit plain instantiation for %d columns at n = %d", C, n);
  if (kernel == LNORM_KERNEL_WIDE && !ny_wide)
    return fail(LNORM_E_UNSUPPORTED, "the 64-bit W-wave walks run in doubles: sum |A| < 2^52, beyond the one-wave "
                "widths up to 512 columns (L2: 256, d = 3: 192, 4: 128, 5..10: 64); got %d columns at n = %d", C, n);
  if (kernel == LNORM_KERNEL_AUTO)
    kernel = ny_plain ? LNORM_KERNEL_PLAIN : ny_wide ? LNORM_KERNEL_WIDE : LNORM_KERNEL_GENERAL;
  if (kernel != LNORM_KERNEL_PLAIN && kernel != LNORM_KERNEL_GENERAL && kernel != LNORM_KERNEL_WIDE)
    return fail(LNORM_E_ARG, "unknown kernel %d", o.kernel);
  const bool general = kernel == LNORM_KERNEL_GENERAL;
  const bool wide = kernel == LNORM_KERNEL_WIDE;
  const int ny = general ? C : wide ? ny_wide : ny_plain;
  const bool fp64 = !general && exact53;

  // matrix padded to the kernel width and the binary delta table
  std::vector<int64_t> hA((size_t)R * ny, 0), hD;
  for (int i = 0; i < R; ++i)
    for (int l = 0; l < C; ++l) hA[(size_t)i * ny + l] = A[(size_t)i * C + l];
  if (!general && p.n <= 2) {
    hD.assign((size_t)2 * std::max(R - 1, 1) * ny, 0);
    const int64_t f = p.n == 1 ? 2 : 1;
    for (int j = 0; j < R - 1; ++j)
      for (int l = 0; l < ny; ++l) {
        const int64_t v = (int64_t)((uint64_t)f * (uint64_t)hA[(size_t)j * ny + l]);
        hD[((size_t)(R - 1) + j) * ny + l] = v;                     // up: +f A
        hD[(size_t)j * ny + l] = (int64_t)(0ull - (uint64_t)v);       // down: -f A
      }
  }

  int dev = 0;
  if (int r = select_device(o, dev)) return r;
  DevCtx &cx = g_ctx[dev];
  std::lock_guard<std::mutex> lock(cx.mu);
  if (int r = ctx_init(cx)) return r;
  hipStream_t st = o.stream ? (hipStream_t)o.stream : cx.stream;
  const int nvec = (p.kind == KIND_L1 || p.kind == KIND_LM) ? 1 : p.n;
  if (int r = ensure(cx.A64, hA.size() * 8)) return r;
  if (int r = ensure(cx.D64, std::max<size_t>(hD.size(), 1) * 8)) return r;
  if (int r = ensure(cx.units64, kUnitsPerChunk * 8)) return r;
  if (int r = ensure(cx.scal, 256)) return r;
  if (general)
    if (int r = ensure(cx.state64, (size_t)kGeneralSlices * nvec * C * 8)) return r;
  if (fp64) {   // the same layouts as doubles (exact: every value is below 2^53 in magnitude)
    for (int64_t &v : hA) { const double x = (double)v; std::memcpy(&v, &x, 8); }
    for (int64_t &v : hD) { const double x = (double)v; std::memcpy(&v, &x, 8); }
  }
  HIPCHK(hipMemcpyAsync(cx.A64.p, hA.data(), hA.size() * 8, hipMemcpyHostToDevice, st));
  if (!hD.empty()) HIPCHK(hipMemcpyAsync(cx.D64.p, hD.data(), hD.size() * 8, hipMemcpyHostToDevice, st));

  lnorm_hip_result rr;
  std::memset(&rr, 0, sizeof rr);
  rr.total = re - rb;
  const Walk walk = general ? WALK_GENERAL : WALK_PLAIN;
  const Geometry g = geometry_for(p.base, R - 1, 20, o.seg_log, walk);
  double ms_kernel = 0;
  int64_t best = INT64_MIN;
  uint64_t first_unit = ~0ull;
  int first_e = g.q;
  // the ranks in runs of aligned units (one run of whole units for the whole
  // range), in rank order; every run on the same walk
  HIPCHK(hipEventRecord(cx.ev[0], st));
  for (const RankRun &run : split_ranks(p.base, g.q, rb, re)) {
    const Geometry gr{run.e, ipow((uint64_t)p.base, run.e), run.n};
    if (run.e == g.q) rr.units += run.n;
    for (uint64_t c0 = 0; c0 < run.n; c0 += kUnitsPerChunk) {
      const uint64_t nch = std::min(kUnitsPerChunk, run.n - c0);
      int64_t b;
      uint64_t f;
      if (int r = run64_phase(cx, p.kind, p.n, ny, general, wide, fp64, MODE_MAX, gr, run.u0 + c0, nch, R, 0, st,
                              ms_kernel, b, f))
        return r;
      if (first_unit == ~0ull || b > best) {
        best = b;
        first_unit = run.u0 + c0 + f;
        first_e = run.e;
      }
    }
  }
  HIPCHK(hipEventRecord(cx.ev[1], st));
  rr.unit_log = g.q;
  rr.kernel = kernel;
  if (o.want_strategy && R > 1) {
    const Geometry g2 = geometry_for(p.base, first_e, 16, 0, walk);
    double ms2 = 0;
    int64_t b;
    uint64_t rank;
    if (int r = run64_phase(cx, p.kind, p.n, ny, general, wide, fp64, MODE_FIRST, g2, first_unit * g2.nseg,
                            g2.nseg, R, best, st, ms2, b, rank))
      return r;
    if (rank == ~0ull) return fail(LNORM_E_HIP, "no rank attains the maximum %lld", (long long)best);
    rr.ms_strategy = ms2;
    rr.best_rank = rank;
    labels_of_rank(rank, R, n, strategy);
  } else {
    rr.best_rank = R == 1 ? 0 : ~0ull;
  }
  // without the strategy phase nothing has waited for the stream since ev[1]
  HIPCHK(hipEventSynchronize(cx.ev[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, cx.ev[0], cx.ev[1]));
  rr.ms_enumerate = ms;
  rr.ms_kernel = ms_kernel;
  rr.Lnorm = (best >= INT32_MIN && best <= INT32_MAX) ? (int32_t)best : 0;
  snprintf(rr.kernel_name, sizeof rr.kernel_name, "%s64<kind=%d,d=%d,ny=%d,q=%d>",
           general ? "general" : wide ? "widef" : fp64 ? "plainf" : "plain", (int)p.kind, p.n, ny, g.q);
  rr.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  *Lnorm = best;
  if (res) *res = rr;
  return LNORM_OK;
}

}  // namespace

extern "C" {

void lnorm_hip_default_opts(lnorm_hip_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->kernel = LNORM_KERNEL_AUTO;
  o->want_strategy = 1;
  o->device = -1;
}

int lnorm_hip_calc_ex(const int32_t *A, int R, int C, int n, int marginal, const lnorm_hip_opts *opts,
                      int32_t *strategy, lnorm_hip_result *res) {
  return calc_impl(A, R, C, n, marginal, opts, strategy, res);
}

int lnorm_hip_calc_noalloc(const int32_t *A, int R, int C, int n, int marginal, const lnorm_hip_opts *opts,
                           int32_t *strategy, lnorm_hip_result *res) {
  return calc_impl(A, R, C, n, marginal, opts, strategy, res, true);
}

int lnorm_hip_calc_ex64(const int64_t *A, int R, int C, int n, int marginal, const lnorm_hip_opts *opts,
                        int32_t *strategy, int64_t *Lnorm, lnorm_hip_result *res) {
  return calc_impl64(A, R, C, n, marginal, opts, strategy, Lnorm, res);
}

int lnorm_hip_calc_Lnorm64(const int64_t *A, int R, int C, int n, int marginal, int num_of_blocks,
                           int num_of_threads_per_block, int64_t *Lnorm, int32_t *strategy) {
  (void)num_of_blocks;
  (void)num_of_threads_per_block;
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  o.want_strategy = strategy != nullptr;
  return calc_impl64(A, R, C, n, marginal, &o, strategy, Lnorm, nullptr);
}

int lnorm_hip_calc_Lnorm(const int32_t *A, int R, int C, int n, int marginal, int num_of_blocks,
                         int num_of_threads_per_block, int32_t *Lnorm, int32_t *strategy) {
  (void)num_of_blocks;
  if (!Lnorm) return fail(LNORM_E_ARG, "Lnorm output is NULL");
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  if (num_of_threads_per_block > 0 && num_of_threads_per_block % 64 == 0 && num_of_threads_per_block <= 1024)
    o.block = num_of_threads_per_block;
  lnorm_hip_result r;
  const int rc = calc_impl(A, R, C, n, marginal, &o, strategy, &r);
  if (rc == 0) *Lnorm = r.Lnorm;
  return rc;
}

int lnorm_hip_calc_batch(lnorm_hip_item *items, int count, const lnorm_hip_opts *opts) {
  return lnorm_hip_calc_batch_ex(items, count, opts, nullptr);
}

int lnorm_hip_calc_batch_ex(lnorm_hip_item *items, int count, const lnorm_hip_opts *opts,
                            lnorm_hip_batch_stats *stats) {
  g_err.clear();
  if (count < 0 || (count > 0 && !items)) return fail(LNORM_E_ARG, "bad batch");
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  if (opts) o = *opts;
  return batch_impl(items, count, o, stats);
}

int lnorm_hip_autotune(const int32_t *A, int R, int C, int n, int marginal, const lnorm_hip_opts *opts,
                       lnorm_hip_opts *best, lnorm_hip_tune_entry *entries, int max_entries) {
  g_err.clear();
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  if (opts) o = *opts;
  if (max_entries < 0 || (max_entries > 0 && !entries)) return fail(LNORM_E_ARG, "bad entries buffer");
  Problem p{A, R, C, n, marginal, KIND_L1, 2, 0};
  if (int r = validate(p)) return r;
  int dev = 0;
  if (int r = select_device(o, dev)) return r;
  DevCtx &cx = g_ctx[dev];
  std::lock_guard<std::mutex> lock(cx.mu);
  if (int r = ctx_init(cx)) return r;
  hipStream_t st = o.stream ? (hipStream_t)o.stream : cx.stream;
  Choice ch;
  if (int r = choose_kernel(p, o, ch)) return r;
  std::vector<lnorm_hip_tune_entry> log;
  TuneVal tv;
  if (int r = tune_single(cx, p, ch, o, st, tv, &log)) return r;
  if (!log.empty()) {
    std::lock_guard<std::mutex> tl(g_tune_mu);
    g_tune[tune_key(dev, p, ch)] = tv;
  }
  if (best) {
    *best = o;
    best->kernel = ch.kernel;
    if (tv.P) best->patterns_per_lane = tv.P;
    if (tv.s) best->wave_log2 = tv.s;
    if (ch.kernel == LNORM_KERNEL_BLOCKED && !tv.P) {   // the default shape kept: report it
      best->patterns_per_lane = ch.bp.P;
      best->wave_log2 = ch.bp.s;
    }
    if (tv.block) best->block = tv.block;
  }
  for (int i = 0; i < (int)log.size() && i < max_entries; ++i) entries[i] = log[i];
  return (int)log.size();
}

int lnorm_hip_autotune_batch(const lnorm_hip_item *items, int count, const lnorm_hip_opts *opts,
                             lnorm_hip_opts *best, lnorm_hip_tune_entry *entries, int max_entries) {
  g_err.clear();
  if (count < 0 || (count > 0 && !items)) return fail(LNORM_E_ARG, "bad batch");
  if (max_entries < 0 || (max_entries > 0 && !entries)) return fail(LNORM_E_ARG, "bad entries buffer");
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  if (opts) o = *opts;
  std::vector<lnorm_hip_tune_entry> log;
  lnorm_hip_opts b;
  const int r = tune_batch(items, count, o, b, &log);
  if (r < 0) return r;
  if (best) *best = b;
  for (int i = 0; i < (int)log.size() && i < max_entries; ++i) entries[i] = log[i];
  return (int)log.size();
}

const char *lnorm_hip_last_error(void) { return g_err.c_str(); }

int lnorm_hip_abi_version(void) { return LNORM_HIP_ABI_VERSION; }

int lnorm_hip_device_count(void) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) return fail(LNORM_E_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
  return c;
}

void lnorm_hip_release(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
  DevCtx &c = g_ctx[dev];
  std::lock_guard<std::mutex> lock(c.mu);
  if (c.bGraph) (void)hipGraphExecDestroy(c.bGraph);
  c.bGraph = nullptr;
  c.bGraphSig.clear();
  for (Buf *b : {&c.A, &c.D, &c.units, &c.scal, &c.table, &c.dpk, &c.seed, &c.dfull, &c.lo, &c.state, &c.bA, &c.bD,
                 &c.bUnits, &c.bScal, &c.bTasks, &c.bCnt, &c.A64, &c.D64, &c.units64, &c.state64}) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->cap = 0;
  }
  if (c.init) {
    for (auto &e : c.ev) (void)hipEventDestroy(e);
    for (int ph = 0; ph < 2; ++ph) {
      (void)hipEventDestroy(c.bFork[ph]);
      for (auto &e : c.bJoin[ph]) (void)hipEventDestroy(e);
    }
    for (auto &st : c.side) (void)hipStreamDestroy(st);
    (void)hipStreamDestroy(c.stream);
    c.init = false;
  }
}

int lnorm_hip_selftest(void) {
  g_err.clear();
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  int dev = 0;
  if (int r = select_device(o, dev)) return r;
  DevCtx &cx = g_ctx[dev];
  std::lock_guard<std::mutex> lock(cx.mu);
  if (int r = ctx_init(cx)) return r;
  // edge values of both halves plus a SplitMix64 stream
  const uint32_t edge[] = {0u, 1u, 2u, 0x7FFFu, 0x8000u, 0x8001u, 0xFFFFu, 0x10000u, 0x7FFF8000u,
                           0x80008000u, 0x7FFFFFFFu, 0x80000000u, 0x80000001u, 0xFFFFFFFFu, 0x12345678u};
  const int ne = (int)(sizeof edge / sizeof edge[0]);
  const int n = ne * ne + 4096;
  std::vector<uint32_t> a(n), b(n), c(n), out(2 * (size_t)n);
  uint64_t x = 0x243F6A8885A308D3ull;
  auto next = [&]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
  };
  for (int i = 0; i < n; ++i) {
    if (i < ne * ne) {
      a[i] = edge[i / ne];
      b[i] = edge[i % ne];
      c[i] = (i & 1) ? 0u : 0x00100000u;
    } else {
      a[i] = next();
      b[i] = next();
      c[i] = next() >> 8;
    }
  }
  Buf bufs;
  if (int r = ensure(bufs, (size_t)5 * n * 4)) return r;
  uint32_t *da = (uint32_t *)bufs.p, *db = da + n, *dc = db + n, *dout = dc + n;
  HIPCHK(hipMemcpyAsync(da, a.data(), (size_t)n * 4, hipMemcpyHostToDevice, cx.stream));
  HIPCHK(hipMemcpyAsync(db, b.data(), (size_t)n * 4, hipMemcpyHostToDevice, cx.stream));
  HIPCHK(hipMemcpyAsync(dc, c.data(), (size_t)n * 4, hipMemcpyHostToDevice, cx.stream));
  HIPCHK(launch_sad_probe(da, db, dc, dout, n, cx.stream));
  HIPCHK(hipMemcpyAsync(out.data(), dout, (size_t)2 * n * 4, hipMemcpyDeviceToHost, cx.stream));
  HIPCHK(hipStreamSynchronize(cx.stream));
  (void)hipFree(bufs.p);
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    const uint32_t e32 = (a[i] > b[i] ? a[i] - b[i] : b[i] - a[i]) + c[i];
    const uint32_t al = a[i] & 0xFFFF, ah = a[i] >> 16, bl = b[i] & 0xFFFF, bh = b[i] >> 16;
    const uint32_t e16 = (al > bl ? al - bl : bl - al) + (ah > bh ? ah - bh : bh - ah) + c[i];
    if (out[2 * (size_t)i] != e32 || out[2 * (size_t)i + 1] != e16) {
      if (!bad)
        fail(LNORM_E_HIP, "sad mismatch at a=%08x b=%08x c=%08x: sad32 %08x (want %08x), sad16 %08x (want %08x)",
             a[i], b[i], c[i], out[2 * (size_t)i], e32, out[2 * (size_t)i + 1], e16);
      ++bad;
    }
  }
  return bad;
}

int lnorm_hip_valu_rate(int op, double *lane_ops_per_s, double *ms) {
  g_err.clear();
  if (op < 0 || op > 2) return fail(LNORM_E_ARG, "op must be 0 (v_sad_u16), 1 (v_sad_u32) or 2 (v_add_u32)");
  if (!lane_ops_per_s) return fail(LNORM_E_ARG, "lane_ops_per_s is NULL");
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  int dev = 0;
  if (int r = select_device(o, dev)) return r;
  DevCtx &cx = g_ctx[dev];
  std::lock_guard<std::mutex> lock(cx.mu);
  if (int r = ctx_init(cx)) return r;
  // 8 waves per SIMD (2048 workgroups of 4 waves over 256 CUs), about 120 ms
  // at the published rate, after a launch of half that length that brings the
  // clock to its sustained level
  const int blocks = 2048, iters = 32768;
  Buf out;
  if (int r = ensure(out, (size_t)blocks * 256 * 4)) return r;
  float t = 0;
  hipError_t e = launch_valu_rate(op, (uint32_t *)out.p, blocks, iters / 2, cx.stream);
  if (e == hipSuccess) e = hipEventRecord(cx.ev[4], cx.stream);
  if (e == hipSuccess) e = launch_valu_rate(op, (uint32_t *)out.p, blocks, iters, cx.stream);
  if (e == hipSuccess) e = hipEventRecord(cx.ev[5], cx.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(cx.stream);
  if (e == hipSuccess) e = hipEventElapsedTime(&t, cx.ev[4], cx.ev[5]);
  (void)hipFree(out.p);
  if (e != hipSuccess) return fail(LNORM_E_HIP, "VALU rate probe: %s", hipGetErrorString(e));
  if (t <= 0) return fail(LNORM_E_HIP, "VALU rate probe: no elapsed time");
  *lane_ops_per_s = (double)valu_rate_lane_ops(blocks, iters) / (t / 1e3);
  if (ms) *ms = t;
  return 0;
}

int lnorm_hip_debug_table(const int32_t *A, int R, int C, int marginal, int P, int word_bits, uint32_t *out,
                          int out_words) {
  g_err.clear();
  Problem p{A, R, C, 1, marginal, KIND_L1, 2, 0};
  if (int r = validate(p)) return r;
  lnorm_hip_opts o;
  lnorm_hip_default_opts(&o);
  o.patterns_per_lane = P;
  o.word_bits = word_bits;
  BlockedPlan bp;
  std::string why;
  if (!blocked_plan(p, o, bp, why)) return fail(LNORM_E_UNSUPPORTED, "blocked kernel not applicable: %s", why.c_str());
  const int words = bp.P * bp.NW * 64;
  if (!out || out_words < words) return fail(LNORM_E_ARG, "output needs %d words", words);
  int dev = 0;
  if (int r = select_device(o, dev)) return r;
  DevCtx &cx = g_ctx[dev];
  std::lock_guard<std::mutex> lock(cx.mu);
  if (int r = ctx_init(cx)) return r;
  std::vector<int8_t> lo;
  split_planes(p, bp.kl, bp.planes, lo);
  if (int r = ensure(cx.lo, lo.size())) return r;
  if (int r = ensure(cx.table, (size_t)words * 4)) return r;
  HIPCHK(hipMemcpyAsync(cx.lo.p, lo.data(), lo.size(), hipMemcpyHostToDevice, cx.stream));
  TableArgs ta{(const int8_t *)cx.lo.p, bp.planes, bp.kl, C, marginal, bp.np, bp.P, (uint32_t *)cx.table.p,
               bp.wide ? 1 : 0};
  HIPCHK(launch_table(ta, cx.stream));
  HIPCHK(hipMemcpyAsync(out, cx.table.p, (size_t)words * 4, hipMemcpyDeviceToHost, cx.stream));
  HIPCHK(hipStreamSynchronize(cx.stream));
  return words;
}

}  // extern "C"

// lnorm_internal.h -- shared declarations of the HIP engine (kernels and the
// C-ABI layer).  Device code targets gfx950 (wave64) only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lnorm {

// Which objective a kernel evaluates (the reference's kernel selection,
// L_OpenMP.c:251-290 / L_CUDA.cu:299-333).
enum Kind : int { KIND_L1 = 0, KIND_LM = 1, KIND_L2 = 2, KIND_LD = 3 };

// Phase of the two-phase search: MAX finds the norm and the best value of
// every unit (a contiguous rank range); FIRST re-walks one unit and returns
// the smallest Gray rank attaining a given target value.
enum Mode : int { MODE_MAX = 0, MODE_FIRST = 1 };

constexpr uint32_t kBias32 = 0x80000000u;   // y stored as y + 2^31 for v_sad_u32
constexpr uint32_t kBias16 = 0x8000u;       // per half for v_sad_u16
constexpr uint32_t kBias16x2 = 0x80008000u;

// Plain Gray walk: lane-per-segment.  The ranks [rank0, rank0 + nseg*L) are
// split into nseg segments of L = base^q ranks; segment `seg` (0-based,
// relative) starts at rank rank0 + seg*L.  rank0 must be a multiple of L.
struct PlainArgs {
  const int32_t *A;        // R x ny_pad, row-major, zero-padded columns
  const int32_t *D;        // binary: [2][R-1][ny_pad] signed row deltas
  int R;                   // rows (row R-1 has the fixed label)
  int ny_pad;              // padded column count (template NY)
  int d;                   // labels (1 = L1/LM, 2, 3..10)
  int q;                   // segment = d^q ranks (2^q for d <= 2)
  uint64_t seg0;           // first segment index (rank0 / L)
  uint64_t nseg;           // number of segments
  int32_t target;          // FIRST mode: value to find
  int32_t *unit_best;      // MAX mode: best per segment [nseg]
  int32_t *gbest;          // MAX mode: atomic max
  unsigned long long *gfirst;  // FIRST mode: atomic min of the rank
  uint32_t *state;         // general kernel: per-wave scratch, slices * nvec * ny words
  uint64_t slices;         // general kernel: concurrent waves (state slices); 0 = kGeneralSlices
};

// The int_type = long long variant: the same geometry with 64-bit values.
struct Plain64Args {
  const int64_t *A;        // R x ny_pad (general kernel: R x ny)
  const int64_t *D;        // binary: [2][R-1][ny_pad] signed row deltas
  int R, ny_pad, d, q;
  uint64_t seg0, nseg;
  int64_t target;          // FIRST mode: value to find
  int64_t *unit_best;      // MAX mode: best per segment [nseg]
  long long *gbest;        // MAX mode: atomic max
  unsigned long long *gfirst;
  uint64_t *state;         // general kernel: per-wave scratch, kGeneralSlices * nvec * ny words
  int fp64;                // binary walk: A and D hold doubles (sum |A| < 2^52), k_plainf64_bin
};

// The general kernel runs at most this many waves at once (its scratch size).
constexpr uint64_t kGeneralSlices = 2048;

// Blocked L_1 / L_M walk (low-row table + scalar high walk).
struct BlockedArgs {
  const uint32_t *table;   // [P][NP+MARG][64] packed low-pattern words (see kernel)
  const uint32_t *dpk;     // [2][nh][NW] packed high-row deltas applied to the scalar state
  const uint32_t *seed;    // [NW] packed scalar state at high rank 0 (before any flips)
  const uint32_t *dfull;   // [nh][NW] packed sign-flip deltas: seed flip of high row m
  int nh;                  // high rows walked by the Gray code
  int s;                   // high ranks per wave = 2^s
  uint64_t wave0;          // first wave index
  uint64_t nwaves;         // waves of this launch (< 2^32 - 1)
  unsigned long long *gkey;   // atomic max over the launch's waves of blocked_key(value, wave - wave0)
};

// The blocked kernels' result key of one wave (or workgroup) w of a launch:
// the value, made order-preserving as unsigned, above the complement of w, so
// the largest key holds the launch's maximum and the first wave attaining it.
// A zero key (the host's initial value) is below every wave's key.
__host__ __device__ inline unsigned long long blocked_key(int32_t v, uint64_t w) {
  return ((unsigned long long)((uint32_t)v ^ 0x80000000u) << 32) | (0xFFFFFFFFull - (uint32_t)w);
}
__host__ __device__ inline int32_t blocked_key_value(unsigned long long k) {
  return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
}
__host__ __device__ inline uint64_t blocked_key_wave(unsigned long long k) {
  return 0xFFFFFFFFull - (k & 0xFFFFFFFFull);
}

struct TableArgs {
  const int8_t *lo;        // low rows as int8 planes: [planes][kl][ny] (hi plane first)
  int planes;              // 1 (entries fit int8) or 2 (int16: hi/lo planes)
  int kl;                  // low rows (table has 2^kl patterns)
  int ny;                  // columns of the table (before packing)
  int marginal;            // LM: column 0 kept as an int32 word
  int np;                  // |.| words per pattern (excluding the marginal word)
  int P;                   // patterns per lane
  uint32_t *out;           // packed table
  int wide;                // 0: two biased 16-bit halves per word; 1: one Y + 2^31 word per column
};

// Launchers (return hipSuccess or the launch error).  `grid_hint` <= 0 means
// choose automatically.
// largest packed width (words of two columns) instantiated per d: d * NW <= 160
constexpr int dary16_max_words(int d) {
  return d == 2 ? 64 : d == 3 ? 48 : (d == 4 || d == 5) ? 32 : d == 6 ? 24 : (d >= 7 && d <= 10) ? 16 : 0;
}

// One item of a batch for one phase (device memory, read by the persistent
// batch kernels).  Tasks [task0, task0 + 2*ceil(nseg/128)) are its waves.
struct BatchTask {
  const int32_t *A;            // binary walk: R x NY; packed walk: R x NW words
  const int32_t *D;            // binary walk: [2][R-1][NY] deltas; tabulated walk: 2^M x NW subset table
  int R, q;                    // rows; segment = base^q ranks
  uint64_t nseg;               // segments of this item in this phase
  uint64_t seg_base;           // MODE_FIRST: segments per unit (seg0 = *unit * seg_base)
  uint64_t task0;              // first task of this item
  int32_t *unit_best;          // MODE_MAX: this item's unit maxima (nseg entries)
  int32_t *gbest;              // this item's maximum
  unsigned long long *gfirst;  // MODE_FIRST: this item's smallest maximising rank
  const unsigned long long *unit;  // MODE_FIRST: the first unit attaining the maximum
};

// The task counter of a persistent launch is sharded by XCD: shard x hands
// out tasks x, x + 8, x + 16, ... and lives in its own 4 KiB block, so the
// device-scope atomics of different XCDs do not queue on one address.
constexpr int kCounterShards = 8;
constexpr int kCounterStride = 1024;   // uint32 words between shards

// One persistent launch over the items of a group.
struct BatchLaunch {
  const BatchTask *tasks;      // nitems descriptors of this group (device)
  int nitems;
  uint64_t ntasks;             // < 2^31
  unsigned int *counter;       // kCounterShards shards, kCounterStride apart, zeroed before the launch
  Kind kind;
  int d;
  int width;                   // binary: NY; packed: 2 * NW
  bool packed;
  bool table;                  // packed and tabulated (walk_blk16)
};

hipError_t launch_batch_walk(Mode mode, const BatchLaunch &b, hipStream_t st);
hipError_t launch_batch_t16(Mode mode, const BatchLaunch &b, hipStream_t st);   // the tabulated d-ary groups
int batch_t16_words(int d, int ny);   // a batch item's tabulated width in words (0: none)
int batch_tb16_words(int nabs);       // a batch L1 / LM item's tabulated width in words (0: none)
hipError_t launch_batch_first_unit(const BatchTask *tasks, int nitems, unsigned long long *first, hipStream_t st);

hipError_t launch_plain(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st, int block);
// Packed 16-bit d-ary walk (d = 2..10): A holds R x (ny_pad/2) packed words.
hipError_t launch_plain16(Mode mode, const PlainArgs &a, hipStream_t st, int block);
int wide_ny_pad(Kind kind, int ny);   // 0: no W-wave instantiation for this width
hipError_t launch_wide(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st);
int wide16_words(int d, int ny);      // packed W-wave width in words (0: none)
int blk16_words(int d, int ny);       // tabulated packed walk: width in words (0: none)
int blk16_rows(int d, int words);     // its tabulated rows M
int wt16_words(int d, int ny);        // the W-wave tabulated walk: width in words (0: none)
int wt16_rows(int d, int words);      // its tabulated rows M
hipError_t launch_wide_t16(Mode mode, const PlainArgs &a, hipStream_t st);
int blk16_blocks_per_cu(int d, int words, int block);             // resident workgroups of the
int tab16_bin_blocks_per_cu(Kind kind, int words, int block);     // tabulated walks' kernels (0: unknown)
int wtb16_words(int nabs);            // the W-wave tabulated L1 / LM walk: width in words (0: none)
int wtb16_rows(int words);            // its tabulated rows M
hipError_t launch_wide_tb16(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st);
hipError_t launch_blk16(Mode mode, const PlainArgs &a, hipStream_t st, int block);
int tab16_bin_words(int nabs);        // tabulated L1/LM walk: width in words for nabs |.| columns (0: none)
// its table rows: the sign patterns of rows 0..M-1; 4 beyond 64 words keeps
// the table (2^M x NW scalar words) and the unrolled block within bounds
constexpr int tab_bin_rows(int nw) { return nw <= 64 ? 5 : 4; }
hipError_t launch_tab16_bin(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st, int block);
// The tabulated walks in 32-bit words (one column per word, biased by 2^31,
// v_sad_u32): int32 inputs whose column sums exceed the 16-bit walks' bound.
// Their widths are in words, one column each (PlainArgs::ny_pad = words).
// L1 / LM table rows in one wave: 5 up to 128 words, 4 beyond.  The table
// (2^M x NW words, read by scalar loads every block) stays within 16 KiB: 6
// rows at 128 words (32 KiB) ran 1.4x slower than 5.
constexpr int tab32_bin_rows(int nw) { return nw <= 128 ? 5 : 4; }
int tab32_bin_words(int nabs);        // one wave, up to 192 |.| columns (0: none)
int tab32_bin_blocks_per_cu(Kind kind, int words, int block);
hipError_t launch_tab32_bin(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st, int block);
int wtb32_words(int nabs);            // W waves (the 16-bit walk's shapes): 193..1024 |.| columns (0: none)
int wtb32_rows(int words);
hipError_t launch_wide_tb32(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st);
int blk32_words(int d, int ny);       // d-ary, one wave: up to 96 columns at d = 2 ... 16 at d = 10 (0: none)
int blk32_rows(int d, int words);
int blk32_blocks_per_cu(int d, int words, int block);
hipError_t launch_blk32(Mode mode, const PlainArgs &a, hipStream_t st, int block);
int wt32_words(int d, int ny);        // d-ary, W waves of 32 (d = 8, 9: 24; 10: 16) words (0: none)
int wt32_rows(int d, int words);
hipError_t launch_wide_t32(Mode mode, const PlainArgs &a, hipStream_t st);
hipError_t launch_wide16(Mode mode, const PlainArgs &a, hipStream_t st);
int plain16_words(int d, int ny);               // packed width instantiated for ny columns, 0 if none
bool plain_supported(Kind kind, int d, int ny_pad);
int plain_ny_pad(Kind kind, int d, int ny);     // smallest instantiated width >= ny, 0 if none

hipError_t launch_blocked(int np, int P, int marginal, const BlockedArgs &a, hipStream_t st);
bool blocked_supported(int np, int P, int marginal);
int blocked_np_pad(int np, int P, int marginal);

// 32-bit blocked walk: nc words per wave, W waves per workgroup (nwaves in
// BlockedArgs counts workgroups, one unit each).
hipError_t launch_blocked32(int nc, int P, int W, int marginal, const BlockedArgs &a, hipStream_t st);
bool blocked32_supported(int nc, int P, int W, int marginal);

hipError_t launch_table(const TableArgs &a, hipStream_t st);

hipError_t launch_general(Kind kind, Mode mode, const PlainArgs &a, int ny, hipStream_t st);

hipError_t launch_fill_i32(int32_t *p, int32_t v, uint64_t n, hipStream_t st);
hipError_t launch_fill_i64(long long *p, long long v, uint64_t n, hipStream_t st);
hipError_t launch_first_unit64(const int64_t *unit_best, uint64_t n, const long long *gbest,
                               unsigned long long *first, hipStream_t st);
int plain64_ny_pad(Kind kind, int d, int ny);   // 0: no register-resident 64-bit width
hipError_t launch_plain64(Kind kind, Mode mode, const Plain64Args &a, hipStream_t st);
hipError_t launch_general64(Kind kind, Mode mode, const Plain64Args &a, int ny, hipStream_t st);
// the walks in doubles over W waves (a.fp64 set; binary and d-ary): the
// width for ny columns (0: none) and the launch
int widef64_ny_pad(Kind kind, int d, int ny);
hipError_t launch_widef64(Kind kind, Mode mode, const Plain64Args &a, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, uint64_t n, hipStream_t st);
hipError_t launch_first_unit(const int32_t *unit_best, uint64_t n, const int32_t *gbest,
                             unsigned long long *first, hipStream_t st);
hipError_t launch_sad_probe(const uint32_t *a, const uint32_t *b, const uint32_t *c, uint32_t *out, int n,
                            hipStream_t st);
// lnorm_peak.hip: the VALU rate probe (op 0: v_sad_u16, 1: v_sad_u32, 2:
// v_add_u32; 256-thread workgroups) and the lane-ops one launch issues
hipError_t launch_valu_rate(int op, uint32_t *out, int blocks, int iters, hipStream_t st);
uint64_t valu_rate_lane_ops(int blocks, int iters);

}  // namespace lnorm

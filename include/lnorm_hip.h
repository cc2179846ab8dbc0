/*
 * lnorm_hip.h -- C ABI of the MI355X (gfx950) L_d enumeration engine.
 *
 * Drop-in for the reference's compute boundary
 *     void calc_Lnorm(item_calc* second, int* num_of_blocks,
 *                     int* num_of_threads_per_block)        (L_CUDA.cu:281-359)
 *     void calc_Lnorm(item_calc* second)                    (L_OpenMP.c:241-297)
 * which reads second->{n, iRows_reduced, iCols_reduced, marginal, mtx_as_vec}
 * (functions.h:7-19) and writes second->Lnorm and second->strategy.
 *
 * Plain C types only; no HIP or torch types cross this boundary.  All entry
 * points return 0 on success or a negative LNORM_E_* code; the message of the
 * last failure on the calling thread is lnorm_hip_last_error().  Unlike the
 * reference, whose CUDA errors are reported with perror and then ignored
 * (L_CUDA.cu:292-358, README.md:98), a failure here never yields a norm.
 */
#ifndef LNORM_HIP_H
#define LNORM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LNORM_HIP_ABI_VERSION 7

/* error codes */
#define LNORM_OK 0
#define LNORM_E_ARG (-1)        /* invalid argument (sizes, d, NULL pointers) */
#define LNORM_E_BOUND (-2)      /* sum |A| > INT32_MAX: the reference's int arithmetic would overflow */
#define LNORM_E_SIZE (-3)       /* beyond the reference limits (rows > maxRows, cols > 4096, d > 10) */
#define LNORM_E_HIP (-4)        /* HIP runtime error (no device, launch or copy failure) */
#define LNORM_E_UNSUPPORTED (-5) /* requested kernel cannot run this problem */

/* kernel selection */
#define LNORM_KERNEL_AUTO 0     /* n = 1: blocked from 2^33 sign vectors at up to 40 |.| columns, else
                                   table16; n >= 2: table16; then table32, blocked (32-bit words), packed
                                   plain16, plain, wide, columns, general */
#define LNORM_KERNEL_PLAIN 1    /* register-resident Gray walk, lane per segment */
#define LNORM_KERNEL_BLOCKED 2  /* L1/LM: MFMA-built low-row table + scalar high walk (16- or 32-bit words) */
#define LNORM_KERNEL_GENERAL 3  /* any width: wave per segment, columns over lanes */
#define LNORM_KERNEL_PLAIN16 4  /* n >= 2: packed 16-bit Gray walk (column sums <= 32767) */
#define LNORM_KERNEL_WIDE 5     /* W waves split the columns: n <= 2 at 129..1024 columns (L2: ..512);
                                   n = 3..10 packed, column sums <= 32767, up to 640 (d = 3) .. 192 (d >= 9) */
#define LNORM_KERNEL_TABLE16 6  /* packed walk over the high rows; the 1..5 fastest rows are tabulated:
                                   per-label subset tables for n = 2..10 (column sums <= 32767; one wave
                                   up to 192 columns at d = 2 ... 32 at d = 10, W waves up to 512 at
                                   d <= 6 ... 128 at d = 10), sign-pattern tables for n = 1 (|.| column
                                   sums <= 32767, one wave up to 384 columns, 8 to 16 waves up to 2048) */
#define LNORM_KERNEL_TABLE32 7  /* the tabulated walks in 32-bit words (one column per word, v_sad_u32): any
                                   int32 input; n = 1 up to 192 |.| columns in one wave, 1024 over W waves;
                                   n = 2..10 up to 96 (d = 2) .. 16 (d = 10) columns in one wave, 256
                                   (d <= 6) .. 64 (d = 10) over W waves.  AUTO takes it where the 16-bit
                                   walks do not apply (column sums beyond 32767) */
#define LNORM_KERNEL_COLUMNS 8  /* the column walk: one wave per segment, the columns over the lanes, label sums
                                   in registers, the rows in LDS; any int32 input up to 4096 columns for n <= 2,
                                   2048 at n = 3, 4 ... 768 at n = 9, 10, while the rows fit one workgroup's LDS
                                   (R x 64 x ceil(C / 64) x 4 B <= 160 KiB).  AUTO takes it where no other walk
                                   but the general kernel applies */

typedef struct lnorm_hip_opts {
    int32_t kernel;             /* LNORM_KERNEL_* */
    int32_t want_strategy;      /* nonzero: recover the strategy (smallest maximising Gray rank) */
    int32_t seg_log;            /* plain/general: segment = base^seg_log ranks per lane; 0 = auto */
    int32_t patterns_per_lane;  /* blocked: 1, 2, 4 or 8 low patterns per lane; 0 = auto */
    int32_t wave_log2;          /* blocked: 2^wave_log2 high ranks per wave; 0 = auto */
    int32_t block;              /* plain kernels: threads per workgroup, 64..256 in steps of 64; else 256 */
    int32_t device;             /* HIP device ordinal; -1 = current device */
    int32_t word_bits;          /* blocked: 16 (two halves per word) or 32; 0 = 16 when column sums allow;
                                   the long long variant: 64 = integer walk, 0 = doubles while exact */
    void *stream;               /* hipStream_t to run on; NULL = the library's own stream */
    int32_t batch_graph;        /* batches: capture the launch sequence in a HIP graph and replay it
                                   while the batch's structure (groups, sizes) stays the same */
    int32_t autotune;           /* nonzero: use launch shapes chosen by lnorm_hip_autotune (cached) */
    int32_t batch_streams;      /* batches: streams the kernel groups spread over, 1..4; 0 = auto */
    int32_t batch_cap_log;      /* batches: at most ~2^cap ranks per lane segment, 8..20; 0 = auto */
    /* Single problems: enumerate only the Gray ranks [rank_begin, rank_end): one shard of the rank space,
       as calc_jmin_jmax hands a worker its words (L_OpenMP.c:6-13).  The result is that range's maximum
       and its smallest maximising rank.  rank_end = 0 means total_num_to_calc, the reference's count
       (L_OpenMP.c:352-367): 2^(k-1) for n <= 2, (unsigned long long) pow(n, k-1) beyond, which is inexact
       from 2^53 on; the engine enumerates [0, min(that count, n^(k-1))), which gives the reference's
       result either way (ranks from n^(k-1) on repeat earlier labelings and never win its strict
       comparison).  Batches refuse nonzero values. */
    uint64_t rank_begin;
    uint64_t rank_end;
} lnorm_hip_opts;

typedef struct lnorm_hip_result {
    int32_t Lnorm;              /* second->Lnorm */
    int32_t kernel;             /* LNORM_KERNEL_* actually used */
    uint64_t total;             /* ranks enumerated: rank_end - rank_begin (by default total_num_to_calc,
                                   L_OpenMP.c:352-367) */
    uint64_t best_rank;         /* smallest Gray rank attaining Lnorm (if the strategy was recovered) */
    double ms_enumerate;        /* HIP-event time of the enumeration phase (table, walk, reduction) */
    double ms_kernel;           /* HIP-event time of the enumeration kernel launches alone */
    double ms_strategy;         /* HIP-event time of the strategy recovery */
    double ms_total;            /* host wall time of the call, copies included */
    uint64_t units;             /* units (segments or waves) of the enumeration launch */
    int32_t unit_log;           /* log_base of ranks per lane-segment (plain) or per wave (blocked) */
    int32_t patterns_per_lane;  /* blocked: P used */
    char kernel_name[64];       /* instantiation that ran, e.g. "blocked16<np=20,P=4,marg=0>" */
} lnorm_hip_result;

/*
 * calc_Lnorm replacement (L_CUDA.cu:281).
 *   mtx_as_vec : host int32, iRows x iCols row-major, as convert_mtx_to_vec
 *                leaves it (functions.h:291-329): transposed for d=1 when
 *                needed, marginal row swapped to the last position.
 *   n          : effective order second->n (1..10) after calc_Parameters.
 *   marginal   : second->marginal (only meaningful for n == 1).
 *   num_of_blocks, num_of_threads_per_block : the CLI's launch hints
 *                (L_CUDA.cu:361-412).  The result does not depend on them: the
 *                engine sizes its own launch, and the strategy is always the
 *                one the reference reports (smallest maximising Gray rank).
 *   Lnorm      : out, second->Lnorm.
 *   strategy   : out, caller-allocated iRows-1 ints (second->strategy):
 *                0/1 for n <= 2 (1 = +1 sign / label 1), labels 0..n-1 for n >= 3.
 */
int lnorm_hip_calc_Lnorm(const int32_t *mtx_as_vec, int iRows, int iCols, int n, int marginal,
                         int num_of_blocks, int num_of_threads_per_block,
                         int32_t *Lnorm, int32_t *strategy);

/* Same computation with explicit options and timing (opts may be NULL;
 * strategy may be NULL when opts->want_strategy == 0; res may be NULL). */
int lnorm_hip_calc_ex(const int32_t *mtx_as_vec, int iRows, int iCols, int n, int marginal,
                      const lnorm_hip_opts *opts, int32_t *strategy, lnorm_hip_result *res);

/*
 * lnorm_hip_calc_ex without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code that
 * must not call hipMalloc/hipFree, such as tight serving loops or harnesses
 * that forbid it).  Arguments, results and kernels are those of
 * lnorm_hip_calc_ex; the caller still passes the matrix, the strategy and
 * the result in host memory.
 *   Device memory: every buffer (matrix copy, tables, unit maxima, results,
 *   general-kernel state) is carved from one static array of
 *   LNORM_NOALLOC_SCRATCH_BYTES per device, which each device that loads the
 *   library holds until the process ends.  The call never calls hipMalloc or
 *   hipFree.  Its first call on a device creates that device's stream and
 *   timing events and looks up the array's address.
 *   Limits: enumeration launches hold at most 2^20 units (longer problems
 *   run in more launches); the general kernel runs as many concurrent waves
 *   as its per-wave state fits in the rest of the array.  A problem whose
 *   tables do not fit returns LNORM_E_SIZE with the bytes needed; within the
 *   reference's limits (63 rows, 4096 columns, d <= 10) every int32 problem
 *   fits.
 *   Sharing: all calls on one device share its array, which keeps the last
 *   call's contents; each call writes every buffer it reads before reading
 *   it.  Calls that overlapped on one device would race on the array: the
 *   library runs them one at a time (a per-device lock, and each call's
 *   device work completes before it returns).
 */
#define LNORM_NOALLOC_SCRATCH_BYTES (8u << 20)
int lnorm_hip_calc_noalloc(const int32_t *mtx_as_vec, int iRows, int iCols, int n, int marginal,
                           const lnorm_hip_opts *opts, int32_t *strategy, lnorm_hip_result *res);

/*
 * The int_type = long long variant of calc_Lnorm (README.md:100-105:
 * functions.h:5 typedef long long int_type, abs -> llabs): entries and the
 * norm in int64, every sum exact while sum |A| <= INT64_MAX (else
 * LNORM_E_BOUND).  Problems inside the int32 contract run on the int32
 * kernels (the same result); the rest on 64-bit walks (no blocked or packed
 * 16-bit counterpart: opts->kernel may be AUTO, PLAIN, WIDE or GENERAL).
 * The plain walks run in doubles while sum |A| < 2^52, where every value is
 * an integer below 2^53 and each add is exact (kernel name plainf64<...>),
 * and there the walks also run over W waves beyond one wave's width (WIDE,
 * widef64<...>: L1 / LM up to 512 columns, L2 up to 256, d = 3 up to 192,
 * d = 4 up to 128, d = 5..10 up to 64);
 * opts->word_bits = 64 keeps the 64-bit integer walks (plain64<...>, and the
 * general kernel beyond their widths), 0 chooses; other values are
 * LNORM_E_ARG.
 * res->Lnorm holds the norm only when it fits int32 (else 0); *Lnorm always.
 */
int lnorm_hip_calc_Lnorm64(const int64_t *mtx_as_vec, int iRows, int iCols, int n, int marginal,
                           int num_of_blocks, int num_of_threads_per_block,
                           int64_t *Lnorm, int32_t *strategy);
int lnorm_hip_calc_ex64(const int64_t *mtx_as_vec, int iRows, int iCols, int n, int marginal,
                        const lnorm_hip_opts *opts, int32_t *strategy, int64_t *Lnorm,
                        lnorm_hip_result *res);

/* One problem of a batch (cfg: many matrices of uneven shape and d). */
typedef struct lnorm_hip_item {
    const int32_t *mtx_as_vec;
    int32_t iRows, iCols, n, marginal;
    int32_t *strategy;          /* iRows-1 ints, or NULL */
    int32_t Lnorm;              /* out */
    int32_t status;             /* out: LNORM_OK or LNORM_E_* for this item */
} lnorm_hip_item;

/* Computes every item; returns LNORM_OK when all items succeeded, else the
 * first failing item's code (every item's own status is filled in).
 * Items that fit the batched walks (at most 2^32 labelings; n = 1 with at
 * most 128 columns; n >= 2 with column sums <= 32767) run together: one
 * persistent launch per kernel
 * instantiation for the enumeration and one for the strategies, with no host
 * synchronisation in between (opts->batch_graph replays it as a HIP graph).
 * Other items are computed one by one.  Results equal lnorm_hip_calc_Lnorm's. */
int lnorm_hip_calc_batch(lnorm_hip_item *items, int count, const lnorm_hip_opts *opts);

typedef struct lnorm_hip_batch_stats {
    int32_t items, batched, fallback, groups;
    int32_t graph_replayed;     /* 1 when an already captured graph was replayed */
    int32_t reserved;
    uint64_t labelings;         /* sum of total_num_to_calc over the batched items */
    double ms_device;           /* HIP-event time of the batched launch sequence */
    double ms_total;            /* host wall time of the call */
} lnorm_hip_batch_stats;

/* lnorm_hip_calc_batch with statistics (stats may be NULL). */
int lnorm_hip_calc_batch_ex(lnorm_hip_item *items, int count, const lnorm_hip_opts *opts,
                            lnorm_hip_batch_stats *stats);

/* One timed candidate of the autotuner. */
typedef struct lnorm_hip_tune_entry {
    int32_t kernel;             /* LNORM_KERNEL_* of the candidate */
    int32_t patterns_per_lane;  /* blocked: P */
    int32_t wave_log2;          /* blocked: 2^wave_log2 high ranks per wave */
    int32_t block;              /* plain kernels: threads per workgroup */
    int32_t batch_streams;      /* batches: streams the groups spread over */
    int32_t batch_cap_log;      /* batches: lane-segment cap */
    double ms;                  /* HIP-event time of the probe */
    double labelings_per_s;     /* labelings the probe covered per second */
} lnorm_hip_tune_entry;

/*
 * Launch-shape autotuner.  No counterpart in the reference, whose launch
 * shape is the CLI's blocks x threads (L_CUDA.cu:361-412) and does not
 * change results here either.
 *
 * Times the candidate shapes of the kernel opts->kernel (AUTO: the one AUTO
 * picks) would run this problem with:
 *   blocked: patterns per lane 1/2/4/8 x 2^10..2^14 high ranks per wave;
 *   plain, plain16: 64/128/256 threads per workgroup.
 * Each probe is that kernel over a prefix of this problem's own units.
 * The fastest shape is stored in the library's per-device cache, keyed by
 * (kernel, kind, d, padded width, marginal, word size), and replaces any
 * earlier entry.  Calls with opts->autotune != 0 use the cache (tuning on a
 * miss); options the caller sets explicitly win.
 *   best    : out (may be NULL), opts with the tuned fields filled in.
 *   entries : out (may be NULL), up to max_entries timed candidates.
 * Returns the number of candidates timed (0 for kernels without tunable
 * shapes), or a negative LNORM_E_* code.
 */
int lnorm_hip_autotune(const int32_t *mtx_as_vec, int iRows, int iCols, int n, int marginal,
                       const lnorm_hip_opts *opts, lnorm_hip_opts *best,
                       lnorm_hip_tune_entry *entries, int max_entries);

/*
 * Batch tuner: times batch_streams 1/2/4 x batch_cap_log 10/12/14/16 on
 * every 8th item of the batch (all items when fewer than 64), through the
 * batched path without strategies.  The fastest pair is cached under the
 * batch's shape (the multiset of item sizes, orders and marginal flags) and
 * used by lnorm_hip_calc_batch(_ex) when opts->autotune != 0 and the caller
 * leaves batch_streams and batch_cap_log at 0.  best and entries as above.
 * Returns the number of candidates timed or a negative LNORM_E_* code.
 * Items are not modified.
 */
int lnorm_hip_autotune_batch(const lnorm_hip_item *items, int count, const lnorm_hip_opts *opts,
                             lnorm_hip_opts *best, lnorm_hip_tune_entry *entries, int max_entries);

/* Fills opts with the defaults (AUTO kernel, strategy on, current device). */
void lnorm_hip_default_opts(lnorm_hip_opts *opts);

/* Message of the last failure on the calling thread ("" if none). */
const char *lnorm_hip_last_error(void);

/* LNORM_HIP_ABI_VERSION of the loaded library. */
int lnorm_hip_abi_version(void);

/* Number of visible HIP devices, or a negative LNORM_E_HIP. */
int lnorm_hip_device_count(void);

/* Releases the library's device buffers and stream on the current device. */
void lnorm_hip_release(void);

/* Device self-test of the absolute-difference instructions the kernels are
 * built on (v_sad_u32, v_sad_u16) against their definitions on edge and
 * pseudo-random operands.  Returns the number of mismatches (0 = pass) or a
 * negative LNORM_E_* code. */
int lnorm_hip_selftest(void);

/* Measurement aid, with no counterpart in the reference (it has no device
 * code to measure): the sustained rate of one VALU instruction the walks are
 * built on, on the current device, for the roofline's denominator.
 *   op 0: v_sad_u16, 1: v_sad_u32 (each with a wave-uniform SGPR operand, as
 *   in the walks), 2: v_add_u32.  Eight independent chains per lane, 8 waves
 *   per SIMD, no memory traffic in the loop; one launch of about 60 ms to
 *   bring the clock to its sustained level, then one timed launch of about
 *   120 ms.
 *   *lane_ops_per_s: lane-operations per second (the published peak is
 *   256 CUs x 64 lanes x 2.4 GHz = 3.93e13); *ms (may be NULL): the timed
 *   launch.  Returns 0 or a negative LNORM_E_* code. */
int lnorm_hip_valu_rate(int op, double *lane_ops_per_s, double *ms);

/* Diagnostic: the low-pattern table the blocked kernel would use
 * (P*(words+marginal)*64 words) for the first kl = 6+log2(P) rows of A, in
 * the 16-bit packed (word_bits 16) or 32-bit (word_bits 32) format; 0 = the
 * format AUTO would pick.  Returns the number of words written, or a
 * negative LNORM_E_* code. */
int lnorm_hip_debug_table(const int32_t *mtx_as_vec, int iRows, int iCols, int marginal, int P,
                          int word_bits, uint32_t *out, int out_words);

#ifdef __cplusplus
}
#endif
#endif /* LNORM_HIP_H */

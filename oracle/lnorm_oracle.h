/*
 * lnorm_oracle.h -- CPU restatement of the reference enumeration.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline).  The product (the HIP library and
 * the L_HIP host) never links it.
 *
 * Matrix convention: A is the reference's mtx_as_vec, row-major R x C int32
 * (int64 for the *64 functions)
 * (functions.h:291-329).  Row R-1 carries the fixed label (+1 for d=1,
 * label 0 for d>=2); rows 0..R-2 are enumerated.  n is the effective order
 * (item_calc.n after calc_Parameters, L_OpenMP.c:346-350).
 */
#ifndef LNORM_ORACLE_H
#define LNORM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* total_num_to_calc, as calc_Parameters computes it (L_OpenMP.c:352-367). */
uint64_t lo_total(int R, int n);

/* maxRows, as calc_Parameters computes it (L_OpenMP.c:354-366). */
int lo_max_rows(int n);

/* calc_jmin_jmax (L_OpenMP.c:6-13): inclusive Gray-rank segment of worker idx. */
void lo_segment(int idx, uint64_t steps, uint64_t rem, uint64_t *jmin, uint64_t *jmax);

/* Label of row i at Gray rank N: BRGC bit for n<=2, reflected n-ary digit
 * P[(N/n^i) mod 2n] for n>=3 (L_OpenMP.c:63,101,140,204). */
int lo_digit(uint64_t N, int i, int n);

/* One worker's walk over the inclusive rank range [jmin, jmax], following the
 * incremental updates of L1/LM/L2/L3/Ln (L_OpenMP.c:15-239).  Writes the best
 * objective, the rank where it first occurred (strict '>'), and that
 * strategy (R-1 labels).  Returns 0, or -1 on bad arguments. */
int lo_walk(const int32_t *A, int R, int C, int n, int marginal,
            uint64_t jmin, uint64_t jmax,
            int32_t *best, uint64_t *best_rank, int32_t *strategy);

/* calc_Lnorm (L_OpenMP.c:241-297): copyNum = min(threads, total) OpenMP
 * workers on calc_jmin_jmax segments, then the first-maximum argmax. */
int lo_calc_lnorm(const int32_t *A, int R, int C, int n, int marginal, int threads,
                  int32_t *norm, int32_t *strategy, uint64_t *rank_out);

/* Independent brute force for tiny inputs: every rank N in [0,total) is
 * decoded directly (no Gray increments) and its objective recomputed from
 * scratch; returns the maximum and the smallest rank attaining it. */
int lo_brute(const int32_t *A, int R, int C, int n, int marginal,
             int32_t *norm, uint64_t *rank_out);

/* Objective of one labeling (labels[R-1], row R-1 implied), recomputed from
 * scratch: sum_l |y_l| (L1), y_0 + sum_{l>=1}|y_l| (LM), or
 * sum_c sum_l |t_{c,l}| (n>=2). */
int32_t lo_objective(const int32_t *A, int R, int C, int n, int marginal,
                     const int32_t *labels);

/* The int_type = long long variant (README.md:100-105, abs -> llabs):
 * the same functions on an int64 matrix with int64 objectives.  The int32
 * entry points above widen the matrix and call these. */
int lo_walk64(const int64_t *A, int R, int C, int n, int marginal,
              uint64_t jmin, uint64_t jmax,
              int64_t *best, uint64_t *best_rank, int32_t *strategy);
int lo_calc_lnorm64(const int64_t *A, int R, int C, int n, int marginal, int threads,
                    int64_t *norm, int32_t *strategy, uint64_t *rank_out);
int lo_brute64(const int64_t *A, int R, int C, int n, int marginal,
               int64_t *norm, uint64_t *rank_out);
int64_t lo_objective64(const int64_t *A, int R, int C, int n, int marginal,
                       const int32_t *labels);

#ifdef __cplusplus
}
#endif
#endif

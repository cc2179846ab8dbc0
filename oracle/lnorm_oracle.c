/*
 * lnorm_oracle.c -- CPU restatement of the reference enumeration
 * (L_OpenMP.c, functions.h).  TEST INFRASTRUCTURE ONLY: it is the checker
 * for the HIP engine and the CPU baseline of bench.py; the product never
 * links it.
 *
 * Pinning (see tests/test_oracle.py):
 *   - lo_calc_lnorm / lo_walk: golden vectors produced by the reference
 *     L_OpenMP built from source (tests/golden/), the independent brute
 *     force lo_brute, closed forms (CHSH, identity, all-ones, rank one,
 *     d >= k cap) and invariants (transpose, permutations, sign flips,
 *     scaling, worker-count independence of the strategy).
 *   - lo_brute: closed forms and hand-checked tiny cases; it shares no code
 *     path with the Gray walks (direct digit decoding, full recompute).
 *
 * Integer arithmetic: the reference computes in int (int_type,
 * functions.h:5).  The engine's contract requires sum|A| <= INT32_MAX, under
 * which every partial sum and objective fits in int32; here updates are
 * done in uint32 (wrapping, so 2*A[i] never invokes signed-overflow UB) and
 * objectives are summed in int64, which is identical inside the contract.
 */
#include "lnorm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LO_MAX_LABELS 10   /* RANK_OF_NORM, functions.h:3 */

static inline int32_t iabs32(int32_t v) { return v < 0 ? -v : v; }

uint64_t lo_total(int R, int n)
{
    /* L_OpenMP.c:352-367 */
    if (R <= 0) return 0;
    if (n <= 2) return (uint64_t)1 << (R - 1);
    return (uint64_t)pow((double)n, (double)(R - 1));
}

int lo_max_rows(int n)
{
    /* L_OpenMP.c:354, 358, 362, 366 */
    if (n <= 2) return 64 - 1;
    return (int)(floor(64.0 / log2((double)n)) + 1);
}

void lo_segment(int idx, uint64_t steps, uint64_t rem, uint64_t *jmin, uint64_t *jmax)
{
    /* L_OpenMP.c:6-13: the first `rem` workers own steps+1 ranks. */
    uint64_t i = (uint64_t)idx;
    uint64_t lo = i * steps + (i < rem ? i : rem);
    uint64_t cnt = steps + (i < rem ? 1 : 0);
    *jmin = lo;
    *jmax = lo + cnt - 1;
}

int lo_digit(uint64_t N, int i, int n)
{
    if (n <= 2) {
        /* L_OpenMP.c:63: floor((N + 2^i) / 2^(i+1)) mod 2 == bit i of N^(N>>1) */
        return (int)(((N + ((uint64_t)1 << i)) >> (i + 1)) & 1);
    }
    /* L_OpenMP.c:140,204 with the pattern of functions.h:350-359:
     * P[(N / n^i) mod 2n], P = 0,1,..,n-1,n-1,..,0 */
    uint64_t p = 1;
    for (int t = 0; t < i; ++t) p *= (uint64_t)n;
    uint64_t r = (N / p) % (uint64_t)(2 * n);
    return r < (uint64_t)n ? (int)r : (int)(2 * n - 1 - r);
}

int32_t lo_objective(const int32_t *A, int R, int C, int n, int marginal, const int32_t *labels)
{
    int64_t total = 0;
    if (n == 1) {
        for (int l = 0; l < C; ++l) {
            uint32_t y = (uint32_t)A[(size_t)(R - 1) * C + l];
            for (int i = 0; i < R - 1; ++i) {
                uint32_t a = (uint32_t)A[(size_t)i * C + l];
                y = labels[i] ? y + a : y - a;
            }
            int32_t v = (int32_t)y;
            total += (marginal && l == 0) ? v : iabs32(v);
        }
        return (int32_t)total;
    }
    for (int c = 0; c < n; ++c) {
        for (int l = 0; l < C; ++l) {
            uint32_t t = c == 0 ? (uint32_t)A[(size_t)(R - 1) * C + l] : 0u;
            for (int i = 0; i < R - 1; ++i)
                if (labels[i] == c) t += (uint32_t)A[(size_t)i * C + l];
            total += iabs32((int32_t)t);
        }
    }
    return (int32_t)total;
}

/* ---- the Gray walks (L_OpenMP.c:15-239) ---------------------------------- */

static int32_t sum_abs_l1(const uint32_t *y, int C, int marginal)
{
    int64_t s = 0;
    for (int l = 0; l < C; ++l) {
        int32_t v = (int32_t)y[l];
        s += (marginal && l == 0) ? v : iabs32(v);
    }
    return (int32_t)s;
}

static int32_t sum_abs_labels(const uint32_t *t, int n, int C)
{
    int64_t s = 0;
    for (int c = 0; c < n; ++c)
        for (int l = 0; l < C; ++l) s += iabs32((int32_t)t[(size_t)c * C + l]);
    return (int32_t)s;
}

int lo_walk(const int32_t *A, int R, int C, int n, int marginal,
            uint64_t jmin, uint64_t jmax,
            int32_t *best_out, uint64_t *best_rank, int32_t *strategy)
{
    if (R < 1 || C < 1 || n < 1 || n > LO_MAX_LABELS || jmax < jmin) return -1;
    const int nv = R - 1;
    const size_t Cs = (size_t)C;
    int32_t vect[64];
    int32_t best;
    uint64_t brank = jmin;

    if (n <= 2) {
        /* binary reflected Gray code: L1/LM (L_OpenMP.c:15-85), L2 (87-124) */
        uint32_t *t = (uint32_t *)calloc(2 * Cs, sizeof(uint32_t));
        if (!t) return -1;
        uint32_t *t0 = t, *t1 = t + Cs;  /* L1: t0 is y; L2: t0,t1 */
        for (int l = 0; l < C; ++l) t0[l] = (uint32_t)A[(size_t)nv * Cs + l];
        for (int i = 0; i < nv; ++i) {
            vect[i] = lo_digit(jmin, i, 2);
            const int32_t *row = A + (size_t)i * Cs;
            for (int l = 0; l < C; ++l) {
                if (n == 1) t0[l] = vect[i] ? t0[l] + (uint32_t)row[l] : t0[l] - (uint32_t)row[l];
                else if (vect[i]) t1[l] += (uint32_t)row[l];
                else t0[l] += (uint32_t)row[l];
            }
        }
        best = n == 1 ? sum_abs_l1(t0, C, marginal) : sum_abs_labels(t, 2, C);
        if (strategy) memcpy(strategy, vect, sizeof(int32_t) * (size_t)nv);
        for (uint64_t N = jmin + 1; N <= jmax && N != 0; ++N) {
            int i = __builtin_ctzll(N);
            const uint32_t *row = (const uint32_t *)A + (size_t)i * Cs;
            int up = vect[i] == 0;
            vect[i] = up;
            int64_t prod = 0;
            if (n == 1) {
                /* L_OpenMP.c:77-78 (LM: 40-41, column 0 without abs) */
                uint32_t m = up ? 2u : (uint32_t)-2;
                int l0 = 0;
                if (marginal) { t0[0] += m * row[0]; prod += (int32_t)t0[0]; l0 = 1; }
                for (int l = l0; l < C; ++l) {
                    t0[l] += m * row[l];
                    prod += iabs32((int32_t)t0[l]);
                }
            } else {
                /* L_OpenMP.c:116-117 */
                uint32_t m = up ? 1u : (uint32_t)-1;
                for (int l = 0; l < C; ++l) {
                    uint32_t a = m * row[l];
                    t1[l] += a;
                    t0[l] -= a;
                    prod += iabs32((int32_t)t0[l]) + iabs32((int32_t)t1[l]);
                }
            }
            if ((int32_t)prod > best) {
                best = (int32_t)prod;
                brank = N;
                if (strategy) memcpy(strategy, vect, sizeof(int32_t) * (size_t)nv);
            }
            if (N == jmax) break;
        }
        free(t);
    } else {
        /* reflected n-ary Gray code: L3 (L_OpenMP.c:126-183), Ln (185-239) */
        uint64_t pw[64];
        pw[0] = 1;
        for (int i = 1; i < 64; ++i) pw[i] = pw[i - 1] * (uint64_t)n;
        uint32_t *t = (uint32_t *)calloc((size_t)n * Cs, sizeof(uint32_t));
        if (!t) return -1;
        for (int l = 0; l < C; ++l) t[l] = (uint32_t)A[(size_t)nv * Cs + l];
        for (int i = 0; i < nv; ++i) {
            vect[i] = lo_digit(jmin, i, n);
            const int32_t *row = A + (size_t)i * Cs;
            for (int l = 0; l < C; ++l) t[(size_t)vect[i] * Cs + l] += (uint32_t)row[l];
        }
        best = sum_abs_labels(t, n, C);
        if (strategy) memcpy(strategy, vect, sizeof(int32_t) * (size_t)nv);
        for (uint64_t N = jmin + 1; N <= jmax && N != 0; ++N) {
            for (int i = 0; i < nv; ++i) {
                uint64_t q = N / pw[i];
                if (q % (uint64_t)n) {
                    uint64_t r = q % (uint64_t)(2 * n);
                    int nw = r < (uint64_t)n ? (int)r : (int)(2 * n - 1 - r);
                    const int32_t *row = A + (size_t)i * Cs;
                    uint32_t *from = t + (size_t)vect[i] * Cs, *to = t + (size_t)nw * Cs;
                    for (int l = 0; l < C; ++l) { from[l] -= (uint32_t)row[l]; to[l] += (uint32_t)row[l]; }
                    vect[i] = nw;
                    break;
                }
            }
            int32_t prod = sum_abs_labels(t, n, C);
            if (prod > best) {
                best = prod;
                brank = N;
                if (strategy) memcpy(strategy, vect, sizeof(int32_t) * (size_t)nv);
            }
            if (N == jmax) break;
        }
        free(t);
    }
    *best_out = best;
    if (best_rank) *best_rank = brank;
    return 0;
}

int lo_calc_lnorm(const int32_t *A, int R, int C, int n, int marginal, int threads,
                  int32_t *norm, int32_t *strategy, uint64_t *rank_out)
{
    uint64_t total = lo_total(R, n);
    if (total == 0 || threads < 1) return -1;
    uint64_t copy = (uint64_t)threads > total ? total : (uint64_t)threads;  /* L_OpenMP.c:368 */
    uint64_t steps = total / copy, rem = total % copy;                       /* :369-370 */
    int nv = R - 1 > 0 ? R - 1 : 0;
    int32_t *vals = (int32_t *)malloc(sizeof(int32_t) * copy);
    uint64_t *ranks = (uint64_t *)malloc(sizeof(uint64_t) * copy);
    int32_t *strat = (int32_t *)calloc(copy * (size_t)(nv ? nv : 1), sizeof(int32_t));
    int err = 0;
    if (!vals || !ranks || !strat) { free(vals); free(ranks); free(strat); return -1; }
#pragma omp parallel for schedule(static, 1) num_threads((int)copy) reduction(| : err)
    for (long long w = 0; w < (long long)copy; ++w) {
        uint64_t jmin, jmax;
        lo_segment((int)w, steps, rem, &jmin, &jmax);
        err |= lo_walk(A, R, C, n, marginal, jmin, jmax, &vals[w], &ranks[w],
                       strat + (size_t)w * (size_t)nv);
    }
    if (!err) {
        /* L_OpenMP.c:291-294: first maximum wins */
        uint64_t im = 0;
        for (uint64_t w = 1; w < copy; ++w)
            if (vals[im] < vals[w]) im = w;
        *norm = vals[im];
        if (strategy && nv) memcpy(strategy, strat + im * (size_t)nv, sizeof(int32_t) * (size_t)nv);
        if (rank_out) *rank_out = ranks[im];
    }
    free(vals); free(ranks); free(strat);
    return err ? -1 : 0;
}

int lo_brute(const int32_t *A, int R, int C, int n, int marginal, int32_t *norm, uint64_t *rank_out)
{
    uint64_t total = lo_total(R, n);
    if (total == 0 || n > LO_MAX_LABELS) return -1;
    int nv = R - 1;
    int32_t labels[64];
    int32_t best = 0;
    uint64_t brank = 0;
    for (uint64_t N = 0; N < total; ++N) {
        for (int i = 0; i < nv; ++i) {
            if (n <= 2) {
                labels[i] = (int32_t)(((N >> i) ^ (N >> (i + 1))) & 1);
            } else {
                /* explicit base-n expansion with reflection of odd blocks */
                uint64_t q = N;
                for (int t = 0; t < i; ++t) q /= (uint64_t)n;
                uint64_t r = q % (uint64_t)n;
                uint64_t blk = q / (uint64_t)n;
                labels[i] = (int32_t)((blk & 1) ? (uint64_t)(n - 1) - r : r);
            }
        }
        int32_t v = lo_objective(A, R, C, n, marginal, labels);
        if (N == 0 || v > best) { best = v; brank = N; }
    }
    *norm = best;
    if (rank_out) *rank_out = brank;
    return 0;
}

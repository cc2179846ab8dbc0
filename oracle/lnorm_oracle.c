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
 * Integer arithmetic: the reference computes in int_type (functions.h:5),
 * int by default and long long in its documented variant (README.md:100-105,
 * with abs -> llabs).  The walks here run in 64-bit two's complement
 * (uint64 updates, so 2*A[i] never invokes signed-overflow UB): that is the
 * long long variant exactly whenever sum|A| <= INT64_MAX, and the int
 * version exactly whenever sum|A| <= INT32_MAX (every partial sum and
 * objective then fits in int32).  The int32 entry points widen the matrix and
 * narrow the results; the *64 entry points are the long long variant.
 *
 * Extra pins of the 64-bit arithmetic (tests/test_oracle.py): the scaling
 * identity L(c*A) = |c| L(A) with |c| * L(A) > 2^31, the rank-one closed form
 * with entries beyond int32, and lo_brute64 against lo_calc_lnorm64 on
 * matrices with entries up to 2^40.
 */
#include "lnorm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LO_MAX_LABELS 10   /* RANK_OF_NORM, functions.h:3 */

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

static inline int64_t iabs64(int64_t v) { return v < 0 ? -v : v; }

int64_t lo_objective64(const int64_t *A, int R, int C, int n, int marginal, const int32_t *labels)
{
    int64_t total = 0;
    if (n == 1) {
        for (int l = 0; l < C; ++l) {
            uint64_t y = (uint64_t)A[(size_t)(R - 1) * C + l];
            for (int i = 0; i < R - 1; ++i) {
                uint64_t a = (uint64_t)A[(size_t)i * C + l];
                y = labels[i] ? y + a : y - a;
            }
            int64_t v = (int64_t)y;
            total += (marginal && l == 0) ? v : iabs64(v);
        }
        return total;
    }
    for (int c = 0; c < n; ++c) {
        for (int l = 0; l < C; ++l) {
            uint64_t t = c == 0 ? (uint64_t)A[(size_t)(R - 1) * C + l] : 0u;
            for (int i = 0; i < R - 1; ++i)
                if (labels[i] == c) t += (uint64_t)A[(size_t)i * C + l];
            total += iabs64((int64_t)t);
        }
    }
    return total;
}

/* ---- the Gray walks (L_OpenMP.c:15-239) ---------------------------------- */

static int64_t sum_abs_l1(const uint64_t *y, int C, int marginal)
{
    int64_t s = 0;
    for (int l = 0; l < C; ++l) {
        int64_t v = (int64_t)y[l];
        s += (marginal && l == 0) ? v : iabs64(v);
    }
    return s;
}

static int64_t sum_abs_labels(const uint64_t *t, int n, int C)
{
    int64_t s = 0;
    for (int c = 0; c < n; ++c)
        for (int l = 0; l < C; ++l) s += iabs64((int64_t)t[(size_t)c * C + l]);
    return s;
}

int lo_walk64(const int64_t *A, int R, int C, int n, int marginal,
              uint64_t jmin, uint64_t jmax,
              int64_t *best_out, uint64_t *best_rank, int32_t *strategy)
{
    if (R < 1 || C < 1 || n < 1 || n > LO_MAX_LABELS || jmax < jmin) return -1;
    const int nv = R - 1;
    const size_t Cs = (size_t)C;
    int32_t vect[64];
    int64_t best;
    uint64_t brank = jmin;

    if (n <= 2) {
        /* binary reflected Gray code: L1/LM (L_OpenMP.c:15-85), L2 (87-124) */
        uint64_t *t = (uint64_t *)calloc(2 * Cs, sizeof(uint64_t));
        if (!t) return -1;
        uint64_t *t0 = t, *t1 = t + Cs;  /* L1: t0 is y; L2: t0,t1 */
        for (int l = 0; l < C; ++l) t0[l] = (uint64_t)A[(size_t)nv * Cs + l];
        for (int i = 0; i < nv; ++i) {
            vect[i] = lo_digit(jmin, i, 2);
            const int64_t *row = A + (size_t)i * Cs;
            for (int l = 0; l < C; ++l) {
                if (n == 1) t0[l] = vect[i] ? t0[l] + (uint64_t)row[l] : t0[l] - (uint64_t)row[l];
                else if (vect[i]) t1[l] += (uint64_t)row[l];
                else t0[l] += (uint64_t)row[l];
            }
        }
        best = n == 1 ? sum_abs_l1(t0, C, marginal) : sum_abs_labels(t, 2, C);
        if (strategy) memcpy(strategy, vect, sizeof(int32_t) * (size_t)nv);
        for (uint64_t N = jmin + 1; N <= jmax && N != 0; ++N) {
            int i = __builtin_ctzll(N);
            const uint64_t *row = (const uint64_t *)A + (size_t)i * Cs;
            int up = vect[i] == 0;
            vect[i] = up;
            int64_t prod = 0;
            if (n == 1) {
                /* L_OpenMP.c:77-78 (LM: 40-41, column 0 without abs) */
                uint64_t m = up ? 2u : (uint64_t)-2;
                int l0 = 0;
                if (marginal) { t0[0] += m * row[0]; prod += (int64_t)t0[0]; l0 = 1; }
                for (int l = l0; l < C; ++l) {
                    t0[l] += m * row[l];
                    prod += iabs64((int64_t)t0[l]);
                }
            } else {
                /* L_OpenMP.c:116-117 */
                uint64_t m = up ? 1u : (uint64_t)-1;
                for (int l = 0; l < C; ++l) {
                    uint64_t a = m * row[l];
                    t1[l] += a;
                    t0[l] -= a;
                    prod += iabs64((int64_t)t0[l]) + iabs64((int64_t)t1[l]);
                }
            }
            if (prod > best) {
                best = prod;
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
        uint64_t *t = (uint64_t *)calloc((size_t)n * Cs, sizeof(uint64_t));
        if (!t) return -1;
        for (int l = 0; l < C; ++l) t[l] = (uint64_t)A[(size_t)nv * Cs + l];
        for (int i = 0; i < nv; ++i) {
            vect[i] = lo_digit(jmin, i, n);
            const int64_t *row = A + (size_t)i * Cs;
            for (int l = 0; l < C; ++l) t[(size_t)vect[i] * Cs + l] += (uint64_t)row[l];
        }
        best = sum_abs_labels(t, n, C);
        if (strategy) memcpy(strategy, vect, sizeof(int32_t) * (size_t)nv);
        for (uint64_t N = jmin + 1; N <= jmax && N != 0; ++N) {
            for (int i = 0; i < nv; ++i) {
                uint64_t q = N / pw[i];
                if (q % (uint64_t)n) {
                    uint64_t r = q % (uint64_t)(2 * n);
                    int nw = r < (uint64_t)n ? (int)r : (int)(2 * n - 1 - r);
                    const int64_t *row = A + (size_t)i * Cs;
                    uint64_t *from = t + (size_t)vect[i] * Cs, *to = t + (size_t)nw * Cs;
                    for (int l = 0; l < C; ++l) { from[l] -= (uint64_t)row[l]; to[l] += (uint64_t)row[l]; }
                    vect[i] = nw;
                    break;
                }
            }
            int64_t prod = sum_abs_labels(t, n, C);
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

int lo_calc_lnorm64(const int64_t *A, int R, int C, int n, int marginal, int threads,
                    int64_t *norm, int32_t *strategy, uint64_t *rank_out)
{
    uint64_t total = lo_total(R, n);
    if (total == 0 || threads < 1) return -1;
    uint64_t copy = (uint64_t)threads > total ? total : (uint64_t)threads;  /* L_OpenMP.c:368 */
    uint64_t steps = total / copy, rem = total % copy;                       /* :369-370 */
    int nv = R - 1 > 0 ? R - 1 : 0;
    int64_t *vals = (int64_t *)malloc(sizeof(int64_t) * copy);
    uint64_t *ranks = (uint64_t *)malloc(sizeof(uint64_t) * copy);
    int32_t *strat = (int32_t *)calloc(copy * (size_t)(nv ? nv : 1), sizeof(int32_t));
    int err = 0;
    if (!vals || !ranks || !strat) { free(vals); free(ranks); free(strat); return -1; }
#pragma omp parallel for schedule(static, 1) num_threads((int)copy) reduction(| : err)
    for (long long w = 0; w < (long long)copy; ++w) {
        uint64_t jmin, jmax;
        lo_segment((int)w, steps, rem, &jmin, &jmax);
        err |= lo_walk64(A, R, C, n, marginal, jmin, jmax, &vals[w], &ranks[w],
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

int lo_brute64(const int64_t *A, int R, int C, int n, int marginal, int64_t *norm, uint64_t *rank_out)
{
    uint64_t total = lo_total(R, n);
    if (total == 0 || n > LO_MAX_LABELS) return -1;
    int nv = R - 1;
    int32_t labels[64];
    int64_t best = 0;
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
        int64_t v = lo_objective64(A, R, C, n, marginal, labels);
        if (N == 0 || v > best) { best = v; brank = N; }
    }
    *norm = best;
    if (rank_out) *rank_out = brank;
    return 0;
}

/* ---- int (the reference default): widen, walk in 64 bits, narrow -------- */

static int64_t *widen(const int32_t *A, int R, int C)
{
    if (R < 1 || C < 1) return NULL;
    int64_t *w = (int64_t *)malloc(sizeof(int64_t) * (size_t)R * (size_t)C);
    if (!w) return NULL;
    for (size_t i = 0; i < (size_t)R * (size_t)C; ++i) w[i] = A[i];
    return w;
}

int32_t lo_objective(const int32_t *A, int R, int C, int n, int marginal, const int32_t *labels)
{
    int64_t *w = widen(A, R, C);
    if (!w) return 0;
    int64_t v = lo_objective64(w, R, C, n, marginal, labels);
    free(w);
    return (int32_t)v;
}

int lo_walk(const int32_t *A, int R, int C, int n, int marginal,
            uint64_t jmin, uint64_t jmax,
            int32_t *best_out, uint64_t *best_rank, int32_t *strategy)
{
    int64_t *w = widen(A, R, C);
    if (!w) return -1;
    int64_t b = 0;
    int r = lo_walk64(w, R, C, n, marginal, jmin, jmax, &b, best_rank, strategy);
    free(w);
    if (!r) *best_out = (int32_t)b;
    return r;
}

int lo_calc_lnorm(const int32_t *A, int R, int C, int n, int marginal, int threads,
                  int32_t *norm, int32_t *strategy, uint64_t *rank_out)
{
    int64_t *w = widen(A, R, C);
    if (!w) return -1;
    int64_t b = 0;
    int r = lo_calc_lnorm64(w, R, C, n, marginal, threads, &b, strategy, rank_out);
    free(w);
    if (!r) *norm = (int32_t)b;
    return r;
}

int lo_brute(const int32_t *A, int R, int C, int n, int marginal, int32_t *norm, uint64_t *rank_out)
{
    int64_t *w = widen(A, R, C);
    if (!w) return -1;
    int64_t b = 0;
    int r = lo_brute64(w, R, C, n, marginal, &b, rank_out);
    free(w);
    if (!r) *norm = (int32_t)b;
    return r;
}

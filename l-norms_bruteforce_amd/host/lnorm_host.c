/*
 * lnorm_host.c -- host side of the drop-in (include/lnorm_host.h).
 *
 * Behaviour follows the reference host code:
 *   reader            functions.h:40-79
 *   zero lines        functions.h:109-151
 *   row / col merges  functions.h:237-289, single-sign columns 193-235
 *   compaction        functions.h:153-191 (a stable filter of the kept lines)
 *   layout            functions.h:291-329
 *   work sizing       L_OpenMP.c:344-376, L_CUDA.cu:242-279
 *   printer           functions.h:371-412
 *   argv              L_OpenMP.c:299-342, L_CUDA.cu:361-412
 * Arithmetic on entries is int32 with wrap-around (the reference's int on
 * x86); proportionality tests use 64-bit products as the reference does.
 * Rows shorter than the last row (undefined behaviour in the reference) are
 * padded with zeros.
 */
#define _GNU_SOURCE
#include "lnorm_host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NUM_OF_BITS 64

static int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

static int is_token_char(int g) {
    return (g >= '0' && g <= '9') || g == 'e' || g == 'E' || g == '.' || g == '+' || g == '-';
}

/* x86 (int) of a double: truncation, INT_MIN when out of range or NaN */
static int32_t dtoi(double d) {
    if (!(d > -2147483649.0 && d < 2147483648.0)) return INT32_MIN;
    return (int32_t)d;
}

/* ---- reader ---------------------------------------------------------------- */
typedef struct {
    const unsigned char *text;
    size_t len, pos;
    FILE *fp;
} src_t;

static int src_get(src_t *s) {
    int c;
    if (s->fp) c = fgetc(s->fp);
    else c = s->pos < s->len ? s->text[s->pos++] : EOF;
    /* `char g = fgetc(fp)` (functions.h:52): byte 0xFF compares equal to EOF */
    return (c == EOF || c == 0xFF) ? EOF : c;
}

static int read_matrix(lnorm_host_problem *p, src_t *s, FILE *out) {
    int32_t **rows = NULL;
    int *lens = NULL;
    int nrows = 0, cap_rows = 0;
    int32_t *row = NULL;
    int j = 0, cap_row = 0;
    char *tok = NULL;
    size_t ntok = 0, cap_tok = 0;
    int has_e = 0;
    int value = 0;
    double dvalue = 0.0;
    int icols = 0;
    int g;
    do {
        g = src_get(s);
        if (g != EOF && is_token_char(g)) {
            if (ntok + 2 > cap_tok) {
                cap_tok = cap_tok ? 2 * cap_tok : 64;
                char *t = (char *)realloc(tok, cap_tok);
                if (!t) goto oom;
                tok = t;
            }
            tok[ntok++] = (char)g;
            if (g == 'e' || g == 'E') has_e = 1;
        } else {
            if (ntok) {
                tok[ntok] = '\0';
                if (has_e) {
                    sscanf(tok, "%le", &dvalue);
                    value = dtoi(dvalue);
                    has_e = 0;
                } else {
                    sscanf(tok, "%d", &value);
                }
                ntok = 0;
                if (j + 1 > cap_row) {
                    cap_row = cap_row ? 2 * cap_row : 16;
                    int32_t *r = (int32_t *)realloc(row, sizeof(int32_t) * (size_t)cap_row);
                    if (!r) goto oom;
                    row = r;
                }
                row[j++] = value;
            }
            if ((g == '\n' || g == EOF) && j > 0) {
                icols = j;
                if (nrows + 1 > cap_rows) {
                    cap_rows = cap_rows ? 2 * cap_rows : 16;
                    int32_t **rr = (int32_t **)realloc(rows, sizeof(int32_t *) * (size_t)cap_rows);
                    int *ll = (int *)realloc(lens, sizeof(int) * (size_t)cap_rows);
                    if (rr) rows = rr;
                    if (ll) lens = ll;
                    if (!rr || !ll) goto oom;
                }
                rows[nrows] = row;
                lens[nrows] = j;
                ++nrows;
                row = NULL;
                j = 0;
                cap_row = 0;
            }
        }
    } while (g != EOF);
    free(row);
    free(tok);
    p->iRows = nrows;
    p->iCols = icols;
    p->matrix = (int32_t *)calloc((size_t)(nrows ? nrows : 1) * (size_t)(icols ? icols : 1), sizeof(int32_t));
    if (!p->matrix) goto oom_rows;
    for (int i = 0; i < nrows; ++i) {
        const int n = lens[i] < icols ? lens[i] : icols;
        memcpy(p->matrix + (size_t)i * icols, rows[i], sizeof(int32_t) * (size_t)n);
        free(rows[i]);
    }
    free(rows);
    free(lens);
    fprintf(out, "rows: %d, cols: %d\n", p->iRows, p->iCols);
    return 0;
oom:
    free(row);
    free(tok);
oom_rows:
    for (int i = 0; i < nrows; ++i) free(rows[i]);
    free(rows);
    free(lens);
    return -1;
}

#define M(p, i, k) ((p)->matrix[(size_t)(i) * (size_t)(p)->iCols + (size_t)(k)])

static int flag_in(char stat, const char *set) { return stat != '\0' && strchr(set, stat) != NULL; }

/* ---- preprocessing ------------------------------------------------------------ */
static void eliminate_zero_lines(lnorm_host_problem *p, FILE *out) {
    p->R = p->iRows;
    p->C = p->iCols;
    p->marginal = p->marginal_req;
    if (!flag_in(p->stat, "yYrRcC")) return;
    int cr = 0, cc = 0;
    for (int i = 0; i < p->iRows; ++i) {
        int zero = 1;
        for (int k = 0; k < p->iCols && zero; ++k) zero = M(p, i, k) == 0;
        if (zero) {
            p->original_r[i] = p->iRows;
            if (p->marginal_req == 0 || i > 0) {
                ++cr;
                fprintf(out, "row %d entries are zeros.\n", i + 1);
            }
        }
    }
    p->R = p->iRows - cr;
    for (int i = 0; i < p->iCols; ++i) {
        int zero = 1;
        for (int k = 0; k < p->iRows && zero; ++k) zero = M(p, k, i) == 0;
        if (zero) {
            p->original_c[i] = p->iCols;
            if (p->marginal_req == 0 || i > 0) {
                ++cc;
                fprintf(out, "col %d entries are zeros.\n", i + 1);
            }
        }
    }
    p->C = p->iCols - cc;
    if (p->marginal_req == 1) {
        if (p->original_r[0] == p->iRows && p->original_c[0] == p->iCols) {
            p->marginal = 0;
            p->R--;
            p->C--;
            fprintf(out, "The entries of the first row and column are zeros.\n");
        } else {
            p->original_r[0] = 0;
            p->original_c[0] = 0;
        }
    }
}

/* Proportionality of two lines u, v of length L read through get(i, k):
 * the first position where not both vanish anchors the ratio, which must hold
 * for every later position (64-bit cross products).  Returns mult and the
 * anchor index. */
typedef int32_t (*getter)(const lnorm_host_problem *, int line, int k);
static int32_t get_row(const lnorm_host_problem *p, int line, int k) { return M(p, line, k); }
static int32_t get_col(const lnorm_host_problem *p, int line, int k) { return M(p, k, line); }

static int proportional(const lnorm_host_problem *p, getter get, int u, int v, int L, int *anchor) {
    int k = 0;
    while (k < L && get(p, u, k) == 0 && get(p, v, k) == 0) ++k;
    if (k >= L) { *anchor = k; return 0; }
    int mult = ((get(p, u, k) == 0) != (get(p, v, k) == 0)) ? 0 : 1;
    const int idx = k;
    for (; k < L - 1 && mult; ++k)
        mult = (int64_t)get(p, u, idx) * get(p, v, k + 1) == (int64_t)get(p, v, idx) * get(p, u, k + 1);
    *anchor = idx;
    return mult;
}

static void merge_row(lnorm_host_problem *p, int i, int j, int sign) {
    for (int k = 0; k < p->iCols; ++k) M(p, i, k) = sign > 0 ? wadd(M(p, i, k), M(p, j, k)) : wsub(M(p, i, k), M(p, j, k));
}

static void merge_col(lnorm_host_problem *p, int i, int j, int sign) {
    for (int k = 0; k < p->iRows; ++k) M(p, k, i) = sign > 0 ? wadd(M(p, k, i), M(p, k, j)) : wsub(M(p, k, i), M(p, k, j));
}

static void reduce_rows(lnorm_host_problem *p, FILE *out) {
    int cr = 0;
    for (int i = p->marginal_req; i < p->iRows; ++i) {
        if (p->original_r[i] != 0) continue;
        for (int j = i + 1; j < p->iRows; ++j) {
            if (p->original_r[j] != 0) continue;
            int idx;
            const int mult = proportional(p, get_row, i, j, p->iCols, &idx);
            if (!mult) {
                p->original_r[j] = 0;
            } else if ((int64_t)M(p, i, idx) * M(p, j, idx) < 0) {
                if (p->n_original == 1) {
                    p->original_r[j] = -(i + 1);
                    ++cr;
                    merge_row(p, i, j, -1);
                    fprintf(out, "row %d was subtracted from %d\n", j + 1, i + 1);
                }
            } else {
                p->original_r[j] = i + 1;
                ++cr;
                merge_row(p, i, j, +1);
                fprintf(out, "row %d was added to %d\n", j + 1, i + 1);
            }
        }
    }
    p->R -= cr;
}

static void reduce_cols(lnorm_host_problem *p, FILE *out) {
    int cc = 0;
    for (int i = p->marginal_req; i < p->iCols; ++i) {
        if (p->original_c[i] != 0) continue;
        for (int j = i + 1; j < p->iCols; ++j) {
            if (p->original_c[j] != 0) continue;
            int idx;
            const int mult = proportional(p, get_col, i, j, p->iRows, &idx);
            if (!mult) {
                p->original_c[j] = 0;
            } else if ((int64_t)M(p, idx, i) * M(p, idx, j) < 0) {
                p->original_c[j] = -(i + 1);
                ++cc;
                merge_col(p, i, j, -1);
                fprintf(out, "col %d was subtracted from %d\n", j + 1, i + 1);
            } else {
                p->original_c[j] = i + 1;
                ++cc;
                merge_col(p, i, j, +1);
                fprintf(out, "col %d was added to %d\n", j + 1, i + 1);
            }
        }
    }
    p->C -= cc;
}

static int reduce_cols_sign(lnorm_host_problem *p, FILE *out) {
    int *sign = (int *)calloc((size_t)(p->iCols ? p->iCols : 1), sizeof(int));
    if (!sign) return -1;
    int cnt = 0;
    for (int i = 0; i < p->iCols; ++i) {
        if (p->original_c[i] != 0) continue;
        int pos = 0, neg = 0;
        for (int k = 0; k < p->iRows && (!pos || !neg); ++k) {
            if (M(p, k, i) > 0) pos = 1;
            if (M(p, k, i) < 0) neg = 1;
        }
        if (pos && neg) sign[i] = 0;
        else if (neg) { sign[i] = -1; ++cnt; }
        else { sign[i] = 1; ++cnt; }
    }
    if (cnt > 1) {
        p->C = p->C - cnt + 1;
        int index = 0;
        while (sign[index] == 0) ++index;
        for (int i = index + 1; i < p->iCols; ++i) {
            if (p->original_c[i] != 0) continue;
            const int mult = sign[index] * sign[i];
            if (mult < 0) {
                p->original_c[i] = -(index + 1);
                merge_col(p, index, i, -1);
                fprintf(out, "col %d was subtracted from %d\n", i + 1, index + 1);
            }
            if (mult > 0) {
                p->original_c[i] = index + 1;
                merge_col(p, index, i, +1);
                fprintf(out, "col %d was added to %d\n", i + 1, index + 1);
            }
        }
    }
    free(sign);
    return 0;
}

/* ---- layout (functions.h:291-329) ----------------------------------------------- */
static int layout(lnorm_host_problem *p, FILE *out) {
    const int pre = flag_in(p->stat, "yYrRcC");
    int *keep_r = (int *)malloc(sizeof(int) * (size_t)(p->iRows + 1));
    int *keep_c = (int *)malloc(sizeof(int) * (size_t)(p->iCols + 1));
    if (!keep_r || !keep_c) { free(keep_r); free(keep_c); return -1; }
    int nr = 0, nc = 0;
    for (int i = 0; i < p->iRows && nr < p->R; ++i)
        if (!pre || p->original_r[i] == 0) keep_r[nr++] = i;
    for (int j = 0; j < p->iCols && nc < p->C; ++j)
        if (!pre || p->original_c[j] == 0) keep_c[nc++] = j;
    const int R = p->R, C = p->C;
    p->mtx_as_vec = (int32_t *)calloc((size_t)(R > 0 ? R : 1) * (size_t)(C > 0 ? C : 1), sizeof(int32_t));
    if (!p->mtx_as_vec) { free(keep_r); free(keep_c); return -1; }
    if (p->n_original == 1 && R > C) {
        for (int j = 0; j < C; ++j)
            for (int i = 0; i < R; ++i) p->mtx_as_vec[(size_t)j * R + i] = M(p, keep_r[i], keep_c[j]);
        p->R = C;
        p->C = R;
        p->original = p->original_c;
        p->original_length = p->iCols;
        p->transposed = 1;
        fprintf(out, "Matrix was transposed.\n");
    } else {
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < C; ++j) p->mtx_as_vec[(size_t)i * C + j] = M(p, keep_r[i], keep_c[j]);
        p->original = p->original_r;
        p->original_length = p->iRows;
    }
    if (p->marginal == 1 && p->R > 0) {
        int32_t *first = p->mtx_as_vec, *last = p->mtx_as_vec + (size_t)(p->R - 1) * p->C;
        for (int l = 0; l < p->C; ++l) {
            const int32_t t = first[l];
            first[l] = last[l];
            last[l] = t;
        }
    }
    free(keep_r);
    free(keep_c);
    return 0;
}

static int load_src(lnorm_host_problem *p, src_t *s, int n_original, char stat, int marginal, FILE *out) {
    memset(p, 0, sizeof *p);
    p->n_original = n_original;
    p->stat = stat;
    p->marginal_req = marginal;
    if (read_matrix(p, s, out)) return -1;
    p->original_r = (int *)calloc((size_t)(p->iRows ? p->iRows : 1), sizeof(int));
    p->original_c = (int *)calloc((size_t)(p->iCols ? p->iCols : 1), sizeof(int));
    if (!p->original_r || !p->original_c) return -1;
    eliminate_zero_lines(p, out);
    if (flag_in(stat, "yYrR")) reduce_rows(p, out);
    if (flag_in(stat, "yYcC")) {
        reduce_cols(p, out);
        if (n_original > 1 && reduce_cols_sign(p, out)) return -1;
    }
    return layout(p, out);
}

int lnorm_host_load(lnorm_host_problem *p, const char *path, int n_original, char stat, int marginal, FILE *out) {
    FILE *fp = fopen(path, "r");
    if (!fp) {
        memset(p, 0, sizeof *p);
        return -1;
    }
    src_t s = {NULL, 0, 0, fp};
    const int r = load_src(p, &s, n_original, stat, marginal, out);
    fclose(fp);
    return r;
}

int lnorm_host_load_text(lnorm_host_problem *p, const char *text, size_t len, int n_original, char stat,
                         int marginal, FILE *out) {
    src_t s = {(const unsigned char *)text, len, 0, NULL};
    return load_src(p, &s, n_original, stat, marginal, out);
}

/* ---- work sizing ------------------------------------------------------------------ */
static uint64_t dtou64(double d) {
    /* (unsigned long long) of a double as gcc emits it on x86-64 */
    if (d < 9223372036854775808.0) return (uint64_t)(int64_t)d;
    d -= 9223372036854775808.0;
    if (d >= 9223372036854775808.0) return 0;
    return (uint64_t)(int64_t)d ^ (1ull << 63);
}

int lnorm_host_params(lnorm_host_problem *p, int flavor, int *threads, int *blocks, FILE *out) {
    if (p->n_original < p->R || p->n_original == 1) {
        p->n = p->n_original;
    } else {
        p->n = p->R > 1 ? p->R : 2;
        fprintf(out, "The preprocessed matrix has a number of rows (%d) less than or equal to the order of the L norm (%d).\n",
                p->R, p->n_original);
    }
    if (p->n <= 2) {
        p->total = p->R == 0 ? 0 : (uint64_t)1 << (p->R - 1);
        p->maxRows = NUM_OF_BITS - 1;
    } else {
        p->total = p->R == 0 ? 0 : dtou64(pow((double)p->n, (double)(p->R - 1)));
        p->maxRows = (int)(floor(NUM_OF_BITS / log2((double)p->n)) + 1);
    }
    if (flavor == LNORM_HOST_OMP) {
        const uint64_t t = (uint64_t)(*threads > 0 ? *threads : 1);
        p->copyNum = t > p->total ? p->total : t;
        fprintf(out, "number of threads: %llu\n", (unsigned long long)p->copyNum);
    } else {
        int t = *threads, b = *blocks;
        if ((uint64_t)t > p->total) t = (int)p->total;
        if (t == 0) b = 0;
        else b = p->total / (uint64_t)t > (uint64_t)b ? b : (int)(p->total / (uint64_t)t);
        *threads = t;
        *blocks = b;
        p->copyNum = (uint64_t)b * (uint64_t)t;
        fprintf(out, "num_ofBlock: %d, num_ofThread in block: %d\n", b, t);
    }
    p->steps = p->copyNum ? p->total / p->copyNum : 0;
    p->steps_remainder = p->copyNum ? p->total % p->copyNum : 0;
    fprintf(out, "maximum length of strategy vector: %d\n", p->maxRows);
    if (p->maxRows < p->R) {
        fprintf(out, "Matrix is too large. The matrix has %d rows after reduction. The maximum number of rows cannot be more than %d.\n",
                p->R, p->maxRows);
        return -1;
    }
    if (p->C > LNORM_HOST_LENGTH) {
        fprintf(out, "Matrix is too large. The length variable %d should be greater or equal than %d.\n",
                LNORM_HOST_LENGTH, p->C);
        return -1;
    }
    return 0;
}

/* ---- printer (functions.h:371-412) ------------------------------------------------ */
const char *lnorm_host_strategy_name(const lnorm_host_problem *p) {
    static __thread char name[64];
    if (p->n_original == 1 && p->marginal_req == 1) snprintf(name, sizeof name, "strategy_LM.txt");
    else snprintf(name, sizeof name, "strategy_L%d.txt", p->n_original);
    return name;
}

int lnorm_host_print_results(const lnorm_host_problem *p, int32_t Lnorm, const int32_t *strategy,
                             const char *dir, FILE *out) {
    if (p->n_original == 1 && p->marginal_req == 1) fprintf(out, "LM is: %d\n", Lnorm);
    else fprintf(out, "L%d is: %d\n", p->n_original, Lnorm);
    char path[2048];
    const char *name = lnorm_host_strategy_name(p);
    if (dir && dir[0]) snprintf(path, sizeof path, "%s/%s", dir, name);
    else snprintf(path, sizeof path, "%s", name);
    FILE *fp = fopen(path, "w");
    if (!fp) return -1;
    int *vals = (int *)calloc((size_t)(p->original_length ? p->original_length : 1), sizeof(int));
    if (!vals) { fclose(fp); return -1; }
    int j = 0;
    for (int i = 0; i < p->original_length; ++i) {
        const int o = p->original[i];
        int val;
        if (o == 0) {
            if (p->marginal == 1 && p->n == 1) val = j == 0 ? 1 : (j == p->R - 1 ? strategy[0] : strategy[j]);
            else val = j == p->R - 1 ? (p->n == 1 ? 1 : 0) : strategy[j];
            ++j;
        } else if (o == p->original_length) {
            val = p->n == 1 ? 1 : 0;
        } else if (o < 0) {
            val = (vals[-o - 1] + 1) & 1;
        } else {
            val = vals[o - 1];
        }
        vals[i] = val;
        fprintf(fp, "%d\n", p->n == 1 ? 2 * val - 1 : val);
    }
    free(vals);
    return fclose(fp) == 0 ? 0 : -1;
}

void lnorm_host_free(lnorm_host_problem *p) {
    free(p->matrix);
    free(p->original_r);
    free(p->original_c);
    free(p->mtx_as_vec);
    memset(p, 0, sizeof *p);
}

/* ---- argv -------------------------------------------------------------------------- */
int lnorm_host_parse_args(int argc, char **argv, int flavor, int max_threads_per_block, lnorm_host_args *a,
                          FILE *out) {
    /* The usage text is the reference program's own (the drop-in prints
     * every line as the program it replaces does). */
    const char *msg = flavor == LNORM_HOST_OMP
        ? "Use the following command: ./L_OpenMP number_of_threads filename_of_matrix order_of_the_L_norm"
        : "Use the following command: ./L_CUDA number_of_blocks number_of_threads_per_block filename_of_matrix order_of_the_L_norm";
    const int off = flavor == LNORM_HOST_OMP ? 0 : 1;  /* CUDA argv has the extra block count */
    /* As in the reference, a field is refused when sscanf converts nothing
     * (sd == 0) or the value is below 1; an empty field (sd == EOF) keeps
     * the value of the field before it (0 before the first). */
    int t = 0, sd;
    memset(a, 0, sizeof *a);
    a->flavor = flavor;
    if (argc < 4 + off) {
        fprintf(out, "Incorrect number of input arguments. %s\n", msg);
        return -1;
    }
    if (flavor == LNORM_HOST_CUDA) {
        sd = sscanf(argv[1], "%d", &t);
        if (sd == 0 || t < 1) {
            fprintf(out, "Please make sure that the number of blocks is a positive integer. %s\n", msg);
            return -1;
        }
        a->blocks = t;
    }
    sd = sscanf(argv[1 + off], "%d", &t);
    if (sd == 0 || t < 1) {
        fprintf(out, "Please make sure that the number of threads is a positive integer. %s\n", msg);
        return -1;
    }
    a->threads = t;
    snprintf(a->file, sizeof a->file, "%s", argv[2 + off]);
    FILE *fp = fopen(a->file, "r");
    if (!fp) {
        /* the reference prints this line and then calls fclose(NULL), which
         * ends the process; this one exits with -1 after the same line */
        fprintf(out, "Please make sure that the file containig the matrix exists within this directory. %s\n", msg);
        return -1;
    }
    fclose(fp);
    sd = sscanf(argv[3 + off], "%d", &t);
    if (sd == 0 || t < 1) {
        fprintf(out, "Please make sure that the order of the L norm is a positive integer. %s\n", msg);
        return -1;
    }
    a->n_original = t;
    a->stat = argc < 5 + off ? 'n' : argv[4 + off][0];
    if (argc < 5 + off || a->n_original > 1) a->marginal = 0;
    else if (argc == 5 + off) a->marginal = (argv[4 + off][0] == 'm' || argv[4 + off][0] == 'M');
    else a->marginal = (argv[5 + off][0] == 'm' || argv[5 + off][0] == 'M');
    if (t > LNORM_HOST_RANK_OF_NORM) {
        fprintf(out, "The order of the L norm is too large. Please increase the RANK_OF_NORM variable in the code to %d and compile and run it again.\n", t);
        return -1;
    }
    if (flavor == LNORM_HOST_CUDA && max_threads_per_block > 0 && a->threads > max_threads_per_block) {
        fprintf(out, "The maximum number of threads per block cannot be more than %d.\n", max_threads_per_block);
        return -1;
    }
    return 0;
}

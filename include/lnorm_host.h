/*
 * lnorm_host.h -- the host side of the drop-in: matrix file reader,
 * preprocessing, layout, work sizing and result printer of the reference
 * (functions.h, L_OpenMP.c / L_CUDA.cu argument and parameter handling),
 * written anew in C.  Output text and strategy files are byte-identical to
 * the reference's.  Every printing function takes the FILE* to print to.
 */
#ifndef LNORM_HOST_H
#define LNORM_HOST_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LNORM_HOST_LENGTH 4096      /* functions.h:2 `length` */
#define LNORM_HOST_RANK_OF_NORM 10  /* functions.h:3 */

#define LNORM_HOST_OMP 0            /* L_OpenMP flavour: "number of threads: %llu" */
#define LNORM_HOST_CUDA 1           /* L_CUDA flavour: "num_ofBlock: %d, num_ofThread in block: %d" */

/* The reference's `item` + `item_calc` (functions.h:7-38) in one record. */
typedef struct lnorm_host_problem {
    /* request (arguments_OMP / arguments_CUDA) */
    int n_original;             /* d as given */
    char stat;                  /* preprocessing flag: y/Y, r/R, c/C or anything else */
    int marginal_req;           /* first->marginal */
    /* matrix as read (matrix_read, functions.h:40-79) */
    int iRows, iCols;
    int32_t *matrix;            /* iRows x iCols row-major, modified in place by preprocessing */
    int *original_r, *original_c;
    /* reduced problem (item_calc) */
    int R, C;                   /* iRows_reduced, iCols_reduced */
    int n;                      /* effective order */
    int marginal;               /* second->marginal */
    int maxRows;
    uint64_t total;             /* total_num_to_calc */
    uint64_t copyNum, steps, steps_remainder;
    int32_t *mtx_as_vec;        /* R x C row-major (convert_mtx_to_vec) */
    const int *original;        /* original_r or original_c (transposed) */
    int original_length;
    int transposed;
} lnorm_host_problem;

typedef struct lnorm_host_args {
    int flavor;                 /* LNORM_HOST_OMP or LNORM_HOST_CUDA */
    int threads;                /* OMP: number_of_threads; CUDA: number_of_threads_per_block */
    int blocks;                 /* CUDA: number_of_blocks */
    char file[1024];
    int n_original;
    char stat;
    int marginal;
} lnorm_host_args;

/* arguments_OMP (L_OpenMP.c:299-342) with argv = {prog, T, file, d, [p], [m]}
 * or arguments_CUDA (L_CUDA.cu:361-412) with argv = {prog, B, T, file, d, [p], [m]}.
 * Prints the reference's messages to `out`; returns 0, or -1 where the
 * reference exits (and where it would crash on a missing file).
 * max_threads_per_block applies to the CUDA flavour (<= 0: no check). */
int lnorm_host_parse_args(int argc, char **argv, int flavor, int max_threads_per_block,
                          lnorm_host_args *args, FILE *out);

/* load_parameters without calc_Parameters (L_OpenMP.c:378-385): reads the
 * file, preprocesses according to stat, lays the matrix out.  Returns 0, or
 * -1 if the file cannot be opened or memory runs out. */
int lnorm_host_load(lnorm_host_problem *p, const char *path, int n_original, char stat,
                    int marginal, FILE *out);

/* Same, reading the matrix text from memory (len bytes). */
int lnorm_host_load_text(lnorm_host_problem *p, const char *text, size_t len, int n_original,
                         char stat, int marginal, FILE *out);

/* calc_Parameters (L_OpenMP.c:344-376 / L_CUDA.cu:242-279).  For the CUDA
 * flavour *blocks and *threads are clamped in place.  Returns 0, or -1 where
 * the reference exits ("Matrix is too large ..."). */
int lnorm_host_params(lnorm_host_problem *p, int flavor, int *threads, int *blocks, FILE *out);

/* print_results (functions.h:371-412): the norm line to `out` and the
 * strategy file strategy_L<d>.txt / strategy_LM.txt in directory `dir`
 * (NULL or "" = current directory).  `strategy` holds R-1 labels
 * (second->strategy; unused when R == 0).  Returns 0 or -1 if the file cannot
 * be written. */
int lnorm_host_print_results(const lnorm_host_problem *p, int32_t Lnorm, const int32_t *strategy,
                             const char *dir, FILE *out);

/* Name of the strategy file print_results writes for this problem. */
const char *lnorm_host_strategy_name(const lnorm_host_problem *p);

void lnorm_host_free(lnorm_host_problem *p);

#ifdef __cplusplus
}
#endif
#endif /* LNORM_HOST_H */

/*
 * L_HIP -- command-line drop-in for the reference's L_CUDA (L_CUDA.cu:425-439)
 * running the enumeration on an MI355X through include/lnorm_hip.h.
 *
 *   L_HIP number_of_blocks number_of_threads_per_block file d [p] [m]
 *   L_HIP --omp number_of_threads file d [p] [m]
 *
 * The first form takes L_CUDA's argv and prints its "num_ofBlock" line; the
 * second takes L_OpenMP's argv and prints its "number of threads" line, so
 * the whole stdout matches that program byte for byte.  The block/thread
 * numbers only size the printed partition: the GPU launch is the engine's
 * own and the result (norm and strategy) does not depend on it.
 *
 * Exit status: 0 on success; 255 (exit(-1)) where the reference exits, and
 * also when the GPU computation fails -- the reference then prints a silent
 * zero norm (README.md:98), this program prints the error on stderr instead.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lnorm_hip.h"
#include "lnorm_host.h"

int main(int argc, char **argv) {
    int flavor = LNORM_HOST_CUDA;
    char **av = argv;
    int ac = argc;
    if (argc > 1 && strcmp(argv[1], "--omp") == 0) {
        flavor = LNORM_HOST_OMP;
        av = argv + 1;
        ac = argc - 1;
        av[0] = argv[0];
    }
    lnorm_host_args args;
    if (lnorm_host_parse_args(ac, av, flavor, 1024, &args, stdout)) exit(-1);

    lnorm_host_problem p;
    if (lnorm_host_load(&p, args.file, args.n_original, args.stat, args.marginal, stdout)) {
        fprintf(stderr, "L_HIP: cannot read %s\n", args.file);
        exit(-1);
    }
    int threads = args.threads, blocks = args.blocks;
    if (lnorm_host_params(&p, flavor, &threads, &blocks, stdout)) exit(-1);

    int32_t norm = 0;
    int32_t *strategy = (int32_t *)calloc((size_t)(p.R > 1 ? p.R - 1 : 1), sizeof(int32_t));
    if (!strategy) exit(-1);
    if (p.R == 0) {
        printf("This is a zero matrix.\n");
    } else {
        const int rc = lnorm_hip_calc_Lnorm(p.mtx_as_vec, p.R, p.C, p.n, p.marginal, blocks, threads, &norm,
                                            strategy);
        if (rc != LNORM_OK) {
            fflush(stdout);
            fprintf(stderr, "L_HIP: GPU computation failed (%d): %s\n", rc, lnorm_hip_last_error());
            exit(-1);
        }
    }
    if (lnorm_host_print_results(&p, norm, strategy, NULL, stdout)) {
        fprintf(stderr, "L_HIP: cannot write %s\n", lnorm_host_strategy_name(&p));
        exit(-1);
    }
    free(strategy);
    lnorm_host_free(&p);
    return 0;
}

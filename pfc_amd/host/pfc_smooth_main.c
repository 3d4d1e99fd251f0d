/*
 * pfc_smooth_main.c -- the smooth command of the reference's pipelines
 * (src/smooth.c:17-151): Gaussian smoothing of one data file, here on the GPU
 * through pfc_smooth (include/pfc_relax.h).
 *
 *   mpirun -np 4 pfc-smooth in.n out.eta 0.2   (only rank 0 works)
 *
 * PFC_PRECISION=mixed selects FP32 transforms.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pfc_io.h"
#include "pfc_relax.h"

int pfc_rank = 0;

static int launcher_rank(void) {
    static const char* vars[] = {"PMI_RANK", "OMPI_COMM_WORLD_RANK", "PMIX_RANK",
                                 "MV2_COMM_WORLD_RANK"};
    for (size_t i = 0; i < sizeof vars / sizeof vars[0]; i++) {
        const char* v = getenv(vars[i]);
        if (v && *v) return atoi(v);
    }
    return 0;
}

int main(int argc, char** argv) {
    pfc_rank = launcher_rank();
    if (argc == 1) {
        if (pfc_rank == 0) {
            printf("This is a tool for smoothing PFC data. The program expects the names of the input "
                   "and output data files, and the spread of the smoothing kernel in reciprocal "
                   "units.\n\n");
            printf("An example of valid syntax:\n\n");
            printf("mpirun -np 4 smooth [input data file] [output data file] [smoothing kernel "
                   "spread]\n");
        }
        return 0;
    }
    /* the reference reads argv[2] and argv[3] unchecked */
    if (argc < 4) {
        if (pfc_rank == 0) printf("Error: Invalid syntax.\n");
        return 1;
    }
    if (pfc_rank != 0) return 0;
    FILE* file = fopen(argv[1], "r");
    if (!file) {
        printf("Input file not found!\n");
        return 1;
    }
    int W = 0, H = 0;
    double dx = 0.0, dy = 0.0;
    if (fscanf(file, " %d %d %lf %lf ", &W, &H, &dx, &dy) != 4 || W < 1 || H < 1) {
        printf("Error: Invalid input data file header.\n");
        fclose(file);
        return 1;
    }
    const size_t n = (size_t)W * H;
    double* f = (double*)malloc(sizeof(double) * n);
    if (!f) return 1;
    for (size_t k = 0; k < n; k++)
        if (fscanf(file, " %lf ", &f[k]) != 1) {
            printf("Error: Invalid data in data file.\n");
            fclose(file);
            return 1;
        }
    fclose(file);
    const double sigma = atof(argv[3]);
    const char* prec = getenv("PFC_PRECISION");
    const int precision = (prec && strcmp(prec, "mixed") == 0) ? PFC_MIXED : PFC_FP64;
    const int rc = pfc_smooth(f, f, W, H, dx, dy, sigma, precision);
    if (rc) {
        fprintf(stderr, "Error: pfc_smooth failed: %s\n", pfc_strerror(rc));
        free(f);
        return 42;
    }
    if (pfc_write_field_prec(argv[2], f, W, H, dx, dy, 6, 12)) {
        fprintf(stderr, "Error: could not write %s\n", argv[2]);
        free(f);
        return 1;
    }
    free(f);
    return 0;
}

/*
 * pfc_fields_main.c -- the fields command (src/fields.c:362-424): orientation
 * field, and optionally deformation field, of one density file, computed on
 * the GPU through pfc_fields (include/pfc_relax.h).
 *
 *   mpirun -np N pfc-fields in.n out.phi fold sk sb
 *   mpirun -np N pfc-fields in.n out.phi out.chi fold sk sb power
 *
 * Only rank 0 works (rank from the MPI launcher's environment).  Files as
 * fields.c writes them: a "%d %d %lf %lf" header, then "%e %e" (re, im) per
 * point for phi and "%e" per point for chi, rows in order.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pfc_relax.h"

static int launcher_rank(void) {
    static const char* vars[] = {"PMI_RANK", "OMPI_COMM_WORLD_RANK", "PMIX_RANK",
                                 "MV2_COMM_WORLD_RANK"};
    for (size_t i = 0; i < sizeof vars / sizeof vars[0]; i++) {
        const char* v = getenv(vars[i]);
        if (v && *v) return atoi(v);
    }
    return 0;
}

static int write_file(const char* name, int W, int H, double dx, double dy, const double* v,
                      int per_line) {
    FILE* f = fopen(name, "w");
    if (!f) return 1;
    static char buf[1 << 20];
    setvbuf(f, buf, _IOFBF, sizeof buf);
    fprintf(f, "%d %d %lf %lf\n", W, H, dx, dy);
    const size_t n = (size_t)W * H;
    if (per_line == 2)
        for (size_t k = 0; k < n; k++) fprintf(f, "%e %e\n", v[2 * k], v[2 * k + 1]);
    else
        for (size_t k = 0; k < n; k++) fprintf(f, "%e\n", v[k]);
    return fclose(f) != 0;
}

int main(int argc, char** argv) {
    const int rank = launcher_rank();
    if (argc == 1) {
        if (rank == 0) {
            printf("This is a tool for computing the orientation field (and optionally the "
                   "deformation field) of a density field of a periodic system, such as a "
                   "polycrystal, on the GPU.\n\n");
            printf("Valid syntax:\n\n");
            printf("pfc-fields density.n orientation.phi [order of rotational symmetry] [spread of "
                   "complex-valued kernel] [spread of Gaussian smoothing kernel]\n\n");
            printf("pfc-fields density.n orientation.phi deformation.chi [order of rotational "
                   "symmetry] [spread of complex-valued kernel] [spread of Gaussian smoothing "
                   "kernel] [exponent for deformation field]\n\n");
            printf("The order of rotational symmetry is even (2 stripes, 4 square, 6 hexagonal or "
                   "honeycomb); the spreads are in reciprocal units; the density is expected to have "
                   "its characteristic wave number at k = 1. Grid sides: any whose transform fits one "
                   "block's LDS (prime factors up to 13: up to 4864; others up to 2048).\n");
        }
        return 0;
    }
    if (argc != 6 && argc != 8) {
        if (rank == 0) printf("Error: Invalid syntax.\n");
        return 1;
    }
    if (rank != 0) return 0;
    const int with_chi = argc == 8;
    const char* name_n = argv[1];
    const char* name_phi = argv[2];
    const char* name_chi = with_chi ? argv[3] : NULL;
    const int a0 = with_chi ? 4 : 3;
    const int fold = atoi(argv[a0]);
    const double sk = atof(argv[a0 + 1]);
    const double sb = atof(argv[a0 + 2]);
    const double power = with_chi ? atof(argv[7]) : 0.0;

    FILE* file = fopen(name_n, "r");
    if (!file) {
        printf("Error: Input file not found.\n");
        return 1;
    }
    int W = 0, H = 0;
    double dx = 0.0, dy = 0.0;
    if (fscanf(file, " %d %d %lf %lf ", &W, &H, &dx, &dy) != 4 || W < 1 || H < 1) {
        printf("Error: Invalid data.\n");
        fclose(file);
        return 1;
    }
    const size_t n = (size_t)W * H;
    double* p = (double*)calloc(n, sizeof(double));
    double* phi = (double*)malloc(2 * n * sizeof(double));
    double* chi = with_chi ? (double*)malloc(n * sizeof(double)) : NULL;
    if (!p || !phi || (with_chi && !chi)) {
        fprintf(stderr, "Error: out of host memory\n");
        return 1;
    }
    /* values that do not parse leave zeros, as fields.c ignores fscanf's result */
    for (size_t k = 0; k < n; k++)
        if (fscanf(file, "%lf", &p[k]) != 1) break;
    fclose(file);

    const int rc = pfc_fields(p, W, H, dx, dy, fold, sk, sb, power, phi, chi);
    if (rc) {
        fprintf(stderr, "Error: pfc_fields failed: %s\n", pfc_strerror(rc));
        return 42;
    }
    if (write_file(name_phi, W, H, dx, dy, phi, 2) ||
        (with_chi && write_file(name_chi, W, H, dx, dy, chi, 1))) {
        fprintf(stderr, "Error: could not write the output files\n");
        return 1;
    }
    free(p);
    free(phi);
    free(chi);
    return 0;
}

/*
 * pfc_io.h -- input-file grammar and text formats of pfc-relax.
 *
 * Everything here is host-side file handling; the numerics live behind
 * include/pfc_relax.h.  Messages, exit codes and text formats are those of
 * the reference's src/pfc-relax.c (cited per function) so the binary drops
 * into the same pipelines.
 */
#ifndef PFC_IO_H
#define PFC_IO_H

#include <stdio.h>

#include "pfc_relax.h"

#define PFC_LINE 8192   /* the reference's BUFSIZ-sized line and name buffers */

typedef struct run_cfg {
    char input[PFC_LINE];                      /* input file name (argv[1]) */
    int U;
    char innames[PFC_MAX_FIELDS][PFC_LINE];    /* density input files */
    char outnames[PFC_MAX_FIELDS][PFC_LINE];   /* density output files ('#' -> step) */
    char fname[PFC_LINE];                      /* energy density output file */
    int has_fname;
    int Nx, Ny;
    int Tprint, Twrite, pprint, pwrite;
    int* twrite;                               /* write steps, ascending */
    int ntwrite;
    double c, dx, dy, dt;
    int tend, Topt, xopt, yopt;
    pfc_params p;
} run_cfg;

/* Prints the reference's message for fail code i on rank 0 and exits with
 * status i (src/pfc-relax.c:242-411; MPI_Abort(i) becomes exit(i)). */
void pfc_fail(int i);

/* Parses the input file (src/pfc-relax.c:414-725).  Fails like the reference. */
void pfc_read_input(run_cfg* cfg);

/* Reads the U density files into fields (U x Ny x Nx), failing with codes
 * 37 (missing), 4 (header) or 38 (data) (src/pfc-relax.c:728-763). */
void pfc_read_data(const run_cfg* cfg, double* fields);

/* Output file name: the first '#' replaced by the decimal step
 * (src/pfc-relax.c:785-792). */
void pfc_step_name(const char* pattern, int t, char* out, size_t n);

/* Writes one field file: header "Nx Ny dx dy" at pwrite decimals, then one
 * "%.*e" value per line, row by row (src/pfc-relax.c:766-858). */
int pfc_write_field(const char* path, const double* f, int Nx, int Ny, double dx, double dy,
                    int pwrite);

/* The same with separate precisions for the header (%.*lf) and the values
 * (%.*e); smooth writes "%lf" and "%.12e" (src/smooth.c:123-130). */
int pfc_write_field_prec(const char* path, const double* f, int Nx, int Ny, double dx, double dy,
                         int phead, int pvalue);

/* The value lines alone: one "%.*e" per value, formatted in parallel. */
void pfc_write_values(FILE* out, const double* f, size_t n, int pvalue);

/* Log line of print_output (src/pfc-relax.c:861-897) to stdout and to the
 * ".out" file named after the input file. */
void pfc_print_output(const run_cfg* cfg, int t, int elapsed, double dx, double dy, double fave,
                      const double* nave);

#endif

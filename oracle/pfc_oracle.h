/*
 * pfc_oracle.h -- CPU FP64 restatement of pfc-relax's relaxation loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU engine is compared
 * against; nothing in the product (libpfcgpu, the pfc-relax binary, the
 * Python package) links, loads or calls it.  Only tests/, the smoke check
 * and bench.py's CPU-baseline leg use it.
 *
 * Every function follows a function of the reference's src/pfc-relax.c
 * (single process, so MPI ranks R = 1) and cites it.  Layouts follow the
 * reference: real fields are Ny rows of Nx values; k-space arrays are the
 * transposed half spectrum, index kx*Ny + ky with kx in [0, Nx/2].
 */
#ifndef PFC_ORACLE_H
#define PFC_ORACLE_H

#include "../include/pfc_relax.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

/* ---- FFT building blocks (FFTW conventions: forward e^{-2 pi i jk/n},
 *      backward e^{+...}, both unnormalised) ---- */

/* 1D complex DFT of any length n >= 1; data as interleaved re,im.
 * sign = -1 forward, +1 backward. */
int oracle_fft1d(int n, const double* in, double* out, int sign);

/* 2D real-to-complex: in is Ny x Nx real (row-major); out is the transposed
 * half spectrum (Nx/2+1) x Ny complex, interleaved.  This is what
 * fftw_mpi_plan_dft_r2c_2d(Ny, Nx, ..., FFTW_MPI_TRANSPOSED_OUT) computes on
 * one rank (src/pfc-relax.c:227-228). */
int oracle_r2c2d(int Nx, int Ny, const double* in, double* out);

/* 2D complex-to-real, inverse of the above up to the factor NxNy
 * (fftw_mpi_plan_dft_c2r_2d with FFTW_MPI_TRANSPOSED_IN, :231-235).  The
 * imaginary parts of the kx = 0 and kx = Nx/2 entries after the column pass
 * are ignored, as a 1D c2r ignores them. */
int oracle_c2r2d(int Nx, int Ny, const double* in, double* out);

/* ---- solver ---- */

oracle* oracle_create(const pfc_params* p, int Nx, int Ny);
void oracle_destroy(oracle* o);
int oracle_set_fields(oracle* o, const double* fields);     /* U x Ny x Nx */
int oracle_get_fields(const oracle* o, double* fields);
int oracle_get_energy(const oracle* o, double* f);           /* Ny x Nx */
int oracle_get_spectral(const oracle* o, int u, double* out); /* N_u, (Nx/2+1) x Ny complex */
int oracle_get_AB(const oracle* o, int u, double* A, double* B);
int oracle_set_box(oracle* o, double dx, double dy);         /* update_AB */
/* One step at relaxation step t.  Returns 39 (the reference's fail code)
 * if m[0][0] is NaN, else 0. */
int oracle_step(oracle* o, int t);
int oracle_evaluate(oracle* o, double* fave, double* nave);
/* Box-size optimisation; *d is the sampling step carried between calls. */
int oracle_optimize(oracle* o, double* d, int xopt, int yopt);
/* The optimisation arithmetic with the energy supplied by a callback
 * f(dx, dy, user) -- the same code path as oracle_optimize. */
typedef double (*oracle_energy_fn)(double dx, double dy, void* user);
int oracle_optimize_with(int Nx, int Ny, double* dx, double* dy, double* d,
                         int xopt, int yopt, oracle_energy_fn fn, void* user);
void oracle_get_box(const oracle* o, double* dx, double* dy);

/* The relax() loop without file output: records one row per printed step,
 * [t, dx, dy, fave, nave_0 .. nave_{U-1}] into log (row stride 4+U), at most
 * max_rows rows.  snap_steps (ascending, n_snap entries) lists steps at which
 * the fields are copied into snaps (n_snap x U x Ny x Nx), mirroring
 * write_data, and, if fsnaps is not NULL, the energy density of the last
 * evaluate into fsnaps (n_snap x Ny x Nx), mirroring write_energy.  Returns
 * the number of rows, or -39 on a NaN. */
int oracle_relax(oracle* o, int tend, int Tprint, int Topt, int xopt, int yopt,
                 double* log, int max_rows, const int* snap_steps, int n_snap,
                 double* snaps, double* fsnaps);

#ifdef __cplusplus
}
#endif

#endif

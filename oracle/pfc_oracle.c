/*
 * pfc_oracle.c -- CPU FP64 restatement of pfc-relax's relaxation loop.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU engine (see
 * pfc_oracle.h).  Compiled with -ffp-contract=off so products and sums round
 * as in the reference's x86-64 build (mpicc -O3, no FMA).
 *
 * Function-by-function correspondence with the reference's src/pfc-relax.c:
 *   oracle_set_box     <- update_AB      :900-957
 *   oracle_step        <- step           :1133-1247
 *   oracle_evaluate    <- evaluate       :987-1130 (+ descale :967-984)
 *   oracle_optimize    <- optimize       :1250-1289
 *   oracle_relax       <- relax          :1292-1312
 *   oracle_r2c2d/c2r2d <- the FFTW-MPI plans of plan_FFTs :196-239
 * Deliberate differences (all roundoff-level or defect-related):
 *   - fave/nave sums are compensated (Neumaier) instead of the reference's
 *     naive running sums (:1114-1121), so the oracle is the better estimate;
 *   - the FFT is this file's own Stockham transform, not FFTW;
 *   - evaluate's beta_uv gradient term uses mu2 (the documented model) unless
 *     p->energy_compat == PFC_ENERGY_NU2, which restates the reference's
 *     nu2-for-mu2 alias (:1019, :1064) with 0 past the end of nu2[u].
 */
#include "pfc_oracle.h"

/* ORACLE_MUTANT plants one known mistake, for tests/test_oracle_mutants.py,
 * which builds such copies and checks that the closed-form pins reject them:
 *   1  step: m_u and m_v swapped in the gamma_uv term of mc[v] (:1196-1197)
 *   2  relax: the step taken before the print and write of the same t
 *   3  relax: the print taken before the optimisation of the same t
 *   4  relax: the snapshot taken after the step of the same t
 * The shipped build (the Makefile's liboracle.so) has ORACLE_MUTANT 0. */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_L 3.141592653589793238462643383279502884L

typedef struct {
    double re, im;
} cplx;

/* ------------------------------------------------------------------ FFT */

typedef struct {
    int n;
    int nr;          /* number of stages */
    int radix[40];
    cplx* tw;        /* tw[m] = exp(-2 pi i m / n), computed in long double */
} fftplan;

/* Stages: radix 4, 2 and 3 first, then every other prime factor of n as a
 * stage of its own, computed as a direct DFT (any n; a prime n is one O(n^2)
 * stage, which the small grids of the tests afford). */
static int plan_init(fftplan* pl, int n) {
    int m = n, k = 0;
    if (n < 1) return -1;
    while (m % 4 == 0) { pl->radix[k++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->radix[k++] = 2; m /= 2; }
    while (m % 3 == 0) { pl->radix[k++] = 3; m /= 3; }
    for (int p = 5; m > 1 && p * p <= m; p += 2)
        while (m % p == 0) { pl->radix[k++] = p; m /= p; }
    if (m > 1) pl->radix[k++] = m;
    pl->n = n;
    pl->nr = k;
    pl->tw = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int i = 0; i < n; i++) {
        long double a = -2.0L * PI_L * (long double)i / (long double)n;
        pl->tw[i].re = (double)cosl(a);
        pl->tw[i].im = (double)sinl(a);
    }
    return 0;
}

static void plan_free(fftplan* pl) { free(pl->tw); pl->tw = NULL; }

/* twiddle exp(sign 2 pi i m / n) */
static inline cplx twid(const fftplan* pl, long m, int sign) {
    cplx w = pl->tw[m % pl->n];
    if (sign > 0) w.im = -w.im;
    return w;
}

static inline cplx cmul(cplx a, cplx b) {
    cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/* Stockham autosort: x (length n, stride 1) -> result in x; work is scratch
 * of length n. */
static void fft_exec(const fftplan* pl, cplx* x, cplx* work, int sign) {
    const int n = pl->n;
    cplx* src = x;
    cplx* dst = work;
    int Ns = 1;
    const double s3 = 0.86602540378443864676; /* sin(pi/3) */
    for (int s = 0; s < pl->nr; s++) {
        const int R = pl->radix[s];
        const int nR = n / R;
        const long twstep = n / (Ns * R);
        if (R > 4) {
            /* direct DFT of a prime radix: y[s] = sum_r v[r] w^(r s), w = e^{sign 2 pi i/R} */
            cplx v[R], y[R];
            const long wstep = n / R;
            for (int j = 0; j < nR; j++) {
                const int k = j % Ns;
                for (int r = 0; r < R; r++) {
                    v[r] = src[j + r * nR];
                    if (r > 0 && k > 0) v[r] = cmul(v[r], twid(pl, (long)k * r * twstep, sign));
                }
                for (int q = 0; q < R; q++) {
                    cplx acc = v[0];
                    for (int r = 1; r < R; r++) {
                        const cplx t = cmul(v[r], twid(pl, (long)((r * (long)q) % R) * wstep, sign));
                        acc.re += t.re;
                        acc.im += t.im;
                    }
                    y[q] = acc;
                }
                const int base = (j / Ns) * Ns * R + k;
                for (int r = 0; r < R; r++) dst[base + r * Ns] = y[r];
            }
            cplx* t = src; src = dst; dst = t;
            Ns *= R;
            continue;
        }
        for (int j = 0; j < nR; j++) {
            const int k = j % Ns;
            cplx v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
            for (int r = 0; r < R; r++) {
                v[r] = src[j + r * nR];
                if (r > 0 && k > 0) v[r] = cmul(v[r], twid(pl, (long)k * r * twstep, sign));
            }
            cplx y[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
            if (R == 2) {
                y[0].re = v[0].re + v[1].re; y[0].im = v[0].im + v[1].im;
                y[1].re = v[0].re - v[1].re; y[1].im = v[0].im - v[1].im;
            } else if (R == 4) {
                cplx a0, a1, b0, b1;
                a0.re = v[0].re + v[2].re; a0.im = v[0].im + v[2].im;
                a1.re = v[0].re - v[2].re; a1.im = v[0].im - v[2].im;
                b0.re = v[1].re + v[3].re; b0.im = v[1].im + v[3].im;
                b1.re = v[1].re - v[3].re; b1.im = v[1].im - v[3].im;
                /* multiply b1 by sign*i: (sign i)(re + i im) = -sign im + i sign re */
                cplx jb1;
                jb1.re = -sign * b1.im; jb1.im = sign * b1.re;
                y[0].re = a0.re + b0.re; y[0].im = a0.im + b0.im;
                y[2].re = a0.re - b0.re; y[2].im = a0.im - b0.im;
                y[1].re = a1.re + jb1.re; y[1].im = a1.im + jb1.im;
                y[3].re = a1.re - jb1.re; y[3].im = a1.im - jb1.im;
            } else { /* R == 3 */
                cplx t1, t2;
                t1.re = v[1].re + v[2].re; t1.im = v[1].im + v[2].im;
                t2.re = v[1].re - v[2].re; t2.im = v[1].im - v[2].im;
                y[0].re = v[0].re + t1.re; y[0].im = v[0].im + t1.im;
                cplx m1;
                m1.re = v[0].re - 0.5 * t1.re; m1.im = v[0].im - 0.5 * t1.im;
                /* sign * i * sin(pi/3) * t2 */
                cplx m2;
                m2.re = -sign * s3 * t2.im; m2.im = sign * s3 * t2.re;
                y[1].re = m1.re + m2.re; y[1].im = m1.im + m2.im;
                y[2].re = m1.re - m2.re; y[2].im = m1.im - m2.im;
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) dst[base + r * Ns] = y[r];
        }
        cplx* t = src; src = dst; dst = t;
        Ns *= R;
    }
    if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)n);
}

int oracle_fft1d(int n, const double* in, double* out, int sign) {
    fftplan pl;
    if (plan_init(&pl, n)) return -1;
    cplx* x = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    memcpy(x, in, sizeof(cplx) * (size_t)n);
    fft_exec(&pl, x, w, sign);
    memcpy(out, x, sizeof(cplx) * (size_t)n);
    free(x); free(w); plan_free(&pl);
    return 0;
}

/* 2D transforms with explicit plans (reused by the solver) */
typedef struct {
    int Nx, Ny, Nh;
    fftplan px, py;
} plan2d;

static int plan2d_init(plan2d* p, int Nx, int Ny) {
    p->Nx = Nx; p->Ny = Ny; p->Nh = Nx / 2 + 1;
    if (plan_init(&p->px, Nx)) return -1;
    if (plan_init(&p->py, Ny)) { plan_free(&p->px); return -1; }
    return 0;
}

static void plan2d_free(plan2d* p) { plan_free(&p->px); plan_free(&p->py); }

/* real Ny x Nx -> transposed half spectrum Nh x Ny */
static void r2c2d(const plan2d* p, const double* in, cplx* out) {
    const int Nx = p->Nx, Ny = p->Ny, Nh = p->Nh;
    #pragma omp parallel
    {
        cplx* row = (cplx*)malloc(sizeof(cplx) * (size_t)Nx);
        cplx* wk = (cplx*)malloc(sizeof(cplx) * (size_t)(Nx > Ny ? Nx : Ny));
        #pragma omp for schedule(static)
        for (int y = 0; y < Ny; y++) {
            for (int x = 0; x < Nx; x++) { row[x].re = in[(size_t)y * Nx + x]; row[x].im = 0.0; }
            fft_exec(&p->px, row, wk, -1);
            for (int kx = 0; kx < Nh; kx++) out[(size_t)kx * Ny + y] = row[kx];
        }
        #pragma omp for schedule(static)
        for (int kx = 0; kx < Nh; kx++) fft_exec(&p->py, out + (size_t)kx * Ny, wk, -1);
        free(row); free(wk);
    }
}

/* transposed half spectrum Nh x Ny -> real Ny x Nx (unnormalised);
 * in is overwritten by the column pass */
static void c2r2d(const plan2d* p, cplx* in, double* out) {
    const int Nx = p->Nx, Ny = p->Ny, Nh = p->Nh;
    #pragma omp parallel
    {
        cplx* row = (cplx*)malloc(sizeof(cplx) * (size_t)Nx);
        cplx* wk = (cplx*)malloc(sizeof(cplx) * (size_t)(Nx > Ny ? Nx : Ny));
        #pragma omp for schedule(static)
        for (int kx = 0; kx < Nh; kx++) fft_exec(&p->py, in + (size_t)kx * Ny, wk, +1);
        #pragma omp for schedule(static)
        for (int y = 0; y < Ny; y++) {
            /* Hermitian extension along x; DC and Nyquist keep their real
             * parts only, as a 1D c2r does */
            for (int kx = 0; kx < Nh; kx++) row[kx] = in[(size_t)kx * Ny + y];
            row[0].im = 0.0;
            if (Nx % 2 == 0) row[Nx / 2].im = 0.0;
            for (int kx = 1; kx < Nh; kx++) {
                if (Nx - kx >= Nh) {
                    row[Nx - kx].re = row[kx].re;
                    row[Nx - kx].im = -row[kx].im;
                }
            }
            fft_exec(&p->px, row, wk, +1);
            for (int x = 0; x < Nx; x++) out[(size_t)y * Nx + x] = row[x].re;
        }
        free(row); free(wk);
    }
}

int oracle_r2c2d(int Nx, int Ny, const double* in, double* out) {
    plan2d p;
    if (plan2d_init(&p, Nx, Ny)) return -1;
    r2c2d(&p, in, (cplx*)out);
    plan2d_free(&p);
    return 0;
}

int oracle_c2r2d(int Nx, int Ny, const double* in, double* out) {
    plan2d p;
    if (plan2d_init(&p, Nx, Ny)) return -1;
    size_t nk = (size_t)(Nx / 2 + 1) * Ny;
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * nk);
    memcpy(tmp, in, sizeof(cplx) * nk);
    c2r2d(&p, tmp, out);
    free(tmp);
    plan2d_free(&p);
    return 0;
}

/* --------------------------------------------------------------- solver */

struct oracle {
    pfc_params p;
    int U, Nx, Ny, Nh;
    size_t nr, nk;      /* real points, k points */
    double dx, dy;
    plan2d pl;
    double** m;         /* real fields (the reference's m) */
    double** n;         /* real scratch (the reference's n in real space) */
    double** e;         /* smoothed densities */
    double** q;         /* q[u*U+v], v>u */
    double* f;          /* energy density */
    cplx** N;           /* spectral state N_u */
    cplx** M;           /* spectral scratch */
    cplx** E;
    cplx** Q;
    double** A;
    double** B;
    int have_f;
};

static double* dalloc(size_t n) { return (double*)calloc(n, sizeof(double)); }
static cplx* calloc_c(size_t n) { return (cplx*)calloc(n, sizeof(cplx)); }

oracle* oracle_create(const pfc_params* p, int Nx, int Ny) {
    if (!p || p->U < 1 || p->U > PFC_MAX_FIELDS || Nx < 2 || Ny < 2) return NULL;
    oracle* o = (oracle*)calloc(1, sizeof(oracle));
    o->p = *p;
    o->U = p->U;
    o->Nx = Nx; o->Ny = Ny; o->Nh = Nx / 2 + 1;
    o->nr = (size_t)Nx * Ny;
    o->nk = (size_t)o->Nh * Ny;
    if (plan2d_init(&o->pl, Nx, Ny)) { free(o); return NULL; }
    const int U = o->U;
    o->m = (double**)calloc(U, sizeof(double*));
    o->n = (double**)calloc(U, sizeof(double*));
    o->e = (double**)calloc(U, sizeof(double*));
    o->q = (double**)calloc((size_t)U * U, sizeof(double*));
    o->N = (cplx**)calloc(U, sizeof(cplx*));
    o->M = (cplx**)calloc(U, sizeof(cplx*));
    o->E = (cplx**)calloc(U, sizeof(cplx*));
    o->Q = (cplx**)calloc((size_t)U * U, sizeof(cplx*));
    o->A = (double**)calloc(U, sizeof(double*));
    o->B = (double**)calloc(U, sizeof(double*));
    for (int u = 0; u < U; u++) {
        o->m[u] = dalloc(o->nr);
        o->n[u] = dalloc(o->nr);
        o->e[u] = dalloc(o->nr);
        o->N[u] = calloc_c(o->nk);
        o->M[u] = calloc_c(o->nk);
        o->E[u] = calloc_c(o->nk);
        o->A[u] = dalloc(o->nk);
        o->B[u] = dalloc(o->nk);
        for (int v = u + 1; v < U; v++) {
            o->q[u * U + v] = dalloc(o->nr);
            o->Q[u * U + v] = calloc_c(o->nk);
        }
    }
    o->f = dalloc(o->nr);
    return o;
}

void oracle_destroy(oracle* o) {
    if (!o) return;
    const int U = o->U;
    for (int u = 0; u < U; u++) {
        free(o->m[u]); free(o->n[u]); free(o->e[u]);
        free(o->N[u]); free(o->M[u]); free(o->E[u]);
        free(o->A[u]); free(o->B[u]);
        for (int v = u + 1; v < U; v++) { free(o->q[u * U + v]); free(o->Q[u * U + v]); }
    }
    free(o->m); free(o->n); free(o->e); free(o->q);
    free(o->N); free(o->M); free(o->E); free(o->Q);
    free(o->A); free(o->B); free(o->f);
    plan2d_free(&o->pl);
    free(o);
}

static void descale(oracle* o, cplx** X);

/* read_data's store into m (:728-763), plus N = FFT(m)/NxNy so that a step
 * that does not refresh (t % 100 != 0) starts from a defined state; in the
 * reference the first step (t = 0) always refreshes. */
int oracle_set_fields(oracle* o, const double* fields) {
    for (int u = 0; u < o->U; u++) {
        memcpy(o->m[u], fields + (size_t)u * o->nr, sizeof(double) * o->nr);
        r2c2d(&o->pl, o->m[u], o->N[u]);
    }
    descale(o, o->N);
    return 0;
}

int oracle_get_fields(const oracle* o, double* fields) {
    for (int u = 0; u < o->U; u++) memcpy(fields + (size_t)u * o->nr, o->m[u], sizeof(double) * o->nr);
    return 0;
}

int oracle_get_energy(const oracle* o, double* f) {
    memcpy(f, o->f, sizeof(double) * o->nr);
    return 0;
}

int oracle_get_spectral(const oracle* o, int u, double* out) {
    if (u < 0 || u >= o->U) return -1;
    memcpy(out, o->N[u], sizeof(cplx) * o->nk);
    return 0;
}

int oracle_get_AB(const oracle* o, int u, double* A, double* B) {
    if (u < 0 || u >= o->U) return -1;
    memcpy(A, o->A[u], sizeof(double) * o->nk);
    memcpy(B, o->B[u], sizeof(double) * o->nk);
    return 0;
}

void oracle_get_box(const oracle* o, double* dx, double* dy) { *dx = o->dx; *dy = o->dy; }

/* ky of row j of a transposed k-space column (src/pfc-relax.c:935-936) */
static inline double ky_of(int j, int Ny, double dky) {
    return (j < Ny / 2) ? j * dky : (j - Ny) * dky;
}

/* update_AB, src/pfc-relax.c:900-957 */
int oracle_set_box(oracle* o, double dx, double dy) {
    const pfc_params* p = &o->p;
    const int U = o->U, Nx = o->Nx, Ny = o->Ny, Nh = o->Nh;
    o->dx = dx; o->dy = dy;
    const double _NxNy = 1.0 / Nx / Ny;
    const double dkx = 2.0 * (double)PI_L / (dx * Nx);
    const double dky = 2.0 * (double)PI_L / (dy * Ny);
    const double c = (p->c != 0.0) ? 1.0 : 0.0;
    const double dt = p->dt;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nh; i++) {
        double kx2 = i * dkx;
        kx2 *= kx2;
        for (int j = 0; j < Ny; j++) {
            const double ky = ky_of(j, Ny, dky);
            const double k2 = kx2 + ky * ky;
            const size_t ij = (size_t)Ny * i + j;
            for (int u = 0; u < U; u++) {
                double P = 1.0;
                for (int w = 0; w < p->W[u]; w++) {
                    const double pw = p->nu2[u][w] - k2;
                    P *= p->omega[u][w] * pw * pw + p->phi[u][w];
                }
                const double a = p->alpha[u][u] + p->beta[u][u] * P;
                const double oo = c - 1.0 - c * k2;
                const double ex = exp(oo * a * dt);
                o->A[u][ij] = ex;
                double b = (a == 0.0) ? oo * dt : (ex - 1.0) / a;
                o->B[u][ij] = b * _NxNy;
            }
        }
    }
    return 0;
}

/* descale, :967-984 */
static void descale(oracle* o, cplx** X) {
    const double _NxNy = 1.0 / o->Nx / o->Ny;
    for (int u = 0; u < o->U; u++) {
        cplx* x = X[u];
        #pragma omp parallel for schedule(static)
        for (size_t i = 0; i < o->nk; i++) { x[i].re *= _NxNy; x[i].im *= _NxNy; }
    }
}

/* step, :1133-1247 */
int oracle_step(oracle* o, int t) {
    const pfc_params* p = &o->p;
    const int U = o->U, Nx = o->Nx, Ny = o->Ny, Nh = o->Nh;
    if (o->m[0][0] != o->m[0][0]) return 39;                      /* :1175 */
    if (t % 100 == 0) {                                           /* :1178-1184 */
        for (int u = 0; u < U; u++) r2c2d(&o->pl, o->m[u], o->N[u]);
        descale(o, o->N);
    }
    /* nonlinear part in direct space, :1187-1208 */
    #pragma omp parallel for schedule(static)
    for (size_t ij = 0; ij < o->nr; ij++) {
        double mc[PFC_MAX_FIELDS];
        for (int u = 0; u < U; u++) mc[u] = 0.0;
        for (int u = 0; u < U - 1; u++) {
            for (int v = u + 1; v < U; v++) {
                const double mu = o->m[u][ij], mv = o->m[v][ij];
                mc[u] += p->alpha[u][v] * mv + 0.5 * p->gamma[u][v] * (2.0 * mu * mv + mv * mv);
#if ORACLE_MUTANT == 1
                mc[v] += p->alpha[u][v] * mu + 0.5 * p->gamma[u][v] * (2.0 * mu * mv + mv * mv);
#else
                mc[v] += p->alpha[u][v] * mu + 0.5 * p->gamma[u][v] * (2.0 * mv * mu + mu * mu);
#endif
            }
        }
        for (int u = 0; u < U; u++) {
            const double mu = o->m[u][ij];
            const double m2 = mu * mu;
            o->m[u][ij] = p->lambda[u] + p->gamma[u][u] * m2 + p->delta[u] * mu * m2 + mc[u];
        }
    }
    for (int u = 0; u < U; u++) r2c2d(&o->pl, o->m[u], o->M[u]);  /* :1211 */
    /* k-space update, :1214-1242 */
    const double NxNy = (double)Nx * Ny;
    const double dkx = 2.0 * (double)PI_L / (o->dx * Nx);
    const double dky = 2.0 * (double)PI_L / (o->dy * Ny);
    const double a = -0.5 / p->sigma / p->sigma;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nh; i++) {
        double kx2 = i * dkx;
        kx2 *= kx2;
        for (int j = 0; j < Ny; j++) {
            const double ky = ky_of(j, Ny, dky);
            const double k2 = kx2 + ky * ky;
            const size_t ij = (size_t)Ny * i + j;
            double G2 = exp(a * k2);
            G2 *= G2;
            for (int u = 0; u < U - 1; u++) {
                for (int v = u + 1; v < U; v++) {
                    double d2 = p->mu2[u][v] - k2;
                    d2 *= d2;
                    const cplx Nv = o->N[v][ij], Nu = o->N[u][ij];
                    const double bu = p->beta[u][v] * d2, eu = p->epsilon[u][v] * G2;
                    o->M[u][ij].re += (bu * Nv.re + eu * Nv.re) * NxNy;
                    o->M[u][ij].im += (bu * Nv.im + eu * Nv.im) * NxNy;
                    o->M[v][ij].re += (bu * Nu.re + eu * Nu.re) * NxNy;
                    o->M[v][ij].im += (bu * Nu.im + eu * Nu.im) * NxNy;
                }
            }
            for (int u = 0; u < U; u++) {
                const double Au = o->A[u][ij], Bu = o->B[u][ij];
                cplx r;
                r.re = Au * o->N[u][ij].re + Bu * o->M[u][ij].re;
                r.im = Au * o->N[u][ij].im + Bu * o->M[u][ij].im;
                o->N[u][ij] = r;
                o->M[u][ij] = r;
            }
        }
    }
    for (int u = 0; u < U; u++) c2r2d(&o->pl, o->M[u], o->m[u]);  /* :1245 */
    return 0;
}

/* Neumaier compensated accumulator */
typedef struct { double s, c; } ksum;
static inline void kadd(ksum* k, double x) {
    double t = k->s + x;
    if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x;
    else k->c += (x - t) + k->s;
    k->s = t;
}

/* evaluate, :987-1130 */
int oracle_evaluate(oracle* o, double* fave_out, double* nave_out) {
    const pfc_params* p = &o->p;
    const int U = o->U, Nx = o->Nx, Ny = o->Ny, Nh = o->Nh;
    const double _NxNy = 1.0 / Nx / Ny;
    const double dkx = 2.0 * (double)PI_L / (o->dx * Nx);
    const double dky = 2.0 * (double)PI_L / (o->dy * Ny);
    const double _3 = 1.0 / 3.0;
    /* n <- m; M <- FFT(m)/NxNy, :1041-1045 */
    for (int u = 0; u < U; u++) {
        memcpy(o->n[u], o->m[u], sizeof(double) * o->nr);
        r2c2d(&o->pl, o->m[u], o->M[u]);
    }
    descale(o, o->M);
    /* operators in k space, :1048-1077 */
    const double a = -0.5 / p->sigma / p->sigma;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nh; i++) {
        double kx2 = i * dkx;
        kx2 *= kx2;
        for (int j = 0; j < Ny; j++) {
            const double ky = ky_of(j, Ny, dky);
            const double k2 = kx2 + ky * ky;
            const size_t ij = (size_t)Ny * i + j;
            const double G = exp(a * k2);
            for (int u = 0; u < U; u++) {
                o->E[u][ij].re = G * o->M[u][ij].re;
                o->E[u][ij].im = G * o->M[u][ij].im;
                for (int v = u + 1; v < U; v++) {
                    double c2;
                    if (p->energy_compat == PFC_ENERGY_NU2)
                        c2 = (v < p->W[u]) ? p->nu2[u][v] : 0.0;
                    else
                        c2 = p->mu2[u][v];
                    const double d = c2 - k2;
                    o->Q[u * U + v][ij].re = d * d * o->M[v][ij].re;
                    o->Q[u * U + v][ij].im = d * d * o->M[v][ij].im;
                }
                double P = 1.0;
                for (int w = 0; w < p->W[u]; w++) {
                    const double pw = p->nu2[u][w] - k2;
                    P *= p->omega[u][w] * pw * pw + p->phi[u][w];
                }
                o->M[u][ij].re *= P;
                o->M[u][ij].im *= P;
            }
        }
    }
    for (int u = 0; u < U; u++) {                                 /* :1080-1086 */
        c2r2d(&o->pl, o->M[u], o->m[u]);
        c2r2d(&o->pl, o->E[u], o->e[u]);
        for (int v = u + 1; v < U; v++) c2r2d(&o->pl, o->Q[u * U + v], o->q[u * U + v]);
    }
    /* direct-space energy, :1093-1117 */
    ksum fs = {0.0, 0.0};
    ksum ns[PFC_MAX_FIELDS];
    for (int u = 0; u < U; u++) { ns[u].s = 0.0; ns[u].c = 0.0; }
    for (size_t ij = 0; ij < o->nr; ij++) {
        double fij = 0.0;
        for (int u = 0; u < U; u++) {
            const double nu = o->n[u][ij];
            const double nunu = nu * nu;
            fij += p->lambda[u] * nu + 0.5 * p->alpha[u][u] * nunu
                 + 0.5 * p->beta[u][u] * nu * o->m[u][ij]
                 + _3 * p->gamma[u][u] * nu * nunu + 0.25 * p->delta[u] * nunu * nunu;
            for (int v = u + 1; v < U; v++) {
                const double nv = o->n[v][ij];
                const double nvnv = nv * nv;
                fij += p->alpha[u][v] * nu * nv + p->beta[u][v] * nu * o->q[u * U + v][ij]
                     + 0.5 * p->gamma[u][v] * (nunu * nv + nu * nvnv)
                     + p->epsilon[u][v] * o->e[u][ij] * o->e[v][ij];
            }
            o->m[u][ij] = nu;
            kadd(&ns[u], nu);
        }
        o->f[ij] = fij;
        kadd(&fs, fij);
    }
    o->have_f = 1;
    const double fave = (fs.s + fs.c) * _NxNy;
    for (int u = 0; u < U; u++) nave_out[u] = (ns[u].s + ns[u].c) * _NxNy;
    *fave_out = fave;
    /* restore the spectral state, :1124-1128 */
    for (int u = 0; u < U; u++) r2c2d(&o->pl, o->n[u], o->N[u]);
    descale(o, o->N);
    return 0;
}

/* optimize, :1250-1289, with the energy behind a callback */
int oracle_optimize_with(int Nx, int Ny, double* pdx, double* pdy, double* pd,
                         int xopt, int yopt, oracle_energy_fn fn, void* user) {
    const double d = *pd;
    const double dxs[5] = {*pdx, *pdx - d, *pdx + d, *pdx, *pdx};
    const double dys[5] = {*pdy, *pdy, *pdy, *pdy - d, *pdy + d};
    double fs[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < 5; i++) {
        if (xopt == 0 && (i == 1 || i == 2)) continue;
        if (yopt == 0 && (i == 3 || i == 4)) continue;
        fs[i] = fn(dxs[i], dys[i], user);
    }
    double dx = dxs[0], dy = dys[0];
    if (xopt != 0)
        dx = (d * (fs[1] - fs[2]) + 2.0 * dxs[0] * (-2.0 * fs[0] + fs[1] + fs[2])) / (2.0 * (fs[1] - 2.0 * fs[0] + fs[2]));
    if (yopt != 0)
        dy = (d * (fs[3] - fs[4]) + 2.0 * dys[0] * (-2.0 * fs[0] + fs[3] + fs[4])) / (2.0 * (fs[3] - 2.0 * fs[0] + fs[4]));
    /* when an axis is not optimised the reference leaves dx/dy at the last
     * sample it evaluated, which is dxs[0]/dys[0] for that axis */
    const double l0 = 4.0 * (double)PI_L / sqrt(3.0);
    const double dw = Nx * (dx - dxs[0]);
    const double dh = Ny * (dy - dys[0]);
    const double dr = sqrt(dw * dw + dh * dh);
    double x = 0.25 * l0;
    if (dr > x) {
        x /= dr;
        dx = x * dx + (1.0 - x) * dxs[0];
        dy = x * dy + (1.0 - x) * dys[0];
    }
    const double ddx = dx - dxs[0];
    const double ddy = dy - dys[0];
    const double ddr = sqrt(ddx * ddx + ddy * ddy);
    double dn = d;
    if (ddr < dn) dn *= 0.5;
    else dn *= 2.0;
    if (dn < 1.0e-6) dn *= 2.0;
    *pdx = dx; *pdy = dy; *pd = dn;
    return 0;
}

typedef struct { oracle* o; } opt_ctx;

static double opt_energy(double dx, double dy, void* user) {
    oracle* o = ((opt_ctx*)user)->o;
    double fave, nave[PFC_MAX_FIELDS];
    o->dx = dx; o->dy = dy;   /* evaluate reads dx, dy; A, B are untouched */
    oracle_evaluate(o, &fave, nave);
    return fave;
}

int oracle_optimize(oracle* o, double* d, int xopt, int yopt) {
    opt_ctx c = {o};
    double dx = o->dx, dy = o->dy;
    oracle_optimize_with(o->Nx, o->Ny, &dx, &dy, d, xopt, yopt, opt_energy, &c);
    oracle_set_box(o, dx, dy);   /* update_AB, :1281 */
    return 0;
}

/* relax, :1292-1312 (file output replaced by in-memory snapshots) */
int oracle_relax(oracle* o, int tend, int Tprint, int Topt, int xopt, int yopt,
                 double* log, int max_rows, const int* snap_steps, int n_snap,
                 double* snaps, double* fsnaps) {
    const int U = o->U;
    int rows = 0, ns = 0;
    double d = 0.0001;
    for (int t = 0; t <= tend; t++) {
#if ORACLE_MUTANT == 2
        if (t < tend && oracle_step(o, t) == 39) return -39;
#endif
#if ORACLE_MUTANT != 3
        if (Topt > 0 && t > 0 && t % Topt == 0) oracle_optimize(o, &d, xopt, yopt);
#endif
        if (Tprint > 0 && t % Tprint == 0) {
            double fave, nave[PFC_MAX_FIELDS];
            oracle_evaluate(o, &fave, nave);
            if (rows < max_rows) {
                double* r = log + (size_t)rows * (4 + U);
                r[0] = t; r[1] = o->dx; r[2] = o->dy; r[3] = fave;
                for (int u = 0; u < U; u++) r[4 + u] = nave[u];
                rows++;
            }
        }
#if ORACLE_MUTANT == 3
        if (Topt > 0 && t > 0 && t % Topt == 0) oracle_optimize(o, &d, xopt, yopt);
#endif
#if ORACLE_MUTANT == 4
        if (t < tend && oracle_step(o, t) == 39) return -39;
#endif
        if (ns < n_snap && t == snap_steps[ns]) {
            oracle_get_fields(o, snaps + (size_t)ns * U * o->nr);
            if (fsnaps) oracle_get_energy(o, fsnaps + (size_t)ns * o->nr);
            ns++;
        }
#if ORACLE_MUTANT != 2 && ORACLE_MUTANT != 4
        if (t < tend) {
            if (oracle_step(o, t) == 39) return -39;
        }
#endif
    }
    return rows;
}

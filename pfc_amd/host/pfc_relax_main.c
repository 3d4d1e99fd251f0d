/*
 * pfc_relax_main.c -- the pfc-relax command: same argv contract, input
 * grammar, files, log lines and exit codes as the reference's
 * src/pfc-relax.c (main :1391-1434, relax :1292-1312, optimize :1250-1289),
 * with the numerics on one GPU through libpfcgpu (include/pfc_relax.h).
 *
 * Launched as `mpirun -np N pfc-relax X.in` (the pipelines' call) only rank
 * 0 works -- the whole grid lives on one GPU -- and other ranks exit 0.
 *
 * Environment (build-only knobs; the command line is unchanged):
 *   PFC_PRECISION=mixed      FP32 transforms/update, FP64 reductions
 *   PFC_ENERGY_COMPAT=1      reproduce evaluate's nu2-for-mu2 alias
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pfc_io.h"
#include "pfc_relax.h"

int pfc_rank = 0;

static const char kUsage[] =
    "This is a program for relaxing PFC systems. It expects a single argument specifying an "
    "input file. An input file in turn specifies the input and output of a relaxation, model "
    "parameters and relaxation parameters. This program supports multiple length scales and "
    "multicomponent systems with multiple density fields and different couplings between them, "
    "including a new coupling that enables straightforward modeling of heterostructures (see "
    "arXiv:1908.05564). Both conserved and nonconserved dynamics are supported. See the source "
    "code for definitions of the model and the dynamics. The program also implements a system "
    "size optimization algorithm; see Sec. 4.3.1 in my thesis (available: "
    "http://urn.fi/URN:ISBN:978-952-60-8608-8). This program uses MPI and FFTW's fast Fourier "
    "transform routines for efficient computation.";

/* Rank of this process under an MPI launcher, from its environment. */
static int launcher_rank(void) {
    static const char* vars[] = {"PMI_RANK", "OMPI_COMM_WORLD_RANK", "PMIX_RANK",
                                 "MV2_COMM_WORLD_RANK"};
    for (size_t i = 0; i < sizeof vars / sizeof vars[0]; i++) {
        const char* v = getenv(vars[i]);
        if (v && *v) return atoi(v);
    }
    return 0;
}

static void gpu_fail(int rc, const char* what) {
    fprintf(stderr, "Error: %s failed: %s\n", what, pfc_strerror(rc));
    exit(42);
}

typedef struct {
    run_cfg* cfg;
    pfc_ctx* ctx;
    time_t t0;
    double* nave;
    double* host;   /* U x Ny x Nx staging for writes */
} run_state;

/* evaluate (src/pfc-relax.c:987-1130).  A write_energy at step tw writes the
 * density of the last evaluate at or before tw, so only an evaluate that is
 * the last one before such a write needs the density field (pfc_evaluate);
 * every other one takes the sums alone (pfc_evaluate_sums). */
static double evaluate(run_state* s, int density) {
    double fave;
    const int rc = density ? pfc_evaluate(s->ctx, &fave, s->nave)
                           : pfc_evaluate_sums(s->ctx, &fave, s->nave);
    if (rc) gpu_fail(rc, density ? "pfc_evaluate" : "pfc_evaluate_sums");
    return fave;
}

/* first step after t that evaluates (a print or an optimisation) */
static int next_evaluate_step(const run_cfg* c, int t) {
    int te = INT_MAX;
    if (c->Tprint > 0) {
        const int k = (t / c->Tprint + 1) * c->Tprint;
        if (k < te) te = k;
    }
    if (c->Topt > 0) {
        const int k = (t / c->Topt + 1) * c->Topt;
        if (k < te) te = k;
    }
    return te;
}

/* whether the last evaluate at step t feeds an energy file: the next write
 * step tw (index nw) has t <= tw < the next evaluating step */
static int density_due(const run_cfg* c, int t, int nw) {
    return c->has_fname && nw < c->ntwrite && c->twrite[nw] >= t &&
           c->twrite[nw] < next_evaluate_step(c, t);
}

static double energy_at(run_state* s, double dx, double dy, int density) {
    const int rc = pfc_set_box(s->ctx, dx, dy);
    if (rc) gpu_fail(rc, "pfc_set_box");
    return evaluate(s, density);
}

/* optimize (src/pfc-relax.c:1250-1289): parabola vertex per axis from
 * energies at dx -+ d and dy -+ d, change limited to a quarter of the
 * lattice constant l0 = 4 pi/sqrt(3), then the sampling step adapts.
 * last_density: the last sample's evaluate feeds an energy file. */
static void optimize(run_state* s, double* d, int last_density) {
    run_cfg* c = s->cfg;
    const double dd = *d;
    const double dxs[5] = {c->dx, c->dx - dd, c->dx + dd, c->dx, c->dx};
    const double dys[5] = {c->dy, c->dy, c->dy, c->dy - dd, c->dy + dd};
    double fs[5] = {0, 0, 0, 0, 0};
    const int last = c->yopt != 0 ? 4 : (c->xopt != 0 ? 2 : 0);
    for (int i = 0; i < 5; i++) {
        if (c->xopt == 0 && (i == 1 || i == 2)) continue;
        if (c->yopt == 0 && (i == 3 || i == 4)) continue;
        fs[i] = energy_at(s, dxs[i], dys[i], i == last && last_density);
    }
    double dx = dxs[0], dy = dys[0];
    if (c->xopt != 0)
        dx = (dd * (fs[1] - fs[2]) + 2.0 * dxs[0] * (-2.0 * fs[0] + fs[1] + fs[2])) /
             (2.0 * (fs[1] - 2.0 * fs[0] + fs[2]));
    if (c->yopt != 0)
        dy = (dd * (fs[3] - fs[4]) + 2.0 * dys[0] * (-2.0 * fs[0] + fs[3] + fs[4])) /
             (2.0 * (fs[3] - 2.0 * fs[0] + fs[4]));
    const double l0 = 4.0 * 3.14159265358979323846264338327 / sqrt(3.0);
    const double dw = c->Nx * (dx - dxs[0]);
    const double dh = c->Ny * (dy - dys[0]);
    const double dr = sqrt(dw * dw + dh * dh);
    double x = 0.25 * l0;
    if (dr > x) {
        x /= dr;
        dx = x * dx + (1.0 - x) * dxs[0];
        dy = x * dy + (1.0 - x) * dys[0];
    }
    c->dx = dx;
    c->dy = dy;
    const int rc = pfc_set_box(s->ctx, dx, dy);
    if (rc) gpu_fail(rc, "pfc_set_box");
    const double ddx = dx - dxs[0], ddy = dy - dys[0];
    const double ddr = sqrt(ddx * ddx + ddy * ddy);
    if (ddr < *d) *d *= 0.5;
    else *d *= 2.0;
    if (*d < 1.0e-6) *d *= 2.0;
}

static void write_outputs(run_state* s, int t) {
    run_cfg* c = s->cfg;
    char name[PFC_LINE + 32];
    const size_t n = (size_t)c->Nx * c->Ny;
    int rc = pfc_get_fields(s->ctx, s->host);
    if (rc) gpu_fail(rc, "pfc_get_fields");
    for (int u = 0; u < c->U; u++) {
        pfc_step_name(c->outnames[u], t, name, sizeof name);
        pfc_write_field(name, s->host + (size_t)u * n, c->Nx, c->Ny, c->dx, c->dy, c->pwrite);
    }
    if (c->has_fname) {
        rc = pfc_get_energy(s->ctx, s->host);
        if (rc) gpu_fail(rc, "pfc_get_energy");
        pfc_step_name(c->fname, t, name, sizeof name);
        pfc_write_field(name, s->host, c->Nx, c->Ny, c->dx, c->dy, c->pwrite);
    }
}

/* relax (src/pfc-relax.c:1292-1312).  Steps between two events (optimise,
 * print, write) go to the GPU as one batch; the NaN check of every step is
 * recorded on the device and read after the batch, before any later output. */
static void relax(run_state* s) {
    run_cfg* c = s->cfg;
    double d = 0.0001;
    int nw = 0;
    int t = 0;
    while (t <= c->tend) {
        const int print = c->Tprint > 0 && t % c->Tprint == 0;
        const int write = nw < c->ntwrite && t == c->twrite[nw];
        if (c->Topt > 0 && t > 0 && t % c->Topt == 0) optimize(s, &d, !print && density_due(c, t, nw));
        int next = c->tend;
        if (c->Tprint > 0) {
            const int k = (t / c->Tprint + 1) * c->Tprint;
            if (k < next) next = k;
        }
        if (c->Topt > 0) {
            const int k = (t / c->Topt + 1) * c->Topt;
            if (k < next) next = k;
        }
        {
            const int w = nw + write;
            if (w < c->ntwrite && c->twrite[w] > t && c->twrite[w] < next) next = c->twrite[w];
        }
        int rc;
        if (print && !write && t < c->tend && !density_due(c, t, nw)) {
            /* a print with nothing written at t: the evaluate rides on the
             * first of the steps that follow it (pfc_step_evaluate) */
            double fave = 0.0;
            rc = pfc_step_evaluate(s->ctx, t, next - t, &fave, s->nave);
            if (rc) gpu_fail(rc, "pfc_step_evaluate");
            pfc_print_output(c, t, (int)(time(NULL) - s->t0), c->dx, c->dy, fave, s->nave);
        } else {
            if (print) {
                const double fave = evaluate(s, density_due(c, t, nw));
                pfc_print_output(c, t, (int)(time(NULL) - s->t0), c->dx, c->dy, fave, s->nave);
            }
            if (write) {
                write_outputs(s, t);
                nw++;
            }
            if (t >= c->tend) break;
            rc = pfc_step(s->ctx, t, next - t);
            if (rc) gpu_fail(rc, "pfc_step");
        }
        int bad = -1;
        rc = pfc_nan_step(s->ctx, &bad);
        if (rc) gpu_fail(rc, "pfc_nan_step");
        if (bad >= 0) pfc_fail(39);
        t = next;
    }
}

int main(int argc, char** argv) {
    pfc_rank = launcher_rank();
    if (argc == 1) {
        if (pfc_rank == 0) printf("%s\n", kUsage);
        return 0;
    }
    if (argc != 2) {
        if (pfc_rank == 0) printf("Error: Invalid syntax.\n");
        return 41;
    }
    if (pfc_rank != 0) return 0;

    run_state s;
    memset(&s, 0, sizeof s);
    s.t0 = time(NULL);
    run_cfg* cfg = (run_cfg*)calloc(1, sizeof(run_cfg));
    snprintf(cfg->input, sizeof cfg->input, "%s", argv[1]);
    s.cfg = cfg;
    pfc_read_input(cfg);

    const char* prec = getenv("PFC_PRECISION");
    const int precision = (prec && strcmp(prec, "mixed") == 0) ? PFC_MIXED : PFC_FP64;
    const char* compat = getenv("PFC_ENERGY_COMPAT");
    cfg->p.energy_compat = (compat && atoi(compat) != 0) ? PFC_ENERGY_NU2 : PFC_ENERGY_MODEL;

    const size_t n = (size_t)cfg->Nx * cfg->Ny;
    s.host = (double*)malloc(sizeof(double) * n * (size_t)(cfg->U > 0 ? cfg->U : 1));
    s.nave = (double*)calloc((size_t)(cfg->U > 0 ? cfg->U : 1), sizeof(double));

    /* initialize (src/pfc-relax.c:959-964): read data, allocate, operators.
     * The data files are parsed before the device is touched so that their
     * errors (codes 4, 37, 38) come first, as they do in the reference. */
    pfc_read_data(cfg, s.host);
    int rc = pfc_create(&s.ctx, &cfg->p, cfg->Nx, cfg->Ny, precision);
    if (rc) gpu_fail(rc, "pfc_create");
    rc = pfc_set_fields(s.ctx, s.host);
    if (rc) gpu_fail(rc, "pfc_set_fields");
    rc = pfc_set_box(s.ctx, cfg->dx, cfg->dy);
    if (rc) gpu_fail(rc, "pfc_set_box");

    relax(&s);

    pfc_destroy(s.ctx);
    free(s.host);
    free(s.nave);
    free(cfg->twrite);
    free(cfg);
    return 0;
}

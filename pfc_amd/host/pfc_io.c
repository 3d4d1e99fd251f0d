/*
 * pfc_io.c -- input grammar, density/energy files and log lines of
 * pfc-relax (reference: src/pfc-relax.c, functions cited below).
 *
 * The parser reads with the same stdio primitives as the reference (token
 * scan " %s", rest-of-line fgets, sscanf of numbers) because the grammar is
 * defined by exactly those primitives: a label's values start on the
 * remainder of the label's own line, each further field takes one more line,
 * and any tokens before a label are skipped.
 */
#include "pfc_io.h"

#include <ctype.h>
#include <errno.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

extern int pfc_rank;   /* set by main from the MPI launcher's environment */

/* Error texts of fail() (src/pfc-relax.c:242-411), index = exit code. */
static const char* const kFail[41] = {
    "Error: Input file not found.",
    "Error: Invalid multicharacter label found.",
    "Error: Failed to read an input data file name.",
    "Error: Input data file not found.",
    "Error: Invalid input data file header.",
    "Error: Invalid mismatching dimensions of input data files.",
    "Error: Invalid number of output data files specified.",
    "Error: Failed to read an output data file name.",
    "Error: Failed to read the free energy density output data file name.",
    "Error: Invalid parameters for writing output.",
    "Error: Invalid write steps specified.",
    "Error: Failed to read the number of density fields.",
    "Error: Mismatching numbers of input data files and density fields specified.",
    "Error: Label for parameters lambda not found.",
    "Error: Invalid syntax for parameters lambda.",
    "Error:  Label for parameters alpha not found.",
    "Error: Invalid syntax for parameters alpha.",
    "Error:  Label for parameters beta not found.",
    "Error: Invalid syntax for parameters beta.",
    "Error:  Label for parameters gamma not found.",
    "Error: Invalid syntax for parameters gamma.",
    "Error: Label for parameters delta not found.",
    "Error: Invalid syntax for parameters delta.",
    "Error: Label for parameters epsilon not found.",
    "Error: Invalid syntax for parameters epsilon.",
    "Error: Label for parameters sigma not found.",
    "Error: Invalid syntax for parameters sigma.",
    "Error: Label for parameters 1/nu not found.",
    "Error: Invalid syntax for parameters 1/nu.",
    "Error: Label for parameters omega not found.",
    "Error: Invalid syntax for parameters omega.",
    "Error: Label for parameters phi not found.",
    "Error: Invalid syntax for parameters phi.",
    "Error: Label for parameters 1/mu not found.",
    "Error: Invalid syntax for parameters 1/mu.",
    "Error: Invalid parameters for relaxation.",
    "Error: Invalid label found.",
    "Error: Input data file not found.",
    "Error: Invalid data in data file.",
    "Error: Not a number value detected.",
    "Error: An error occurred while reading the input file.",
};

void pfc_fail(int i) {
    if (pfc_rank == 0 && i >= 0 && i <= 40) {
        printf("%s\n", kFail[i]);
        fflush(stdout);
    }
    exit(i);
}

/* ------------------------------------------------------------ tokens */

static int count_tokens(const char* s) {
    int n = 0;
    while (*s) {
        while (*s && isspace((unsigned char)*s)) s++;
        if (!*s) break;
        n++;
        while (*s && !isspace((unsigned char)*s)) s++;
    }
    return n;
}

static void read_line(FILE* in, char* buf) {
    if (fgets(buf, PFC_LINE, in) == NULL) pfc_fail(40);
}

/* Skips tokens until `label`; fails with `code` at end of file. */
static void seek_label(FILE* in, const char* label, int code) {
    char tok[PFC_LINE];
    for (;;) {
        if (fscanf(in, " %8191s", tok) != 1) pfc_fail(code);
        if (strcmp(tok, label) == 0) return;
    }
}

/* Reads count doubles from successive whitespace-separated positions of
 * *ps (sscanf " %lf%n"), advancing *ps. */
static int scan_doubles(const char** ps, double* out, int count) {
    for (int i = 0; i < count; i++) {
        int off = 0;
        if (sscanf(*ps, " %lf%n", &out[i], &off) != 1) return -1;
        *ps += off;
    }
    return 0;
}

/* ------------------------------------------------------------- blocks */

/* "I" block (src/pfc-relax.c:442-538) */
static void parse_io_block(FILE* in, run_cfg* cfg) {
    char line[PFC_LINE], tok[PFC_LINE];
    const char* ps;
    int off;

    /* input density files, on the rest of the label line */
    read_line(in, line);
    int U = count_tokens(line);
    if (U > PFC_MAX_FIELDS) U = PFC_MAX_FIELDS + 1;   /* rejected by the engine later */
    cfg->U = U;
    ps = line;
    int Nx = -1, Ny = -1;
    for (int u = 0; u < U && u < PFC_MAX_FIELDS; u++) {
        if (sscanf(ps, "%8191s%n", cfg->innames[u], &off) != 1) pfc_fail(2);
        ps += off;
        FILE* f = fopen(cfg->innames[u], "r");
        if (!f) pfc_fail(3);
        int nx, ny;
        double tmp;
        if (fscanf(f, "%d %d %lf %lf", &nx, &ny, &tmp, &tmp) != 4) pfc_fail(4);
        if ((Nx > 0 && Nx != nx) || (Ny > 0 && Ny != ny)) pfc_fail(5);
        Nx = nx;
        Ny = ny;
        fclose(f);
    }
    cfg->Nx = Nx;
    cfg->Ny = Ny;

    /* output density files, optional energy file */
    read_line(in, line);
    const int V = count_tokens(line);
    if (V < U || V > U + 1) pfc_fail(6);
    ps = line;
    for (int u = 0; u < U && u < PFC_MAX_FIELDS; u++) {
        if (sscanf(ps, "%8191s%n", cfg->outnames[u], &off) != 1) pfc_fail(7);
        ps += off;
    }
    if (V > U) {
        if (sscanf(ps, "%8191s", cfg->fname) != 1) pfc_fail(8);
        cfg->has_fname = 1;
    }

    /* print/write parameters: "Tprint Twrite pprint pwrite" or
     * "Tprint [ t1 ... tM ] pprint pwrite" */
    cfg->Twrite = -1;
    read_line(in, line);
    const int N = count_tokens(line);
    if (N == 4) {
        if (sscanf(line, " %d %d %d %d", &cfg->Tprint, &cfg->Twrite, &cfg->pprint,
                   &cfg->pwrite) != 4)
            pfc_fail(9);
    } else if (N >= 6) {
        const int M = N - 5;
        free(cfg->twrite);
        cfg->twrite = (int*)malloc(sizeof(int) * (size_t)M);
        cfg->ntwrite = M;
        ps = line;
        if (sscanf(ps, "%d%n", &cfg->Tprint, &off) != 1) pfc_fail(9);
        ps += off;
        if (sscanf(ps, "%8191s%n", tok, &off) != 1) pfc_fail(9);
        ps += off;
        if (strcmp(tok, "[") != 0) pfc_fail(9);
        for (int i = 0; i < M; i++) {
            if (sscanf(ps, "%d%n", &cfg->twrite[i], &off) != 1) pfc_fail(9);
            ps += off;
        }
        for (int i = 1; i < M; i++)
            if (cfg->twrite[i] <= cfg->twrite[i - 1]) pfc_fail(10);
        if (sscanf(ps, "%8191s%n", tok, &off) != 1) pfc_fail(9);
        ps += off;
        if (strcmp(tok, "]") != 0) pfc_fail(9);
        if (sscanf(ps, "%d %d", &cfg->pprint, &cfg->pwrite) != 2) pfc_fail(9);
    } else {
        pfc_fail(9);
    }
}

/* "M" block (src/pfc-relax.c:541-704).  Labels come in a fixed order; the
 * values of field u start on the label's line for u = 0. */
static void parse_model_block(FILE* in, run_cfg* cfg) {
    char line[PFC_LINE];
    const char* ps;
    pfc_params* p = &cfg->p;
    int V;
    read_line(in, line);
    if (sscanf(line, "%d", &V) != 1) pfc_fail(11);
    if (cfg->U != V) pfc_fail(12);
    const int U = V;
    memset(p, 0, sizeof(*p));
    p->U = U;
    const int Uc = U < PFC_MAX_FIELDS ? U : PFC_MAX_FIELDS;

    /* lambda: one value per field */
    seek_label(in, "l", 13);
    for (int u = 0; u < U; u++) {
        read_line(in, line);
        double v;
        if (sscanf(line, " %lf", &v) != 1) pfc_fail(14);
        if (u < Uc) p->lambda[u] = v;
    }
    /* alpha, beta, gamma: upper triangle including the diagonal */
    static const struct {
        const char* label;
        int miss, bad;
    } tri[3] = {{"a", 15, 16}, {"b", 17, 18}, {"g", 19, 20}};
    for (int k = 0; k < 3; k++) {
        double(*dst)[PFC_MAX_FIELDS] = (k == 0) ? p->alpha : (k == 1) ? p->beta : p->gamma;
        seek_label(in, tri[k].label, tri[k].miss);
        for (int u = 0; u < U; u++) {
            read_line(in, line);
            ps = line;
            double row[PFC_MAX_FIELDS + 1];
            const int cnt = U - u;
            double* tmp = cnt <= PFC_MAX_FIELDS ? row : (double*)malloc(sizeof(double) * cnt);
            if (scan_doubles(&ps, tmp, cnt)) pfc_fail(tri[k].bad);
            for (int v = u; v < Uc && u < Uc; v++) dst[u][v] = tmp[v - u];
            if (tmp != row) free(tmp);
        }
    }
    /* delta: one value per field */
    seek_label(in, "d", 21);
    for (int u = 0; u < U; u++) {
        read_line(in, line);
        double v;
        if (sscanf(line, " %lf", &v) != 1) pfc_fail(22);
        if (u < Uc) p->delta[u] = v;
    }
    /* epsilon: strict upper triangle */
    seek_label(in, "e", 23);
    for (int u = 0; u < U - 1; u++) {
        read_line(in, line);
        ps = line;
        for (int v = u + 1; v < U; v++) {
            double x;
            if (scan_doubles(&ps, &x, 1)) pfc_fail(24);
            if (u < Uc && v < Uc) p->epsilon[u][v] = x;
        }
    }
    /* sigma */
    seek_label(in, "s", 25);
    read_line(in, line);
    if (sscanf(line, " %lf", &p->sigma) != 1) pfc_fail(26);
    /* 1/nu: the number of values on a field's line is its number of modes */
    seek_label(in, "n", 27);
    int W[PFC_MAX_FIELDS];
    for (int u = 0; u < U; u++) {
        read_line(in, line);
        const int Wu = count_tokens(line);
        ps = line;
        for (int w = 0; w < Wu; w++) {
            double x;
            int off = 0;
            if (sscanf(ps, "%lf%n", &x, &off) != 1) pfc_fail(28);
            ps += off;
            if (u < Uc && w < PFC_MAX_MODES) p->nu2[u][w] = 1.0 / x / x;
        }
        if (u < Uc) W[u] = Wu;
        if (u < Uc) p->W[u] = Wu;
    }
    /* omega and phi: W[u] values per field */
    static const struct {
        const char* label;
        int miss, bad;
    } modes[2] = {{"o", 29, 30}, {"p", 31, 32}};
    for (int k = 0; k < 2; k++) {
        double(*dst)[PFC_MAX_MODES] = (k == 0) ? p->omega : p->phi;
        seek_label(in, modes[k].label, modes[k].miss);
        for (int u = 0; u < U; u++) {
            read_line(in, line);
            ps = line;
            const int Wu = (u < Uc) ? W[u] : 0;
            for (int w = 0; w < Wu; w++) {
                double x;
                int off = 0;
                if (sscanf(ps, "%lf%n", &x, &off) != 1) pfc_fail(modes[k].bad);
                ps += off;
                if (w < PFC_MAX_MODES) dst[u][w] = x;
            }
        }
    }
    /* 1/mu: strict upper triangle; 0 where beta_uv is 0 */
    seek_label(in, "m", 33);
    for (int u = 0; u < U - 1; u++) {
        read_line(in, line);
        ps = line;
        for (int v = u + 1; v < U; v++) {
            double x;
            if (scan_doubles(&ps, &x, 1)) pfc_fail(34);
            if (u < Uc && v < Uc) p->mu2[u][v] = (p->beta[u][v] == 0.0) ? 0.0 : 1.0 / x / x;
        }
    }
}

/* "R" line (src/pfc-relax.c:707-718) */
static void parse_relax_line(FILE* in, run_cfg* cfg) {
    char line[PFC_LINE];
    read_line(in, line);
    if (sscanf(line, "%lf %d %lf %lf %lf %d %d %d", &cfg->c, &cfg->tend, &cfg->dx, &cfg->dy,
               &cfg->dt, &cfg->Topt, &cfg->xopt, &cfg->yopt) != 8)
        pfc_fail(35);
    if (cfg->Twrite > 0) {
        const int n = cfg->tend / cfg->Twrite + 1;
        free(cfg->twrite);
        cfg->twrite = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        cfg->ntwrite = n > 0 ? n : 0;
        for (int i = 0; i < n; i++) cfg->twrite[i] = cfg->Twrite * i;
    }
}

void pfc_read_input(run_cfg* cfg) {
    FILE* in = fopen(cfg->input, "r");
    if (!in) pfc_fail(0);
    char tok[PFC_LINE];
    while (fscanf(in, " %8191s", tok) > 0) {
        if (tok[0] == '#') {
            char rest[PFC_LINE];
            read_line(in, rest);              /* comment: rest of the line */
        } else if (strlen(tok) > 1) {
            pfc_fail(1);
        } else if (tok[0] == 'I') {
            parse_io_block(in, cfg);
        } else if (tok[0] == 'M') {
            parse_model_block(in, cfg);
        } else if (tok[0] == 'R') {
            parse_relax_line(in, cfg);
        } else {
            pfc_fail(36);
        }
    }
    fclose(in);
    cfg->p.c = cfg->c;
    cfg->p.dt = cfg->dt;
}

/* ---------------------------------------------------------- density IO */

static char* slurp(const char* path, size_t* len) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    size_t cap = 1 << 20, n = 0;
    char* buf = (char*)malloc(cap + 1);
    for (;;) {
        if (n == cap) {
            cap *= 2;
            buf = (char*)realloc(buf, cap + 1);
        }
        size_t r = fread(buf + n, 1, cap - n, f);
        n += r;
        if (r == 0) break;
    }
    fclose(f);
    buf[n] = '\0';
    *len = n;
    return buf;
}

/* Parses the next number like fscanf(" %lf"): skips whitespace, converts
 * with strtod (the conversion scanf uses).  Returns 0 on success. */
static int next_double(char** ps, double* out) {
    char* s = *ps;
    while (*s && isspace((unsigned char)*s)) s++;
    if (!*s) return -1;
    char* end;
    errno = 0;
    const double v = strtod(s, &end);
    if (end == s) return -1;
    *out = v;
    *ps = end;
    return 0;
}

void pfc_read_data(const run_cfg* cfg, double* fields) {
    const size_t n = (size_t)cfg->Nx * cfg->Ny;
    for (int u = 0; u < cfg->U; u++) {
        size_t len;
        char* buf = slurp(cfg->innames[u], &len);
        if (!buf) pfc_fail(37);
        char* s = buf;
        double h[4];
        for (int i = 0; i < 4; i++) {
            /* the header's first two entries are %d conversions */
            if (i < 2) {
                while (*s && isspace((unsigned char)*s)) s++;
                char* end;
                strtol(s, &end, 10);
                if (end == s) pfc_fail(4);
                s = end;
            } else if (next_double(&s, &h[i])) {
                pfc_fail(4);
            }
        }
        double* dst = fields + (size_t)u * n;
        for (size_t i = 0; i < n; i++)
            if (next_double(&s, &dst[i])) pfc_fail(38);
        free(buf);
    }
}

/* ------------------------------------------------------------- output */

void pfc_step_name(const char* pattern, int t, char* out, size_t n) {
    const char* x = strchr(pattern, '#');
    if (!x) {
        snprintf(out, n, "%s", pattern);
        return;
    }
    snprintf(out, n, "%.*s%d%s", (int)(x - pattern), pattern, t, x + 1);
}

typedef struct {
    const double* v;
    size_t i0, i1;
    int prec;
    char* buf;
    size_t len;
} fmt_job;

static void* fmt_worker(void* arg) {
    fmt_job* j = (fmt_job*)arg;
    const size_t cap = (j->i1 - j->i0) * (size_t)(j->prec + 16) + 64;
    j->buf = (char*)malloc(cap);
    size_t n = 0;
    for (size_t i = j->i0; i < j->i1; i++) n += (size_t)sprintf(j->buf + n, "%.*e\n", j->prec, j->v[i]);
    j->len = n;
    return NULL;
}

int pfc_write_field(const char* path, const double* f, int Nx, int Ny, double dx, double dy,
                    int pwrite) {
    return pfc_write_field_prec(path, f, Nx, Ny, dx, dy, pwrite, pwrite);
}

int pfc_write_field_prec(const char* path, const double* f, int Nx, int Ny, double dx, double dy,
                         int phead, int pvalue) {
    FILE* out = fopen(path, "w");
    if (!out) return -1;
    fprintf(out, "%d %d %.*lf %.*lf\n", Nx, Ny, phead, dx, phead, dy);
    pfc_write_values(out, f, (size_t)Nx * Ny, pvalue);
    return fclose(out) == 0 ? 0 : -1;
}

void pfc_write_values(FILE* out, const double* f, size_t n, int pvalue) {
    const int prec = pvalue < 0 ? 6 : (pvalue > 300 ? 300 : pvalue);
    long nth = sysconf(_SC_NPROCESSORS_ONLN);
    if (nth < 1) nth = 1;
    if (nth > 16) nth = 16;
    if (n < 65536) nth = 1;
    /* format in slices of at most 1M values so memory stays bounded */
    const size_t slice = (size_t)1 << 20;
    fmt_job jobs[16];
    pthread_t th[16];
    for (size_t base = 0; base < n; base += slice) {
        const size_t end = base + slice < n ? base + slice : n;
        const size_t per = (end - base + nth - 1) / nth;
        int nj = 0;
        for (long k = 0; k < nth; k++) {
            const size_t a = base + k * per;
            if (a >= end) break;
            jobs[nj] = (fmt_job){f, a, a + per < end ? a + per : end, prec, NULL, 0};
            if (nth > 1) pthread_create(&th[nj], NULL, fmt_worker, &jobs[nj]);
            else fmt_worker(&jobs[nj]);
            nj++;
        }
        for (int k = 0; k < nj; k++) {
            if (nth > 1) pthread_join(th[k], NULL);
            fwrite(jobs[k].buf, 1, jobs[k].len, out);
            free(jobs[k].buf);
        }
    }
}

void pfc_print_output(const run_cfg* cfg, int t, int elapsed, double dx, double dy, double fave,
                      const double* nave) {
    if (pfc_rank != 0) return;
    const int pp = cfg->pprint;
    size_t cap = 256 + (size_t)(cfg->U + 3) * (size_t)(pp + 340);
    char* line = (char*)malloc(cap);
    size_t n = (size_t)snprintf(line, cap, "%d %d %.*lf %.*lf %.*lf ", t, elapsed, pp, dx, pp, dy,
                                pp, fave);
    for (int u = 0; u < cfg->U; u++) n += (size_t)snprintf(line + n, cap - n, "%.*lf ", pp, nave[u]);
    printf("%s\n", line);
    fflush(stdout);
    /* ".in" suffix replaced by ".out", else ".out" appended */
    char name[PFC_LINE + 8];
    const size_t L = strlen(cfg->input);
    if (L >= 3 && strcmp(cfg->input + L - 3, ".in") == 0)
        snprintf(name, sizeof name, "%.*s.out", (int)(L - 3), cfg->input);
    else
        snprintf(name, sizeof name, "%s.out", cfg->input);
    FILE* out = fopen(name, t == 0 ? "w" : "a");
    if (out) {
        fprintf(out, "%s\n", line);
        fclose(out);
    }
    free(line);
}

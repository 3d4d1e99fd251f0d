/*
 * pfc_init_main.c -- the pfc-init command: builds an initial density file
 * from a labelled input file, the job of the reference's pfc-init.jar
 * (src/PFCInit.java:96-698), which the demo pipelines run before pfc-relax
 * (hetero.sh:14-16).
 *
 *   pfc-init X.i
 *
 * Labels, in any order after D: D Nx Ny dx dy (grid); S seed (java.util.Random
 * seed); L name type i0 j0 dNx dNy i1 j1 (load part of a saved field); V N and
 * N grains (Voronoi); C (circular grain); P N corners (convex polygon grain);
 * T nave famp (rescale); O name precision (write).  A grain line is
 * x y type, sym a nave amp namp, tx1 ty1 theta xr yr tx2 ty2 (C: x y R type;
 * P: type).  Type 0 overwrites, 1 superposes, 2 leaves the field unchanged.
 * '#' starts a comment.  Messages and exit codes follow the reference.
 *
 * Output: "Nx Ny" and dx, dy in %.{p}e, then one %.{p}e value per line, rows
 * in order.  Values are formatted by C's printf, which rounds the exact binary
 * value; Java's formatter rounds half-up from the shortest decimal, so the two
 * can differ in the last printed digit when that decimal ends in 5 there.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pfc_io.h"
#include "pfc_jrand.h"

int pfc_rank = 0;

static const double kPi = 3.14159265358979323846;

typedef struct {
    char** w;
    int n, cap, k;
} tokens;

static void die(const char* msg) {
    printf("%s\n", msg);
    exit(1);
}

static const char* next_word(tokens* t) {
    if (t->k >= t->n) die("Error: Invalid input file. Unexpected end of input.");
    return t->w[t->k++];
}

/* Integer.parseInt: optional sign and decimal digits, in int range */
static int s2i(const char* s) {
    char* end;
    const long long v = strtoll(s, &end, 10);
    const int ok = *s && *end == 0 && (s[0] == '-' || s[0] == '+' || (s[0] >= '0' && s[0] <= '9')) &&
                   v >= -2147483648LL && v <= 2147483647LL;
    if (!ok) {
        printf("Error: Invalid input file. Integer expected instead of \"%s\".\n", s);
        exit(1);
    }
    return (int)v;
}

static double s2d(const char* s) {
    char* end;
    const double v = strtod(s, &end);
    /* Double.parseDouble also takes a trailing type letter: 1.5d, 2f */
    const int suffix = end != s && (*end == 'd' || *end == 'D' || *end == 'f' || *end == 'F') &&
                       end[1] == 0 && ((end[-1] >= '0' && end[-1] <= '9') || end[-1] == '.');
    if (end == s || (*end && !suffix)) {
        printf("Error: Invalid input file. Decimal number expected instead of \"%s\".\n", s);
        exit(1);
    }
    return v;
}

static int wi(int i, int N) {
    while (i < 0) i += N;
    while (i >= N) i -= N;
    return i;
}

static double wd(double d, double L) {
    while (d < 0) d += L;
    while (d >= L) d -= L;
    return d;
}

/* one-mode lattice ansatz (src/PFCInit.java:44-64) */
static double oma(double x, double y, int sym, double a) {
    const double q = 2.0 * kPi / a;
    if (sym == 3) {
        const double qx = q, qy = q / sqrt(3.0);
        return cos(qx * x) * cos(qy * y) + 0.5 * cos(2.0 * qy * y);
    }
    if (sym == 4) return cos(q * x) * cos(q * y);
    if (sym == 6) {
        const double qx = q, qy = q / sqrt(3.0);
        return -2.0 * cos(qx * x - qy * y) - 2.0 * cos(2.0 * qy * y) - 2.0 * cos(-qx * x - qy * y) +
               cos(-qx * x - 3.0 * qy * y) + cos(2.0 * qx * x) + cos(-qx * x + 3.0 * qy * y);
    }
    die("Error: Invalid lattice symmetry specified. Only triangular, square and hexagonal "
        "symmetries are supported.");
    return 0.0;
}

/* nearest seed under periodic minimum image, first index on ties
 * (src/PFCInit.java:66-93) */
static int closest(double u, double v, const double* x, const double* y, int N, double W, double H) {
    int n = 0;
    double d2 = 0.0;
    for (int m = 0; m < N; m++) {
        double dx = u - x[m];
        if (dx > 0.5 * W) dx -= W;
        else if (dx < -0.5 * W) dx += W;
        double dy = v - y[m];
        if (dy > 0.5 * H) dy -= H;
        else if (dy < -0.5 * H) dy += H;
        const double d = dx * dx + dy * dy;
        if (m == 0 || d < d2) {
            n = m;
            d2 = d;
        }
    }
    return n;
}

typedef struct {
    int sym;
    double a, nave, amp, namp, tx1, ty1, theta, xr, yr, tx2, ty2;
} lattice;

static void read_lattice(tokens* t, lattice* L) {
    L->sym = s2i(next_word(t));
    L->a = s2d(next_word(t));
    L->nave = s2d(next_word(t));
    L->amp = s2d(next_word(t));
    L->namp = s2d(next_word(t));
    L->tx1 = s2d(next_word(t));
    L->ty1 = s2d(next_word(t));
    L->theta = s2d(next_word(t)) / 180.0 * kPi;
    L->xr = s2d(next_word(t));
    L->yr = s2d(next_word(t));
    L->tx2 = s2d(next_word(t));
    L->ty2 = s2d(next_word(t));
}

/* new value of a grid point that held `base` (type * old value): lattice
 * coordinates (x_, y_), already shifted by tx2 + xr and ty2 + yr, are rotated
 * and translated, then the ansatz and one uniform noise draw are added.  The
 * sums associate left to right as in src/PFCInit.java:461, 527, 641, so the
 * doubles match the reference's (built without contraction to FMA). */
static double grain_value(double base, const lattice* L, double x_, double y_, double sine,
                          double cosine, pfc_jrand* rng) {
    const double rx = x_ * cosine - y_ * sine + L->xr - L->tx1;
    const double ry = x_ * sine + y_ * cosine + L->yr - L->ty1;
    return base + L->nave + L->amp * oma(rx, ry, L->sym, L->a) +
           2.0 * L->namp * (pfc_jrand_next_double(rng) - 0.5);
}

static void tokenize(const char* name, tokens* t) {
    FILE* f = fopen(name, "r");
    char* line = NULL;
    size_t cap = 0;
    while (getline(&line, &cap, f) >= 0) {
        char* save = NULL;
        for (char* w = strtok_r(line, " \t\r\n\f\v", &save); w; w = strtok_r(NULL, " \t\r\n\f\v", &save)) {
            if (w[0] == '#') break;   /* rest of the line is a comment */
            if (t->n == t->cap) {
                t->cap = t->cap ? 2 * t->cap : 256;
                t->w = (char**)realloc(t->w, sizeof(char*) * (size_t)t->cap);
            }
            t->w[t->n++] = strdup(w);
        }
    }
    free(line);
    fclose(f);
}

/* every label after D works on the field, which D allocates */
static void need_field(const double* arr) {
    if (!arr) die("Error: Invalid input file. System dimensions (D) expected first.");
}

int main(int argc, char** argv) {
    if (argc != 2) {
        printf("This is a tool for constructing 2D density fields for PFC calculations. It expects "
               "one argument, an input file of labelled steps: D Nx Ny dx dy (dimensions and "
               "discretization, first), S seed (java.util.Random seed), C / P / V (circular, "
               "convex polygon and Voronoi grains), L (load part of a saved field), T nave famp "
               "(transform) and O filename precision (output). '#' starts a comment.\n");
        return 0;
    }
    const char* name = argv[1];
    FILE* probe = fopen(name, "r");
    if (!probe) {
        printf("Error: Input file \"%s\" not found.\n", name);
        return 1;
    }
    fclose(probe);
    tokens t = {NULL, 0, 0, 0};
    tokenize(name, &t);

    pfc_jrand rng;
    pfc_jrand_seed(&rng, (int64_t)time(NULL) ^ (int64_t)clock() * 7919);   /* unseeded: varies */
    int Nx = -1, Ny = -1;
    double dx = 0.0, dy = 0.0;
    double* arr = NULL;
#define A(jj, ii) arr[(size_t)(jj) * Nx + (ii)]
    while (t.k < t.n) {
        const char* lab = next_word(&t);
        if (!strcmp(lab, "D")) {
            Nx = s2i(next_word(&t));
            Ny = s2i(next_word(&t));
            dx = s2d(next_word(&t));
            dy = s2d(next_word(&t));
            free(arr);
            arr = (double*)calloc((size_t)(Nx > 0 ? Nx : 0) * (size_t)(Ny > 0 ? Ny : 0) + 1,
                                  sizeof(double));
        } else if (!strcmp(lab, "S")) {
            pfc_jrand_seed(&rng, (int64_t)s2i(next_word(&t)));
        } else if (!strcmp(lab, "L")) {
            need_field(arr);
            const char* fname = next_word(&t);
            const int type = s2i(next_word(&t));
            if (type == 2) {
                printf("Warning: Phantom type used for the system from data file \"%s\".\n", fname);
                continue;
            }
            if (type < 0 || type > 2)
                printf("Warning: Invalid type used for the system from data file \"%s\". "
                       "Defaulting to overwrite.\n", fname);
            FILE* f = fopen(fname, "r");
            if (!f) {
                printf("Error: Data file \"%s\" not found.\n", fname);
                return 1;
            }
            char hdr[1024];
            if (!fgets(hdr, sizeof hdr, f)) die("Error: Invalid input file. Integer expected instead of \"\".");
            char w0[256] = "", w1[256] = "", w2[256] = "", w3[256] = "";
            sscanf(hdr, "%255s %255s %255s %255s", w0, w1, w2, w3);
            const int Nx2 = s2i(w0), Ny2 = s2i(w1);
            const double dx2 = s2d(w2), dy2 = s2d(w3);
            if (dx2 != dx || dy2 != dy)
                printf("Warning: System from data file \"%s\" has incompatible discretization.\n", fname);
            const int i0 = wi(s2i(next_word(&t)), Nx2);
            const int j0 = wi(s2i(next_word(&t)), Ny2);
            const int dNx = s2i(next_word(&t));
            const int dNy = s2i(next_word(&t));
            const int i1 = wi(s2i(next_word(&t)), Nx);
            const int j1 = wi(s2i(next_word(&t)), Ny);
            double* arr2 = (double*)malloc(sizeof(double) * (size_t)Nx2 * Ny2);
            char vbuf[512];
            for (size_t n = 0; n < (size_t)Nx2 * Ny2; n++) {
                if (!fgets(vbuf, sizeof vbuf, f)) die("Error: Invalid input file. Decimal number expected instead of \"\".");
                vbuf[strcspn(vbuf, "\r\n")] = 0;
                arr2[n] = s2d(vbuf);
            }
            fclose(f);
            /* an invalid type is announced as overwrite but still multiplies
             * the old value, as src/PFCInit.java:339-341, 375 do */
            const int tp = type;
            for (int j = 0; j < dNy; j++) {
                const int js = wi(j0 + j, Ny2), jt = wi(j1 + j, Ny);
                for (int i = 0; i < dNx; i++) {
                    const int is = wi(i0 + i, Nx2), it = wi(i1 + i, Nx);
                    A(jt, it) = tp * A(jt, it) + arr2[(size_t)js * Nx2 + is];
                }
            }
            free(arr2);
        } else if (!strcmp(lab, "V")) {
            need_field(arr);
            const int N = s2i(next_word(&t));
            if (N < 1) die("Error: Invalid input file. At least one Voronoi grain expected.");
            double* x = (double*)malloc(sizeof(double) * (size_t)N);
            double* y = (double*)malloc(sizeof(double) * (size_t)N);
            int* type = (int*)malloc(sizeof(int) * (size_t)N);
            lattice* lat = (lattice*)malloc(sizeof(lattice) * (size_t)N);
            const double W = Nx * dx, H = Ny * dy;
            for (int n = 0; n < N; n++) {
                x[n] = wd(s2d(next_word(&t)), W);
                y[n] = wd(s2d(next_word(&t)), H);
                type[n] = s2i(next_word(&t));
                read_lattice(&t, &lat[n]);
            }
            for (int j = 0; j < Ny; j++) {
                for (int i = 0; i < Nx; i++) {
                    const int n = closest(i * dx, j * dy, x, y, N, W, H);
                    if (type[n] == 2) continue;
                    if (type[n] < 0 || type[n] > 2) {
                        printf("Warning: Invalid type used for a Voronoi grain. Defaulting to overwrite.\n");
                        type[n] = 0;
                    }
                    const lattice* L = &lat[n];
                    const double u = i * dx - L->tx2 - L->xr;
                    const double v = j * dy - L->ty2 - L->yr;
                    A(j, i) = grain_value(type[n] * A(j, i), L, u, v, sin(-L->theta),
                                          cos(-L->theta), &rng);
                }
            }
            free(x);
            free(y);
            free(type);
            free(lat);
        } else if (!strcmp(lab, "C")) {
            need_field(arr);
            const double xc = s2d(next_word(&t)), yc = s2d(next_word(&t)), R = s2d(next_word(&t));
            int type = s2i(next_word(&t));
            if (type == 2) {
                printf("Warning: Phantom type used for a circular grain.\n");
                continue;   /* the grain's remaining words are read as labels, as in the reference */
            }
            if (type < 0 || type > 2) {
                printf("Warning: Invalid type used for a circular grain. Defaulting to overwrite.\n");
                type = 0;
            }
            lattice L;
            read_lattice(&t, &L);
            const double R2 = R * R, sine = sin(-L.theta), cosine = cos(-L.theta);
            const int i1 = (int)((xc - R) / dx), i2 = (int)((xc + R) / dx);
            const int j1 = (int)((yc - R) / dy), j2 = (int)((yc + R) / dy);
            for (int j = j1; j <= j2; j++) {
                const double v = j * dy - yc, y_ = j * dy - L.ty2 - L.yr;
                const int wj = wi(j, Ny);
                for (int i = i1; i <= i2; i++) {
                    const double u = i * dx - xc;
                    const int wx = wi(i, Nx);
                    if (u * u + v * v < R2) {
                        const double x_ = i * dx - L.tx2 - L.xr;
                        A(wj, wx) = grain_value(type * A(wj, wx), &L, x_, y_, sine, cosine, &rng);
                    }
                }
            }
        } else if (!strcmp(lab, "P")) {
            need_field(arr);
            const int N = s2i(next_word(&t));
            if (N < 2) die("Error: Invalid input file. At least two polygon corners expected.");
            double* xc = (double*)malloc(sizeof(double) * (size_t)N);
            double* yc = (double*)malloc(sizeof(double) * (size_t)N);
            for (int n = 0; n < N; n++) {
                xc[n] = s2d(next_word(&t));
                yc[n] = s2d(next_word(&t));
            }
            double xmin = xc[0], xmax = xc[0], ymin = yc[0], ymax = yc[0];
            double ax = xc[0] - xc[N - 1], ay = yc[0] - yc[N - 1];
            double bx = xc[1] - xc[0], by = yc[1] - yc[0];
            double z1 = ax * by - ay * bx;
            for (int n = 1; n < N; n++) {
                if (xc[n] < xmin) xmin = xc[n];
                if (xc[n] > xmax) xmax = xc[n];
                if (yc[n] < ymin) ymin = yc[n];
                if (yc[n] > ymax) ymax = yc[n];
                ax = xc[n] - xc[(n - 1 + N) % N];
                ay = yc[n] - yc[(n - 1 + N) % N];
                bx = xc[(n + 1) % N] - xc[n];
                by = yc[(n + 1) % N] - yc[n];
                const double z2 = ax * by - ay * bx;
                if (z1 * z2 < 0.0) die("Error: Concave grain detected. Convex grain expected.");
                z1 = z2;
            }
            int type = s2i(next_word(&t));
            if (type == 2) {
                printf("Warning: Phantom type used for a polygon grain.\n");
                free(xc);
                free(yc);
                continue;
            }
            if (type < 0 || type > 2) {
                printf("Warning: Invalid type used for a polygon grain. Defaulting to overwrite.\n");
                type = 0;
            }
            lattice L;
            read_lattice(&t, &L);
            const double sine = sin(-L.theta), cosine = cos(-L.theta);
            const int i1 = (int)(xmin / dx), i2 = (int)(xmax / dx);
            const int j1 = (int)(ymin / dy), j2 = (int)(ymax / dy);
            for (int j = j1; j <= j2; j++) {
                const int wj = wi(j, Ny);
                const double yy = j * dy, y_ = yy - L.ty2 - L.yr;
                for (int i = i1; i <= i2; i++) {
                    const int wx = wi(i, Nx);
                    const double xx = i * dx, x_ = xx - L.tx2 - L.xr;
                    double z = (xc[0] - xc[N - 1]) * (yy - yc[N - 1]) - (yc[0] - yc[N - 1]) * (xx - xc[N - 1]);
                    int inside = 1;
                    for (int n = 1; n < N; n++) {
                        const double zz = (xc[n] - xc[n - 1]) * (yy - yc[n]) - (yc[n] - yc[n - 1]) * (xx - xc[n]);
                        if (z * zz < 0.0) {
                            inside = 0;
                            break;
                        }
                        z = zz;
                    }
                    if (inside) A(wj, wx) = grain_value(type * A(wj, wx), &L, x_, y_, sine, cosine, &rng);
                }
            }
            free(xc);
            free(yc);
        } else if (!strcmp(lab, "T")) {
            need_field(arr);
            const double nave = s2d(next_word(&t)), famp = s2d(next_word(&t));
            double nave0 = 0.0;
            for (size_t n = 0; n < (size_t)Nx * Ny; n++) nave0 += arr[n];
            nave0 *= 1.0 / Nx / Ny;
            for (size_t n = 0; n < (size_t)Nx * Ny; n++) arr[n] = (arr[n] - nave0) * famp + nave;
        } else if (!strcmp(lab, "O")) {
            need_field(arr);
            const char* oname = next_word(&t);
            const int prec = s2i(next_word(&t));
            FILE* out = fopen(oname, "w");
            if (!out) {
                printf("Error: could not write \"%s\".\n", oname);
                return 1;
            }
            fprintf(out, "%d %d %.*e %.*e\n", Nx, Ny, prec, dx, prec, dy);
            pfc_write_values(out, arr, (size_t)Nx * Ny, prec);
            fclose(out);
        } else {
            printf("Error: Invalid input file. Label expected instead of \"%s\".\n", lab);
            return 1;
        }
    }
#undef A
    free(arr);
    return 0;
}

/*
 * pfc_coord.c -- the coordinator's algorithm (src/Coordinator.java:29-318),
 * in the reference's loop order so that rings and atoms get the same indices.
 * Arithmetic is plain IEEE double with no contraction (Java semantics); build
 * with -ffp-contract=off.
 */
#include "pfc_coord.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* neighbour offsets of the minimum test (src/Coordinator.java:13-14) */
static const int kU[8] = {1, 1, 0, -1, -1, -1, 0, 1};
static const int kV[8] = {0, 1, 1, 1, 0, -1, -1, -1};

typedef struct {
    int* v;
    int n, cap;
} ilist;

static int il_push(ilist* l, int x) {
    if (l->n == l->cap) {
        int cap = l->cap ? 2 * l->cap : 4;
        int* v = (int*)realloc(l->v, sizeof(int) * (size_t)cap);
        if (!v) return -1;
        l->v = v;
        l->cap = cap;
    }
    l->v[l->n++] = x;
    return 0;
}

static int il_has(const ilist* l, int x) {
    for (int i = 0; i < l->n; i++)
        if (l->v[i] == x) return 1;
    return 0;
}

/* an atom or a ring (Atom.java): bros are same-kind neighbours, hos the
 * other kind */
typedef struct {
    double r[2];
    ilist bros, hos;
} node;

typedef struct {
    node* v;
    int n, cap;
} nlist;

static int nl_push(nlist* l, double x, double y) {
    if (l->n == l->cap) {
        int cap = l->cap ? 2 * l->cap : 1024;
        node* v = (node*)realloc(l->v, sizeof(node) * (size_t)cap);
        if (!v) return -1;
        l->v = v;
        l->cap = cap;
    }
    node* p = &l->v[l->n++];
    memset(p, 0, sizeof(*p));
    p->r[0] = x;
    p->r[1] = y;
    return 0;
}

struct pfc_coord {
    int Nx, Ny, Bx, By;
    double w, h, lA, ux, uy;
    nlist rings, atoms;
    ilist* bins;   /* By x Bx lists of ring indices */
    int flags, bad_ring;
    int order[4], norder;   /* check flags in the order the reference prints them */
};

/* Java's (int) of a double: truncation, NaN -> 0, saturation */
static int j2i(double d) {
    if (d != d) return 0;
    if (d >= 2147483647.0) return 2147483647;
    if (d <= -2147483648.0) return (-2147483647 - 1);
    return (int)d;
}

static int wrapi(int x, int n) {
    if (x < 0) return x + n;
    if (x >= n) return x - n;
    return x;
}

static void wrapv(const pfc_coord* c, double* r) {
    if (r[0] < 0.0) r[0] += c->w;
    else if (r[0] >= c->w) r[0] -= c->w;
    if (r[1] < 0.0) r[1] += c->h;
    else if (r[1] >= c->h) r[1] -= c->h;
}

static void diff(const pfc_coord* c, const double* b, const double* a, double* ab) {
    ab[0] = b[0] - a[0];
    ab[1] = b[1] - a[1];
    if (ab[0] > 0.5 * c->w) ab[0] -= c->w;
    else if (ab[0] < -0.5 * c->w) ab[0] += c->w;
    if (ab[1] > 0.5 * c->h) ab[1] -= c->h;
    else if (ab[1] < -0.5 * c->h) ab[1] += c->h;
}

static double abs2(const double* a) { return a[0] * a[0] + a[1] * a[1]; }

/* bin indices of a position (src/Coordinator.java:96-104) */
static int binx(const pfc_coord* c, const double* r) { return j2i(r[0] / c->w * c->Bx); }
static int biny(const pfc_coord* c, const double* r) { return j2i(r[1] / c->h * c->By); }
static ilist* bin(pfc_coord* c, int nx, int ny) {
    if (nx < 0 || nx >= c->Bx || ny < 0 || ny >= c->By) return NULL;
    return &c->bins[(size_t)ny * c->Bx + nx];
}

/* src/Coordinator.java:123-143 */
static int bros_n_hos(pfc_coord* c, int n, int i, int j, int k) {
    node* A = c->atoms.v;
    node* R = c->rings.v;
    int e = 0;
    e |= il_push(&A[n].hos, i);
    e |= il_push(&A[n].hos, j);
    e |= il_push(&A[n].hos, k);
    e |= il_push(&R[i].hos, n);
    e |= il_push(&R[j].hos, n);
    e |= il_push(&R[k].hos, n);
    if (!il_has(&R[i].bros, j)) {
        e |= il_push(&R[i].bros, j);
        e |= il_push(&R[j].bros, i);
    }
    if (!il_has(&R[i].bros, k)) {
        e |= il_push(&R[i].bros, k);
        e |= il_push(&R[k].bros, i);
    }
    if (!il_has(&R[j].bros, k)) {
        e |= il_push(&R[j].bros, k);
        e |= il_push(&R[k].bros, j);
    }
    return e;
}

/* parabola through (x0,p0), (x1,p1), (x2,p2): vertex, as the reference
 * writes it (src/Coordinator.java:159-186) */
static double vertex(int i, double p0, double p1, double p2, double u) {
    const double x0 = i - 1, x1 = i, x2 = i + 1;
    const double x02 = x0 * x0, x12 = x1 * x1, x22 = x2 * x2;
    const double A = (x2 * (p1 - p0) + x1 * (p0 - p2) + x0 * (p2 - p1));
    const double B = (x22 * (p0 - p1) + x12 * (p2 - p0) + x02 * (p1 - p2));
    double x = -B / (2.0 * A) * u;
    if (fabs(A) < 1.0e-14) x = i;   /* the reference keeps the grid index here */
    return x;
}

static int find_rings(pfc_coord* c, double* psi) {
    const int Nx = c->Nx, Ny = c->Ny;
#define PSI(jj, ii) psi[(size_t)(jj) * Nx + (ii)]
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            int minimum = 1;
            for (int n = 0; n < 8; n++) {
                if (PSI(wrapi(j + kV[n], Ny), wrapi(i + kU[n], Nx)) < PSI(j, i)) {
                    minimum = 0;
                    break;
                }
            }
            if (!minimum) continue;
            double r[2];
            r[0] = vertex(i, PSI(j, wrapi(i - 1, Nx)), PSI(j, i), PSI(j, wrapi(i + 1, Nx)), c->ux);
            r[1] = vertex(j, PSI(wrapi(j - 1, Ny), i), PSI(j, i), PSI(wrapi(j + 1, Ny), i), c->uy);
            wrapv(c, r);
            const int k = c->rings.n;
            if (nl_push(&c->rings, r[0], r[1])) return -1;
            ilist* b = bin(c, binx(c, r), biny(c, r));
            if (!b) return -2;   /* the reference throws here */
            if (il_push(b, k)) return -1;
            /* make this minimum unique among equal neighbours */
            PSI(j, i) -= 2000.0;
            for (int n = 0; n < 8; n++) PSI(wrapi(j + kV[n], Ny), wrapi(i + kU[n], Nx)) -= 1000.0;
        }
    }
#undef PSI
    return 0;
}

/* triplets of mutually near rings whose circumcircle holds no other ring
 * (src/Coordinator.java:210-262) */
static int find_atoms(pfc_coord* c) {
    const double d2 = 2.0 * c->lA;
    for (int i = 0; i < c->rings.n; i++) {
        const double* ri = c->rings.v[i].r;
        const int nx = binx(c, ri), ny = biny(c, ri);
        for (int vj = -1; vj < 2; vj++) {
            for (int uj = -1; uj < 2; uj++) {
                const ilist* bj = bin(c, wrapi(nx + uj, c->Bx), wrapi(ny + vj, c->By));
                if (!bj) return -2;
                for (int jj = 0; jj < bj->n; jj++) {
                    const int j = bj->v[jj];
                    if (j <= i) continue;
                    double ij[2];
                    diff(c, c->rings.v[j].r, ri, ij);
                    if (sqrt(abs2(ij)) > d2) continue;
                    for (int vk = -1; vk < 2; vk++) {
                        for (int uk = -1; uk < 2; uk++) {
                            const ilist* bk = bin(c, wrapi(nx + uk, c->Bx), wrapi(ny + vk, c->By));
                            if (!bk) return -2;
                            for (int kk = 0; kk < bk->n; kk++) {
                                const int k = bk->v[kk];
                                if (k <= j) continue;
                                double ik[2], jk[2];
                                diff(c, c->rings.v[k].r, ri, ik);
                                if (sqrt(abs2(ik)) > d2) continue;
                                diff(c, c->rings.v[k].r, c->rings.v[j].r, jk);
                                if (sqrt(abs2(jk)) > d2) continue;
                                const double D = 2.0 * (ij[0] * ik[1] - ij[1] * ik[0]);
                                if (D == 0.0) continue;
                                const double iD = 1.0 / D;
                                const double it[2] = {
                                    (ik[1] * (ij[0] * ij[0] + ij[1] * ij[1]) -
                                     ij[1] * (ik[0] * ik[0] + ik[1] * ik[1])) * iD,
                                    (ij[0] * (ik[0] * ik[0] + ik[1] * ik[1]) -
                                     ik[0] * (ij[0] * ij[0] + ij[1] * ij[1])) * iD};
                                const double R2 = abs2(it);
                                int inside = 0;
                                for (int vl = -1; vl < 2 && !inside; vl++) {
                                    for (int ul = -1; ul < 2 && !inside; ul++) {
                                        const ilist* bl =
                                            bin(c, wrapi(nx + ul, c->Bx), wrapi(ny + vl, c->By));
                                        if (!bl) return -2;
                                        for (int ll = 0; ll < bl->n; ll++) {
                                            const int l = bl->v[ll];
                                            if (l == i || l == j || l == k) continue;
                                            double il[2];
                                            diff(c, c->rings.v[l].r, ri, il);
                                            const double dl[2] = {il[0] - it[0], il[1] - it[1]};
                                            if (abs2(dl) < R2) {
                                                inside = 1;
                                                break;
                                            }
                                        }
                                    }
                                }
                                if (inside) continue;
                                const double s[2] = {ij[0] + ik[0], ij[1] + ik[1]};
                                const double third = 1.0 / 3.0;
                                double r[2] = {ri[0] + third * s[0], ri[1] + third * s[1]};
                                wrapv(c, r);
                                const int n = c->atoms.n;
                                if (nl_push(&c->atoms, r[0], r[1])) return -1;
                                if (bros_n_hos(c, n, i, j, k)) return -1;
                            }
                        }
                    }
                }
            }
        }
    }
    return 0;
}

/* atoms sharing two rings are neighbours (src/Coordinator.java:264-283) */
static int bond_atoms(pfc_coord* c) {
    node* A = c->atoms.v;
    node* R = c->rings.v;
    for (int n = 0; n < c->atoms.n; n++) {
        for (int a = 0; a < A[n].hos.n; a++) {
            const int i = A[n].hos.v[a];
            for (int b = a + 1; b < A[n].hos.n; b++) {
                const int j = A[n].hos.v[b];
                for (int q = 0; q < R[i].hos.n; q++) {
                    const int m = R[i].hos.v[q];
                    if (n != m && il_has(&A[m].hos, j) && !il_has(&A[n].bros, m)) {
                        if (il_push(&A[n].bros, m) || il_push(&A[m].bros, n)) return -1;
                    }
                }
            }
        }
    }
    return 0;
}

static void flag(pfc_coord* c, int f) {
    if (c->flags & f) return;
    c->flags |= f;
    c->order[c->norder++] = f;
}

/* src/Coordinator.java:297-318; the reference stops at the first ring of
 * flag 8, so later checks are not recorded after it */
static void check(pfc_coord* c) {
    c->bad_ring = -1;
    for (int n = 0; n < c->atoms.n; n++) {
        if (c->atoms.v[n].bros.n != 3) flag(c, 1);
        if (c->atoms.v[n].hos.n != 3) flag(c, 2);
    }
    for (int i = 0; i < c->rings.n; i++) {
        const node* r = &c->rings.v[i];
        if (r->bros.n < 5 || r->bros.n > 7) flag(c, 4);
        if (r->bros.n != r->hos.n) {
            flag(c, 8);
            c->bad_ring = i;
            break;
        }
    }
}

int pfc_coord_messages(const pfc_coord* c, int* order) {
    for (int i = 0; i < c->norder; i++) order[i] = c->order[i];
    return c->norder;
}

pfc_coord* pfc_coord_run(double* psi, int Nx, int Ny, double dx, double dy, double lPFC,
                         double lA) {
    pfc_coord* c = (pfc_coord*)calloc(1, sizeof(pfc_coord));
    if (!c) return NULL;
    c->Nx = Nx;
    c->Ny = Ny;
    c->lA = lA;
    c->ux = dx * lA / lPFC;
    c->uy = dy * lA / lPFC;
    c->w = Nx * c->ux;
    c->h = Ny * c->uy;
    c->Bx = j2i(ceil(0.5 * c->w / lA));
    c->By = j2i(ceil(0.5 * c->h / lA));
    if (c->Bx == 1 || c->By == 1) c->flags |= 16;
    if (c->Bx < 1 || c->By < 1) {
        free(c);
        return NULL;
    }
    c->bins = (ilist*)calloc((size_t)c->Bx * c->By, sizeof(ilist));
    if (!c->bins || find_rings(c, psi) || find_atoms(c) || bond_atoms(c)) {
        pfc_coord_free(c);
        return NULL;
    }
    check(c);
    return c;
}

int pfc_coord_atoms(const pfc_coord* c) { return c->atoms.n; }
int pfc_coord_rings(const pfc_coord* c) { return c->rings.n; }

void pfc_coord_positions(const pfc_coord* c, double* xy) {
    for (int n = 0; n < c->atoms.n; n++) {
        xy[2 * n] = c->atoms.v[n].r[0];
        xy[2 * n + 1] = c->atoms.v[n].r[1];
    }
}

void pfc_coord_box(const pfc_coord* c, double* w, double* h) {
    *w = c->w;
    *h = c->h;
}

int pfc_coord_flags(const pfc_coord* c, int* bad_ring) {
    if (bad_ring) *bad_ring = c->bad_ring;
    return c->flags;
}

void pfc_coord_free(pfc_coord* c) {
    if (!c) return;
    for (int i = 0; i < c->rings.n; i++) {
        free(c->rings.v[i].bros.v);
        free(c->rings.v[i].hos.v);
    }
    for (int i = 0; i < c->atoms.n; i++) {
        free(c->atoms.v[i].bros.v);
        free(c->atoms.v[i].hos.v);
    }
    free(c->rings.v);
    free(c->atoms.v);
    if (c->bins)
        for (int i = 0; i < c->Bx * c->By; i++) free(c->bins[i].v);
    free(c->bins);
    free(c);
}

/* Shortest decimal digits of d that read back to d: the first precision at
 * which %.{p}e round-trips.  Returns the digit string (no dot) and the
 * decimal exponent of the first digit. */
static void shortest(double d, char* digits, int* exp10) {
    char tmp[40];
    for (int p = 0; p < 17; p++) {
        snprintf(tmp, sizeof tmp, "%.*e", p, d);
        if (strtod(tmp, NULL) == d) break;
    }
    /* tmp = "D.DDDDe[+-]XX" */
    int k = 0;
    for (const char* s = tmp; *s && *s != 'e'; s++)
        if (*s >= '0' && *s <= '9') digits[k++] = *s;
    while (k > 1 && digits[k - 1] == '0') k--;
    digits[k] = 0;
    *exp10 = atoi(strchr(tmp, 'e') + 1);
}

void pfc_java_double(double d, char* buf) {
    if (d != d) {
        strcpy(buf, "NaN");
        return;
    }
    if (isinf(d)) {
        strcpy(buf, d > 0 ? "Infinity" : "-Infinity");
        return;
    }
    if (d == 0.0) {
        strcpy(buf, signbit(d) ? "-0.0" : "0.0");
        return;
    }
    char dg[24];
    int e;
    char* o = buf;
    if (d < 0) *o++ = '-';
    shortest(fabs(d), dg, &e);
    const int nd = (int)strlen(dg);
    const double a = fabs(d);
    if (a >= 1.0e-3 && a < 1.0e7) {
        if (e >= 0) {
            for (int i = 0; i <= e; i++) *o++ = (i < nd) ? dg[i] : '0';
            *o++ = '.';
            if (nd > e + 1)
                for (int i = e + 1; i < nd; i++) *o++ = dg[i];
            else
                *o++ = '0';
        } else {
            *o++ = '0';
            *o++ = '.';
            for (int i = 0; i < -e - 1; i++) *o++ = '0';
            for (int i = 0; i < nd; i++) *o++ = dg[i];
        }
        *o = 0;
    } else {
        *o++ = dg[0];
        *o++ = '.';
        if (nd > 1)
            for (int i = 1; i < nd; i++) *o++ = dg[i];
        else
            *o++ = '0';
        sprintf(o, "E%d", e);
    }
}

int pfc_coord_write(const pfc_coord* c, const char* xyz, const char* nonhex) {
    FILE* f = fopen(xyz, "w");
    if (!f) return -1;
    char a[32], b[32];
    pfc_java_double(c->w, a);
    pfc_java_double(c->h, b);
    fprintf(f, "%s %s\n", a, b);
    for (int n = 0; n < c->atoms.n; n++) {
        const node* p = &c->atoms.v[n];
        pfc_java_double(p->r[0], a);
        pfc_java_double(p->r[1], b);
        fprintf(f, "A %s %s 0.0", a, b);
        for (int m = 0; m < p->bros.n; m++) fprintf(f, " %d", p->bros.v[m]);
        fputc('\n', f);
    }
    if (fclose(f)) return -1;
    if (!nonhex) return 0;
    f = fopen(nonhex, "w");
    if (!f) return -1;
    for (int i = 0; i < c->rings.n; i++) {
        const node* r = &c->rings.v[i];
        if (r->bros.n == 6) continue;
        for (int n = 0; n < r->hos.n; n++) fprintf(f, "%d ", r->hos.v[n]);
        fputc('\n', f);
    }
    return fclose(f) ? -1 : 0;
}

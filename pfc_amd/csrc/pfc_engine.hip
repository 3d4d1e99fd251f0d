// pfc_engine.hip -- libpfcgpu: the C ABI of include/pfc_relax.h on gfx950.
//
// Device state of a context (E = complex element, FP64 or FP32 by precision;
// nk = (Nx/2)*P with column pitch P = Ny (+ padding from Ny = 1024), nr = Nx*Ny):
//   I, J, K       U*nk E each (packed row transforms)
//   Nh            U*nkn E, nkn = (Nx/2+1)*Ny (the spectral state)
//   O             (2U + U(U-1)/2)*nk E   evaluate's k-space outputs
//   mreal         U*nr real    densities in real space (materialised on demand)
//   R             (2U + U(U-1)/2)*nr real  evaluate's real-space outputs
//   f             nr doubles   energy density of the last evaluate
//   ktab          Gaussian factors exp(-k^2/sigma^2), per kx then per ky (ktab_gy)
// Everything runs on one HIP stream owned by the context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/pfc_relax.h"
#include "pfc_launch.hpp"

extern template struct pfc::Launch<double>;
extern template struct pfc::Launch<float>;

using namespace pfc;

namespace {

using pfc::Launch;

#define HIPCHK(x)                                 \
    do {                                          \
        hipError_t e_ = (x);                      \
        if (e_ != hipSuccess) return PFC_EDEVICE; \
    } while (0)

template <typename T>
void fill_twiddles(std::vector<cpx<T>>& tw, int n) {
    tw.resize(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < n; m++) {
        const long double a = -2.0L * pi * (long double)m / (long double)n;
        tw[m].x = (T)cosl(a);
        tw[m].y = (T)sinl(a);
    }
}

void fill_devparams(pfc_ctx* c) {
    DevParams& d = c->hp;
    const pfc_params& p = c->p;
#if PFC_PROBE
    unsigned long long* probe = d.probe;
#endif
    std::memset(&d, 0, sizeof(d));
#if PFC_PROBE
    d.probe = probe;
#endif
    d.U = p.U;
    for (int u = 0; u < p.U; u++) d.W[u] = p.W[u];
    d.energy_compat = p.energy_compat;
    d.c1 = (p.c != 0.0) ? 1.0 : 0.0;
    d.dt = p.dt;
    d.sig_a = -0.5 / p.sigma / p.sigma;
    d.NxNy = (double)c->Nx * c->Ny;
    d.inv_NxNy = 1.0 / c->Nx / c->Ny;
    const double pi = 3.14159265358979323846264338327;
    d.dkx = 2.0 * pi / (c->dx * c->Nx);
    d.dky = 2.0 * pi / (c->dy * c->Ny);
    for (int u = 0; u < p.U; u++) {
        d.lambda[u] = p.lambda[u];
        d.delta[u] = p.delta[u];
        for (int v = 0; v < p.U; v++) {
            d.alpha[u][v] = p.alpha[u][v];
            d.beta[u][v] = p.beta[u][v];
            d.gamma[u][v] = p.gamma[u][v];
            d.epsilon[u][v] = p.epsilon[u][v];
            d.mu2[u][v] = p.mu2[u][v];
            if (p.energy_compat == PFC_ENERGY_NU2)
                d.qc2[u][v] = (v < p.W[u]) ? p.nu2[u][v] : 0.0;
            else
                d.qc2[u][v] = p.mu2[u][v];
        }
        for (int w = 0; w < p.W[u]; w++) {
            d.nu2[u][w] = p.nu2[u][w];
            d.omega[u][w] = p.omega[u][w];
            d.phi[u][w] = p.phi[u][w];
        }
    }
}

// Gaussian factor table of the column pass (layout at ktab_gy): for kx =
// 0..Nx/2 and then per row index j (ky = j dky for j < Ny/2, else (j - Ny)
// dky) the factor exp(2 a k^2) with a = -0.5/sigma^2, so that G2 = exp(a k2)^2
// (src/pfc-relax.c:1223-1225) is the product of a kx and a ky factor.
size_t ktab_bytes(const pfc_ctx* c) {
    return (size_t)(ktab_gy(c->Nx / 2) + c->Ny) * (c->esz / 2);
}

template <typename T>
void fill_ktab(pfc_ctx* c) {
    const int H = c->Nx / 2;
    T* k = (T*)c->hktab;
    const double a2 = 2.0 * c->hp.sig_a;
    std::memset(c->hktab, 0, (size_t)(ktab_gy(H) + c->Ny) * sizeof(T));
    for (int i = 0; i <= H; i++) {
        double kx = i * c->hp.dkx;
        kx *= kx;
        k[i] = (T)std::exp(a2 * kx);
    }
    for (int j = 0; j < c->Ny; j++) {
        const double ky = (j < c->Ny / 2) ? j * c->hp.dky : (j - c->Ny) * c->hp.dky;
        k[ktab_gy(H) + j] = (T)std::exp(a2 * (ky * ky));
    }
}

int upload_params(pfc_ctx* c) {
    HIPCHK(hipStreamSynchronize(c->stream));   // the host copies below are reused
    fill_devparams(c);
    if (c->prec == PFC_FP64) fill_ktab<double>(c);
    else fill_ktab<float>(c);
    HIPCHK(hipMemcpyAsync(c->dprm, &c->hp, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ktab, c->hktab, ktab_bytes(c), hipMemcpyHostToDevice, c->stream));
    return PFC_OK;
}

void free_all(pfc_ctx* c) {
    gen_destroy(c);
    void* bufs[] = {c->I, c->J, c->K, c->Nh, c->Nh2, c->O, c->mreal, c->R, c->twx, c->twy,
                    c->ktab, c->f, c->partials, c->sums, c->nan_flag, c->dprm, c->n0};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    free(c->hktab);
    if (c->hsums) (void)hipHostFree(c->hsums);
    if (c->ev_sums) (void)hipEventDestroy(c->ev_sums);
    for (int i = 0; i < c->nev; i++) (void)hipEventDestroy(c->ev[i]);
    free(c->ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// Grows the event pool to hold n recorded steps (3 events each).
int ensure_events(pfc_ctx* c, int n) {
    if (3 * n <= c->nev) return PFC_OK;
    int cap = c->nev ? c->nev : 64;
    while (cap < 3 * n) cap *= 2;
    hipEvent_t* ev = (hipEvent_t*)realloc(c->ev, sizeof(hipEvent_t) * cap);
    if (!ev) return PFC_ENOMEM;
    c->ev = ev;   // the old block may have moved; keep the new one at once
    // nev counts created events only, so that free_all destroys valid handles
    while (c->nev < cap) {
        if (hipEventCreate(&c->ev[c->nev]) != hipSuccess) return PFC_EDEVICE;
        c->nev++;
    }
    return PFC_OK;
}

int dalloc(void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return PFC_ENOMEM;
    }
    return PFC_OK;
}

// Multiplies every real and imaginary part x[i] by 1 + rel * r_i, r_i a
// zero-mean, unit-variance uniform deviate from a counter hash of (seed, i).
// With nkn > 0, x is the spectral state (U fields of nkn complex entries,
// column kx at kx * Ny) and its columns kx = 0 and kx = H are left alone:
// their anti-Hermitian part is invisible to the density and grows with the
// unstable modes until the next refresh, so noise there would be amplified
// ~1e15-fold (the FP32 form projects those columns on their Hermitian part
// for the same reason).
template <typename T>
__global__ void k_perturb(T* x, size_t n, size_t nkn, int Ny, double rel, unsigned long long seed) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (nkn) {
            const size_t j = (i / 2) % nkn;
            if (j < (size_t)Ny || j >= nkn - (size_t)Ny) continue;
        }
        unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);   // splitmix64
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);   // [0, 1)
        x[i] = (T)((double)x[i] * (1.0 + rel * 1.7320508075688772 * (2.0 * u - 1.0)));
    }
}

// Materialise the densities in mreal; if K, also refresh Nh from them.
// (The generic path keeps them there.)
int materialise(pfc_ctx* c, bool refresh) {
    if (c->generic) {
        if (refresh) gen_refresh(c);
        HIPCHK(hipGetLastError());
        return PFC_OK;
    }
    if (c->prec == PFC_FP64) {
        Launch<double>::row_c2r_any(c, c->I, c->sI, c->mreal, refresh ? c->K : nullptr, c->U);
        if (refresh) Launch<double>::col_fwd_any(c, c->K, c->sJ, c->Nh, c->U, c->hp.inv_NxNy);
    } else {
        Launch<float>::row_c2r_any(c, c->I, c->sI, c->mreal, refresh ? c->K : nullptr, c->U);
        if (refresh) Launch<float>::col_fwd_any(c, c->K, c->sJ, c->Nh, c->U, c->hp.inv_NxNy);
    }
    HIPCHK(hipGetLastError());
    if (refresh) c->nh_fresh = 1;
    return PFC_OK;
}

}  // namespace

// ============================================================== C ABI

extern "C" {

const char* pfc_strerror(int code) {
    switch (code) {
        case PFC_OK: return "success";
        case PFC_EINVAL: return "invalid argument";
        case PFC_EDEVICE: return "HIP runtime error";
        case PFC_ENOMEM: return "device allocation failed";
        case PFC_ESIZE:
            return "grid size not supported by this build (each of Nx, Ny must be at least 2 and its "
                   "transform must fit one block's LDS: lengths whose prime factors are at most 13 up "
                   "to 4864 (FP64) or 9728 (mixed), others up to 2048 (FP64) or 4096 (mixed))";
        case PFC_ESTATE: return "call out of order";
        default: return "unknown error";
    }
}

int pfc_create(pfc_ctx** out, const pfc_params* p, int Nx, int Ny, int precision) {
    if (!out || !p) return PFC_EINVAL;
    *out = nullptr;
    if (p->U < 1 || p->U > PFC_MAX_FIELDS) return PFC_EINVAL;
    for (int u = 0; u < p->U; u++)
        if (p->W[u] < 0 || p->W[u] > PFC_MAX_MODES) return PFC_EINVAL;
    if (precision != PFC_FP64 && precision != PFC_MIXED) return PFC_EINVAL;
    // the tuned kernels where they exist, else the generic path
    const bool generic = !(size_ok(Nx) && size_ok(Ny) && p->U <= kMaxU);
    const size_t esz0 = precision == PFC_FP64 ? 16 : 8;
    if (generic && !(gen_plannable(Nx, esz0) && gen_plannable(Ny, esz0))) return PFC_ESIZE;
    int ndev = 0, dev = 0, ncu = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFC_EDEVICE;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        return PFC_EDEVICE;

    pfc_ctx* c = (pfc_ctx*)calloc(1, sizeof(pfc_ctx));
    if (!c) return PFC_ENOMEM;
    c->Nx = Nx;
    c->Ny = Ny;
    c->U = p->U;
    c->prec = precision;
    c->p = *p;
    c->ncu = ncu;
    {
        const char* e = getenv("PFC_PERSIST");
        c->persist = !(e && e[0] == '0');
        // The per-field column pass where it measured faster, with three
        // fields: 4096^2 and 8192^2 (FP64 column pass 399 -> 363 us and
        // 2.40 -> 1.71 ms, FP32 at 8192 831 -> 750 us; FP32 columns of 4096,
        // whose in-place pass already runs two blocks per CU, 162 -> 180 us)
        // and every grid up to 1024^2 (whole steps 13-25 % faster in both
        // forms: U times the blocks; at 2048^2 1-7 % slower);
        // PFC_COLSPLIT=1 takes it wherever it is built, =0 nowhere.
        e = getenv("PFC_COLSPLIT");
        const bool built = !generic && (precision == PFC_FP64 ? Launch<double>::col_split_built(Ny)
                                                              : Launch<float>::col_split_built(Ny));
        const bool pays = p->U >= 2 && (Ny <= 1024 || Ny >= (precision == PFC_FP64 ? 4096 : 8192));
        c->colsplit = built && (e && e[0] ? e[0] == '1' : pays);
    }
    c->nO = 2 * c->U + c->U * (c->U - 1) / 2;
    c->nkn = (size_t)(Nx / 2 + 1) * Ny;   // half spectrum, kx = 0..Nx/2
    c->nr = (size_t)Nx * Ny;
    c->esz = (precision == PFC_FP64) ? 16 : 8;
    c->sC = ilog2(Ny);
    c->sI = ilog2(precision == PFC_FP64 ? row_block<double>(Nx, Ny, PFC_I_BLOCKED)
                                        : row_block<float>(Nx, Ny, PFC_I_BLOCKED));
    c->sJ = ilog2(precision == PFC_FP64 ? row_block<double>(Nx, Ny, PFC_J_BLOCKED)
                                        : row_block<float>(Nx, Ny, PFC_J_BLOCKED));
    {
        // J and K with their columns in groups of 2^x (layout x): the row
        // pass writes runs of 2^x columns' TR rows, 2^x times as long as
        // column-major ones, and the column pass reads parts of lines that
        // its neighbouring column blocks read too.  Measured at 4096^2 and
        // 8192^2 with three fields: FP64 pairs, row pass 355 -> 328 us and
        // 1.98 -> 1.69 ms, column pass 378 -> 390 us and unchanged (groups of
        // 4: column pass slower, step 3.46 -> 3.57 ms at 8192^2); FP32 at
        // 8192 groups of 4, row 0.98 -> 0.73 ms, column 0.76 -> 0.83 ms (step
        // 1.75 -> 1.57 ms; pairs 1.64 ms); FP32 at 4096 the column pass lost
        // more than the row pass gained.  I stays column-major: with the
        // column pass writing half lines it took twice as long.
        // PFC_XPAIR=J takes pairs for every tuned shape, =4 groups of 4, =0
        // neither.
        const char* e = getenv("PFC_XPAIR");
        int x = 0;
        if (e && e[0]) x = e[0] == 'J' ? 1 : e[0] == '4' ? 2 : 0;
        else if (precision == PFC_FP64) x = (Nx >= 4096 && Ny >= 4096) ? 1 : 0;
        else x = (Nx >= 8192 && Ny >= 8192) ? 2 : 0;
        if (!generic) c->sJ |= 32 * x;
    }
    // field stride of I, J, K, O: the largest of their layouts (padded strides)
    c->nk = std::max({lay_elems(c->sC, Nx / 2, Ny, c->esz), lay_elems(c->sI, Nx / 2, Ny, c->esz),
                      lay_elems(c->sJ, Nx / 2, Ny, c->esz)});
    c->generic = generic;
    if (generic) c->nk = c->nkn;   // the transposed half-spectrum layout
    const size_t rsz = c->esz / 2;
    c->dx = 1.0;
    c->dy = 1.0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        free_all(c);
        free(c);
        return PFC_EDEVICE;
    }
    int rc = PFC_OK;
    const size_t spec = c->U * c->nk * c->esz;
    rc |= dalloc(&c->I, spec);
    rc |= dalloc(&c->J, spec);
    rc |= dalloc(&c->K, spec);
    rc |= dalloc(&c->Nh, c->U * c->nkn * c->esz);
    if (c->colsplit || generic) rc |= dalloc(&c->Nh2, c->U * c->nkn * c->esz);
    rc |= dalloc(&c->O, c->nO * c->nk * c->esz);
    rc |= dalloc(&c->mreal, c->U * c->nr * rsz);
    rc |= dalloc(&c->R, c->nO * c->nr * rsz);
    rc |= dalloc((void**)&c->f, c->nr * sizeof(double));
    // energy partials: per block 1 + U sums (f or its local part, each n_u),
    // for up to max(kEnergyBlocks, Ny) blocks (a step's row pass), then from
    // poff the spectral part, up to max(kEnergyBlocks, Nx/2 + 1) blocks (a
    // step's column pass)
    c->poff = (size_t)std::max(kEnergyBlocks, Ny) * 2 * (1 + c->U);
    rc |= dalloc((void**)&c->partials,
                 (c->poff + (size_t)std::max(kEnergyBlocks, Nx / 2 + 1) * 2) * sizeof(double));
    rc |= dalloc((void**)&c->sums, (2 + c->U) * sizeof(double));
    if (hipHostMalloc((void**)&c->hsums, (2 + c->U) * sizeof(double), hipHostMallocDefault) !=
        hipSuccess) {
        c->hsums = nullptr;
        rc |= PFC_ENOMEM;
    }
    if (hipEventCreateWithFlags(&c->ev_sums, hipEventDisableTiming) != hipSuccess) {
        c->ev_sums = nullptr;
        rc |= PFC_EDEVICE;
    }
    rc |= dalloc((void**)&c->nan_flag, sizeof(int));
    rc |= dalloc((void**)&c->dprm, sizeof(DevParams));
    // the k = 0 state and its alternate of the per-field column pass
    rc |= dalloc((void**)&c->n0, 4 * c->U * sizeof(double));
#if PFC_PROBE
    rc |= dalloc((void**)&c->hp.probe, 2 * 65536 * sizeof(unsigned long long));
#endif
    rc |= dalloc(&c->twx, Nx * c->esz);
    rc |= dalloc(&c->twy, Ny * c->esz);
    rc |= dalloc(&c->ktab, ktab_bytes(c));
    c->hktab = malloc(ktab_bytes(c));
    if (!c->hktab) rc |= PFC_ENOMEM;
    if (rc != PFC_OK) {
        free_all(c);
        free(c);
        return PFC_ENOMEM;
    }
    if (generic && (rc = gen_create(c)) != PFC_OK) {
        free_all(c);
        free(c);
        return rc;
    }
    bool ok = true;
    if (generic) {
        // the tuned kernels' tables are not used
    } else if (precision == PFC_FP64) {
        std::vector<cpx<double>> tx, ty;
        fill_twiddles(tx, Nx);
        fill_twiddles(ty, Ny);
        ok &= hipMemcpy(c->twx, tx.data(), Nx * c->esz, hipMemcpyHostToDevice) == hipSuccess;
        ok &= hipMemcpy(c->twy, ty.data(), Ny * c->esz, hipMemcpyHostToDevice) == hipSuccess;
    } else {
        std::vector<cpx<float>> tx, ty;
        fill_twiddles(tx, Nx);
        fill_twiddles(ty, Ny);
        ok &= hipMemcpy(c->twx, tx.data(), Nx * c->esz, hipMemcpyHostToDevice) == hipSuccess;
        ok &= hipMemcpy(c->twy, ty.data(), Ny * c->esz, hipMemcpyHostToDevice) == hipSuccess;
    }
    const int big = INT_MAX;
    ok &= hipMemcpy(c->nan_flag, &big, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemset(c->f, 0, c->nr * sizeof(double)) == hipSuccess;
    if (!ok || upload_params(c) != PFC_OK || hipStreamSynchronize(c->stream) != hipSuccess) {
        free_all(c);
        free(c);
        return PFC_EDEVICE;
    }
    *out = c;
    return PFC_OK;
}

int pfc_destroy(pfc_ctx* c) {
    if (!c) return PFC_EINVAL;
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    free(c);
    return PFC_OK;
}

int pfc_set_box(pfc_ctx* c, double dx, double dy) {
    if (!c) return PFC_EINVAL;
    c->dx = dx;
    c->dy = dy;
    c->have_box = 1;
    return upload_params(c);
}

int pfc_set_fields(pfc_ctx* c, const double* fields) {
    if (!c || !fields) return PFC_EINVAL;
    const size_t n = c->U * c->nr;
    if (c->prec == PFC_FP64) {
        HIPCHK(hipMemcpyAsync(c->mreal, fields, n * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->generic) {
            gen_refresh(c);
        } else {
            Launch<double>::row_r2c_any(c, c->mreal, c->I, c->sI, c->U);
            Launch<double>::col_fwd_any(c, c->I, c->sI, c->Nh, c->U, 1.0 / c->Ny);  // I carries 1/Nx
        }
    } else {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; i++) tmp[i] = (float)fields[i];
        HIPCHK(hipMemcpyAsync(c->mreal, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->generic) {
            gen_refresh(c);
        } else {
            Launch<float>::row_r2c_any(c, c->mreal, c->I, c->sI, c->U);
            Launch<float>::col_fwd_any(c, c->I, c->sI, c->Nh, c->U, 1.0 / c->Ny);  // I carries 1/Nx
        }
        // the k = 0 state in FP64: each field's mean, summed in long double
        std::vector<double> m0(2 * c->U, 0.0);
        for (int u = 0; u < c->U; u++) {
            long double acc = 0.0L;
            const double* f = fields + (size_t)u * c->nr;
            for (size_t i = 0; i < c->nr; i++) acc += f[i];
            m0[2 * u] = (double)(acc / (long double)c->nr);
        }
        HIPCHK(hipMemcpyAsync(c->n0, m0.data(), m0.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_fields = 1;
    c->nh_fresh = !c->generic;   // Nh = FFT2(m)/(Nx Ny), as a refresh computes it
    return PFC_OK;
}

int pfc_step(pfc_ctx* c, int t0, int nsteps) {
    if (!c || nsteps < 0) return PFC_EINVAL;
    if (!c->have_fields || !c->have_box) return PFC_ESTATE;
    bool alt = false;   // the state is in Nh2 (between the two steps of a pair)
    const bool fresh = c->nh_fresh;
    if (nsteps > 0) c->nh_fresh = 0;
    for (int i = 0; i < nsteps; i++) {
        const int t = t0 + i;
        // the refresh of src/pfc-relax.c:1178-1184, unless an evaluate or
        // pfc_set_fields has just computed that same state from the same
        // densities (relax() evaluates before every printed step)
        const bool ref = (t % 100 == 0) && !(i == 0 && fresh);
        hipEvent_t* ev = nullptr;
        if (c->timing) {
            if (ensure_events(c, c->nrec + 1) != PFC_OK) return PFC_EDEVICE;
            ev = c->ev + 3 * c->nrec++;
            HIPCHK(hipEventRecord(ev[0], c->stream));
        }
        if (c->generic) {
            // one sequence of launches, timed as the row pass
            gen_step(c, t);
            if (ev) HIPCHK(hipEventRecord(ev[1], c->stream));
            if (ev) HIPCHK(hipEventRecord(ev[2], c->stream));
            continue;
        }
        if (c->prec == PFC_FP64) Launch<double>::row_step_any(c, t, ref);
        else Launch<float>::row_step_any(c, t, ref);
        if (ev) HIPCHK(hipEventRecord(ev[1], c->stream));
        // Non-refresh steps go through the per-field column pass in pairs,
        // Nh -> Nh2 -> Nh, so that the state is in Nh between calls and on
        // refresh steps; a step without a partner takes the in-place pass.
        const bool split = c->Nh2 && (alt || (!ref && i + 1 < nsteps && (t + 1) % 100 != 0));
        if (split) {
            if (c->prec == PFC_FP64)
                Launch<double>::col_step_split_any(c, alt ? c->Nh2 : c->Nh, alt ? c->Nh : c->Nh2);
            else
                Launch<float>::col_step_split_any(c, alt ? c->Nh2 : c->Nh, alt ? c->Nh : c->Nh2);
            alt = !alt;
        } else if (c->prec == PFC_FP64) {
            Launch<double>::col_step_any(c, ref);
        } else {
            Launch<float>::col_step_any(c, ref);
        }
        if (ev) HIPCHK(hipEventRecord(ev[2], c->stream));
    }
    HIPCHK(hipGetLastError());
    return PFC_OK;
}

int pfc_perturb(pfc_ctx* c, double rel, unsigned long long seed) {
    if (!c || !(rel >= 0.0)) return PFC_EINVAL;
    if (!c->have_fields) return PFC_ESTATE;
    c->nh_fresh = 0;
    // real parts of the density's state (the generic path keeps the density
    // itself) and of Nh
    const size_t ni = c->generic ? c->U * c->nr : 2 * c->U * c->nk, nn = 2 * c->U * c->nkn;
    void* dens = c->generic ? c->mreal : c->I;
    if (c->prec == PFC_FP64) {
        hipLaunchKernelGGL(k_perturb<double>, dim3(2048), dim3(256), 0, c->stream, (double*)dens, ni,
                           (size_t)0, c->Ny, rel, seed);
        hipLaunchKernelGGL(k_perturb<double>, dim3(2048), dim3(256), 0, c->stream, (double*)c->Nh, nn,
                           c->nkn, c->Ny, rel, seed ^ 0x5bd1e995ull);
    } else {
        hipLaunchKernelGGL(k_perturb<float>, dim3(2048), dim3(256), 0, c->stream, (float*)dens, ni,
                           (size_t)0, c->Ny, rel, seed);
        hipLaunchKernelGGL(k_perturb<float>, dim3(2048), dim3(256), 0, c->stream, (float*)c->Nh, nn,
                           c->nkn, c->Ny, rel, seed ^ 0x5bd1e995ull);
    }
    HIPCHK(hipGetLastError());
    return PFC_OK;
}

int pfc_timing(pfc_ctx* c, int enable) {
    if (!c) return PFC_EINVAL;
    c->timing = enable ? 1 : 0;
    if (enable) c->nrec = 0;
    return PFC_OK;
}

#if PFC_PROBE
// Diagnostic builds only: copies the phase timestamps of the last row and
// column pass (2 x 65536 words).
int pfc_probe_read(pfc_ctx* c, unsigned long long* host) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(host, c->hp.probe, 2 * 65536 * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost));
    return PFC_OK;
}
#endif

int pfc_timing_read(pfc_ctx* c, double* ms, int* nsteps) {
    if (!c || !ms || !nsteps) return PFC_EINVAL;
    HIPCHK(hipStreamSynchronize(c->stream));
    ms[0] = ms[1] = 0.0;
    for (int i = 0; i < c->nrec; i++) {
        hipEvent_t* e = c->ev + 3 * i;
        for (int k = 0; k < 2; k++) {
            float v = 0.f;
            HIPCHK(hipEventElapsedTime(&v, e[k], e[k + 1]));
            ms[k] += v;
        }
    }
    *nsteps = c->nrec;
    return PFC_OK;
}

int pfc_evaluate(pfc_ctx* c, double* fave, double* nave) {
    if (!c || !fave || !nave) return PFC_EINVAL;
    if (!c->have_fields || !c->have_box) return PFC_ESTATE;
    if (c->generic) {
        gen_evaluate(c);
    } else if (c->prec == PFC_FP64) {
        Launch<double>::row_c2r_any(c, c->I, c->sI, c->mreal, c->K, c->U);
        Launch<double>::col_eval_any(c);
        Launch<double>::row_c2r_any(c, c->O, c->sC, c->R, nullptr, c->nO);
        Launch<double>::energy_any(c);
    } else {
        Launch<float>::row_c2r_any(c, c->I, c->sI, c->mreal, c->K, c->U);
        Launch<float>::col_eval_any(c);
        Launch<float>::row_c2r_any(c, c->O, c->sC, c->R, nullptr, c->nO);
        Launch<float>::energy_any(c);
    }
    c->nh_fresh = !c->generic;   // k_col_eval stored the refreshed state
    hipLaunchKernelGGL((k_reduce_final<1024>), dim3(1), dim3(1024), 0, c->stream, c->partials,
                       kEnergyBlocks, 1 + c->U, c->sums, c->hp.inv_NxNy);
    HIPCHK(hipGetLastError());
    double h[1 + PFC_MAX_FIELDS];
    HIPCHK(hipMemcpyAsync(h, c->sums, (1 + c->U) * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *fave = h[0];
    for (int u = 0; u < c->U; u++) nave[u] = h[1 + u];
    c->f_stale = 0;
    return PFC_OK;
}

int pfc_evaluate_sums(pfc_ctx* c, double* fave, double* nave) {
    if (!c || !fave || !nave) return PFC_EINVAL;
    if (!c->have_fields || !c->have_box) return PFC_ESTATE;
    // the refreshed state, as the full evaluate leaves it; the tuned path
    // takes the local energy terms from the densities in registers during
    // their row pass and writes no densities
    int nloc = kEnergyBlocks;
    if (c->generic) {
        const int rc = materialise(c, true);
        if (rc) return rc;
        gen_energy_sums(c);
    } else if (c->prec == PFC_FP64) {
        nloc = Launch<double>::eval_local_any(c);
        Launch<double>::col_fwd_any(c, c->K, c->sJ, c->Nh, c->U, c->hp.inv_NxNy);
        Launch<double>::energy_spec_any(c);
    } else {
        nloc = Launch<float>::eval_local_any(c);
        Launch<float>::col_fwd_any(c, c->K, c->sJ, c->Nh, c->U, c->hp.inv_NxNy);
        Launch<float>::energy_spec_any(c);
    }
    if (!c->generic) c->nh_fresh = 1;
    hipLaunchKernelGGL((k_reduce_final<1024>), dim3(1), dim3(1024), 0, c->stream, c->partials,
                       nloc, 1 + c->U, c->sums, c->hp.inv_NxNy);
    hipLaunchKernelGGL((k_reduce_final<1024>), dim3(1), dim3(1024), 0, c->stream,
                       c->partials + c->poff, kEnergyBlocks, 1,
                       c->sums + 1 + c->U, 1.0);
    HIPCHK(hipGetLastError());
    double h[2 + PFC_MAX_FIELDS];
    HIPCHK(hipMemcpyAsync(h, c->sums, (2 + c->U) * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *fave = h[0] + h[1 + c->U];
    for (int u = 0; u < c->U; u++) nave[u] = h[1 + u];
    c->f_stale = 1;
    return PFC_OK;
}

int pfc_step_evaluate(pfc_ctx* c, int t0, int nsteps, double* fave, double* nave) {
    if (!c || nsteps < 1 || !fave || !nave) return PFC_EINVAL;
    if (!c->have_fields || !c->have_box) return PFC_ESTATE;
    if (c->generic) {
        const int rc = pfc_evaluate_sums(c, fave, nave);
        return rc ? rc : pfc_step(c, t0, nsteps);
    }
    // A refresh step t0 (t0 % 100 == 0, no evaluate since the last step)
    // runs with the passes that also write the evaluate's partial sums of the
    // refreshed state: the local terms from the densities the row pass
    // computes, the spectral ones from the state the column pass refreshes
    // (FP64 at 4096^2 with three fields: 14.93 against 15.06 ms for a print
    // and 20 steps).  Otherwise, and in the mixed form (whose FP32 row pass
    // with the sums' registers runs one block per CU instead of two, and
    // which gained nothing: 7.16 against 7.15 ms), the evaluate runs on its
    // own, before the steps.
    if (t0 % 100 != 0 || c->nh_fresh || c->prec != PFC_FP64) {
        const int rc = pfc_evaluate_sums(c, fave, nave);
        return rc ? rc : pfc_step(c, t0, nsteps);
    }
    int nrow, ncol;
    if (c->prec == PFC_FP64) {
        nrow = Launch<double>::row_step_any(c, t0, true, c->partials);
        ncol = Launch<double>::col_step_any(c, true, c->partials + c->poff);
    } else {
        nrow = Launch<float>::row_step_any(c, t0, true, c->partials);
        ncol = Launch<float>::col_step_any(c, true, c->partials + c->poff);
    }
    hipLaunchKernelGGL((k_reduce_final<1024>), dim3(1), dim3(1024), 0, c->stream, c->partials, nrow,
                       1 + c->U, c->sums, c->hp.inv_NxNy);
    hipLaunchKernelGGL((k_reduce_final<1024>), dim3(1), dim3(1024), 0, c->stream,
                       c->partials + c->poff, ncol, 1, c->sums + 1 + c->U, 1.0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->hsums, c->sums, (2 + c->U) * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipEventRecord(c->ev_sums, c->stream));
    c->f_stale = 1;
    if (nsteps > 1) {
        const int rc = pfc_step(c, t0 + 1, nsteps - 1);
        if (rc) return rc;
    }
    HIPCHK(hipEventSynchronize(c->ev_sums));
    *fave = c->hsums[0] + c->hsums[1 + c->U];
    for (int u = 0; u < c->U; u++) nave[u] = c->hsums[1 + u];
    return PFC_OK;
}

int pfc_get_fields(pfc_ctx* c, double* fields) {
    if (!c || !fields) return PFC_EINVAL;
    if (!c->have_fields) return PFC_ESTATE;
    int rc = materialise(c, false);
    if (rc) return rc;
    const size_t n = c->U * c->nr;
    if (c->prec == PFC_FP64) {
        HIPCHK(hipMemcpyAsync(fields, c->mreal, n * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp(n);
        HIPCHK(hipMemcpyAsync(tmp.data(), c->mreal, n * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < n; i++) fields[i] = (double)tmp[i];
    }
    return PFC_OK;
}

int pfc_get_energy(pfc_ctx* c, double* f) {
    if (!c || !f) return PFC_EINVAL;
    if (c->f_stale) return PFC_ESTATE;
    HIPCHK(hipMemcpyAsync(f, c->f, c->nr * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PFC_OK;
}

int pfc_nan_step(pfc_ctx* c, int* step) {
    if (!c || !step) return PFC_EINVAL;
    int v = INT_MAX;
    HIPCHK(hipMemcpyAsync(&v, c->nan_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *step = (v == INT_MAX) ? -1 : v;
    return PFC_OK;
}

int pfc_smooth(const double* in, double* out, int Nx, int Ny, double dx, double dy,
               double sigma, int precision) {
    if (!in || !out || !(sigma > 0.0) || !(dx > 0.0) || !(dy > 0.0)) return PFC_EINVAL;
    pfc_params p;
    std::memset(&p, 0, sizeof p);
    p.U = 1;
    pfc_ctx* c = nullptr;
    int rc = pfc_create(&c, &p, Nx, Ny, precision);
    if (rc) return rc;
    rc = pfc_set_box(c, dx, dy);
    const size_t n = c->nr;
    if (!rc) {
        if (precision == PFC_FP64) {
            if (hipMemcpyAsync(c->mreal, in, n * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
                hipSuccess)
                rc = PFC_EDEVICE;
            if (!rc) {
                if (c->generic) {
                    gen_smooth(c, c->mreal, c->R, sigma);
                } else {
                    Launch<double>::row_r2c_any(c, c->mreal, c->I, c->sC, 1);
                    Launch<double>::col_smooth_any(c, c->I, c->O, sigma);
                    Launch<double>::row_c2r_any(c, c->O, c->sC, c->R, nullptr, 1);
                }
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(out, c->R, n * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream) != hipSuccess ||
                    hipStreamSynchronize(c->stream) != hipSuccess)
                    rc = PFC_EDEVICE;
            }
        } else {
            std::vector<float> tmp(n);
            for (size_t i = 0; i < n; i++) tmp[i] = (float)in[i];
            if (hipMemcpyAsync(c->mreal, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice,
                               c->stream) != hipSuccess)
                rc = PFC_EDEVICE;
            if (!rc) {
                if (c->generic) {
                    gen_smooth(c, c->mreal, c->R, sigma);
                } else {
                    Launch<float>::row_r2c_any(c, c->mreal, c->I, c->sC, 1);
                    Launch<float>::col_smooth_any(c, c->I, c->O, sigma);
                    Launch<float>::row_c2r_any(c, c->O, c->sC, c->R, nullptr, 1);
                }
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(tmp.data(), c->R, n * sizeof(float), hipMemcpyDeviceToHost,
                                   c->stream) != hipSuccess ||
                    hipStreamSynchronize(c->stream) != hipSuccess)
                    rc = PFC_EDEVICE;
            }
            if (!rc)
                for (size_t i = 0; i < n; i++) out[i] = tmp[i];
        }
    }
    pfc_destroy(c);
    return rc;
}

}  // extern "C"

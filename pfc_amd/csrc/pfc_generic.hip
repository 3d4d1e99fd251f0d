// pfc_generic.hip -- the step, evaluation and smoothing for grid shapes and
// field counts the tuned kernels do not cover: any Nx, Ny whose transforms
// fit one block's LDS (fft_generic.hpp) and any U up to PFC_MAX_FIELDS.
//
// Layout: the densities live in real space (mreal, U x Ny x Nx); J, K, Nh
// and O are half spectra of (Nx/2 + 1) columns of Ny values each (column kx
// at kx * Ny), the reference's transposed layout.  A step (src/pfc-relax.c:
// 1133-1247) is three launches:
//   [t % 100 == 0: Nh = FFT2(m) / (Nx Ny), two launches]
//   per row and field: the nonlinear term (m[0] NaN check) and its FFT_x
//     into J                                     k_gen_row_r2c_nl
//   per column and field: FFT_y(J), Nh2 = A Nh + B (FFT_y(J) + couplings)
//     (+ FP32 Hermitian projection of kx = 0, Nx/2), J = IFFT_y(Nh2);
//     Nh and Nh2 then swap                       k_gen_col_step
//   m = C2R_x(J)                                 k_gen_row_c2r
// Every transform is one block per row or column with the sequence in LDS,
// so the real-space and spectral sweeps are separate (more HBM traffic than
// the fused tuned path, which is what this path trades for generality).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/pfc_relax.h"
#include "fft_generic.hpp"
#include "pfc_launch.hpp"

namespace pfc {

namespace {

constexpr int kGenThreads = 256;
constexpr size_t kGenMaxLds = 152 * 1024;   // bytes of LDS one generic transform may use

// Radices of n from {8, 4, 2, 3, 5, 7, 11, 13}; false if another prime is left.
bool gen_factor(int n, GenPlan* p, bool pow2_only) {
    int m = n, k = 0;
    const int rs[] = {8, 4, 2, 3, 5, 7, 11, 13};
    for (int r : rs) {
        if (pow2_only && r != 8 && r != 4 && r != 2) continue;
        while (m % r == 0 && k < kGenMaxStages) {
            p->radix[k++] = r;
            m /= r;
        }
    }
    p->nst = k;
    return m == 1;
}

}  // namespace

bool gen_plan(int L, size_t esz, GenPlan* p) {
    std::memset(p, 0, sizeof *p);
    p->L = L;
    if (L < 2) return false;
    if (gen_factor(L, p, false)) {
        p->M = 0;
    } else {
        int M = 1;
        while (M < 2 * L - 1) M *= 2;
        p->M = M;
        gen_factor(M, p, true);
    }
    return (size_t)gen_lds_elems(*p) * esz <= kGenMaxLds;
}

bool gen_plannable(int L, size_t esz) {
    GenPlan p;
    return gen_plan(L, esz, &p);
}

namespace {

struct GenCtx {
    GenDim x, y;
};

using ld = long double;
const ld kPi = 3.141592653589793238462643383279502884L;

// In-place radix-2 FFT in long double (power-of-two n, forward).
void host_fft(std::vector<ld>& re, std::vector<ld>& im) {
    const size_t n = re.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t i = 0; i < n; i += len) {
            for (size_t k = 0; k < len / 2; k++) {
                const ld a = -2 * kPi * (ld)k / (ld)len;
                const ld wr = cosl(a), wi = sinl(a);
                const size_t u = i + k, v = i + k + len / 2;
                const ld xr = re[v] * wr - im[v] * wi, xi = re[v] * wi + im[v] * wr;
                re[v] = re[u] - xr;
                im[v] = im[u] - xi;
                re[u] += xr;
                im[u] += xi;
            }
        }
    }
}

template <typename T>
int upload(void** d, const std::vector<cpx<T>>& h) {
    if (hipMalloc(d, h.size() * sizeof(cpx<T>)) != hipSuccess) {
        *d = nullptr;
        return PFC_ENOMEM;
    }
    return hipMemcpy(*d, h.data(), h.size() * sizeof(cpx<T>), hipMemcpyHostToDevice) == hipSuccess
               ? PFC_OK
               : PFC_EDEVICE;
}

}  // namespace

template <typename T>
int gen_dim_create(GenDim& g, int L) {
    if (!gen_plan(L, sizeof(cpx<T>), &g.p)) return PFC_ESIZE;
    const int n = g.p.M ? g.p.M : L;
    std::vector<cpx<T>> tw(n);
    for (int m = 0; m < n; m++) {
        const ld a = -2 * kPi * (ld)m / (ld)n;
        tw[m] = {(T)cosl(a), (T)sinl(a)};
    }
    int rc = upload<T>(&g.tw, tw);
    if (rc || !g.p.M) return rc;
    // chirp c[j] = e^{-i pi j^2 / L} (j^2 reduced mod 2L), and
    // B = FFT_M(b) / M with b[m] = conj c[|m|] for |m| < L
    const int M = g.p.M;
    std::vector<cpx<T>> ch(L);
    std::vector<ld> br(M, 0), bi(M, 0);
    for (int j = 0; j < L; j++) {
        const long long q = ((long long)j * j) % (2LL * L);
        const ld a = -kPi * (ld)q / (ld)L;
        ch[j] = {(T)cosl(a), (T)sinl(a)};
        br[j] = cosl(a);
        bi[j] = -sinl(a);
        if (j > 0) {
            br[M - j] = cosl(a);
            bi[M - j] = -sinl(a);
        }
    }
    host_fft(br, bi);
    std::vector<cpx<T>> bh(M);
    for (int k = 0; k < M; k++) bh[k] = {(T)(br[k] / M), (T)(bi[k] / M)};
    rc = upload<T>(&g.chirp, ch);
    if (!rc) rc = upload<T>(&g.bhat, bh);
    return rc;
}

void gen_dim_free(GenDim& g) {
    for (void* p : {g.tw, g.chirp, g.bhat})
        if (p) (void)hipFree(p);
    g.tw = g.chirp = g.bhat = nullptr;
}
template int gen_dim_create<double>(GenDim& g, int L);
template int gen_dim_create<float>(GenDim& g, int L);

namespace {

template <typename T>
GenTables<T> tables(const GenDim& g) {
    return gen_tables<T>(g);
}

// ------------------------------------------------------------------ kernels

// J_b[kx][y] = scale * FFT_x(m_b[y])[kx], kx = 0 .. Nx/2 (one block per row y
// and buffer b)
template <typename T>
__global__ __launch_bounds__(kGenThreads) void k_gen_row_r2c(const T* __restrict__ src,
                                                             cpx<T>* __restrict__ dst, int Nx, int Ny,
                                                             size_t rstride, size_t kstride,
                                                             GenTables<T> gx) {
    extern __shared__ unsigned char smem_r2c[];
    cpx<T>* a = (cpx<T>*)smem_r2c;
    cpx<T>* b = a + gen_lds_elems(gx.p) / 2;
    const int y = blockIdx.x;
    const T* row = src + blockIdx.y * rstride + (size_t)y * Nx;
    for (int x = threadIdx.x; x < Nx; x += blockDim.x) a[x] = {row[x], (T)0};
    const cpx<T>* r = gen_fft<-1, T>(a, b, gx);
    cpx<T>* d = dst + blockIdx.y * kstride;
    for (int k = threadIdx.x; k <= Nx / 2; k += blockDim.x) d[(size_t)k * Ny + y] = r[k];
}

// J_f[kx][y] = FFT_x(nonlinear term of field f of row y)[kx] (one block per
// row y and field f): the nonlinear term of step() (src/pfc-relax.c:
// 1187-1208) from all fields' densities of the row (nonlinear(), as the
// tuned row pass computes it), without writing it back; block (0, 0) makes the NaN check of
// field 0 at (0, 0) (:1175)
template <typename T, int U>
__global__ __launch_bounds__(kGenThreads) void k_gen_row_r2c_nl(
    const T* __restrict__ m, cpx<T>* __restrict__ dst, int Nx, int Ny, size_t rstride,
    size_t kstride, const DevParams* __restrict__ prm, int* __restrict__ nan_step, int tstep,
    GenTables<T> gx) {
    extern __shared__ unsigned char smem_nl[];
    cpx<T>* a = (cpx<T>*)smem_nl;
    cpx<T>* b = a + gen_lds_elems(gx.p) / 2;
    const int y = blockIdx.x, f = blockIdx.y;
    for (int x = threadIdx.x; x < Nx; x += blockDim.x) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = m[u * rstride + (size_t)y * Nx + x];
        if (f == 0 && y == 0 && x == 0 && v[0] != v[0]) atomicMin(nan_step, tstep);
        nonlinear<U, T>(prm, v);
        T w = v[0];
#pragma unroll
        for (int u = 1; u < U; u++) w = u == f ? v[u] : w;
        a[x] = {w, (T)0};
    }
    const cpx<T>* r = gen_fft<-1, T>(a, b, gx);
    cpx<T>* d = dst + f * kstride;
    for (int k = threadIdx.x; k <= Nx / 2; k += blockDim.x) d[(size_t)k * Ny + y] = r[k];
}

// m_b[y] = C2R_x(J_b[.][y]): the row's half spectrum extended by Hermitian
// symmetry, the imaginary parts of kx = 0 and kx = Nx/2 dropped, as a 1D c2r
// does
template <typename T>
__global__ __launch_bounds__(kGenThreads) void k_gen_row_c2r(const cpx<T>* __restrict__ src,
                                                             T* __restrict__ dst, int Nx, int Ny,
                                                             size_t kstride, size_t rstride,
                                                             GenTables<T> gx) {
    extern __shared__ unsigned char smem_c2r[];
    cpx<T>* a = (cpx<T>*)smem_c2r;
    cpx<T>* b = a + gen_lds_elems(gx.p) / 2;
    const int y = blockIdx.x;
    const cpx<T>* s = src + blockIdx.y * kstride;
    for (int k = threadIdx.x; k <= Nx / 2; k += blockDim.x) {
        cpx<T> v = s[(size_t)k * Ny + y];
        if (k == 0 || 2 * k == Nx) v.y = (T)0;
        a[k] = v;
        if (k > 0 && 2 * k < Nx) a[Nx - k] = cconj(v);
    }
    const cpx<T>* r = gen_fft<+1, T>(a, b, gx);
    T* row = dst + blockIdx.y * rstride + (size_t)y * Nx;
    for (int x = threadIdx.x; x < Nx; x += blockDim.x) row[x] = r[x].x;
}

// dst_b column kx = scale * FFT_y (DIR -1) or IFFT_y (DIR +1) of src_b column
// kx (in place allowed: a block reads its whole column first)
template <typename T, int DIR>
__global__ __launch_bounds__(kGenThreads) void k_gen_col_fft(const cpx<T>* src, cpx<T>* dst, int Ny,
                                                             size_t kstride, double scale,
                                                             GenTables<T> gy) {
    extern __shared__ unsigned char smem_col[];
    cpx<T>* a = (cpx<T>*)smem_col;
    cpx<T>* b = a + gen_lds_elems(gy.p) / 2;
    const size_t off = blockIdx.y * kstride + (size_t)blockIdx.x * Ny;
    for (int j = threadIdx.x; j < Ny; j += blockDim.x) a[j] = src[off + j];
    const cpx<T>* r = gen_fft<DIR, T>(a, b, gy);
    for (int j = threadIdx.x; j < Ny; j += blockDim.x) dst[off + j] = cscale(r[j], (T)scale);
}


// k^2 of spectral index idx (column kx = idx / Ny, row j), rounded as the
// reference forms it (src/pfc-relax.c:1217-1222)
__device__ __forceinline__ double gen_k2(size_t idx, int Ny, const DevParams* __restrict__ p) {
    const int kx = (int)(idx / Ny), j = (int)(idx % Ny);
    const double kxd = kx * p->dkx;
    const double ky = ky_of(j, Ny, p->dky);
    return kxd * kxd + ky * ky;
}


// The column pass of a step, per field (one block per column kx and field
// f): FFT_y of J_f, the k-space update of field f (src/pfc-relax.c:
// 1214-1242, kupdate_one: the other fields' old states read from Nsrc, the
// new state of f written to Ndst), in the FP32 form the Hermitian projection
// of columns kx = 0 and Nx/2, then IFFT_y back into J_f.  Replaces a
// forward column transform into K, a separate update sweep and an inverse
// column transform: no K sweep and no update sweep.  The mixed form's k = 0
// state is read from n0src and written to n0dst.
template <typename T, int U>
__global__ __launch_bounds__(kGenThreads) void k_gen_col_step(
    cpx<T>* J, const cpx<T>* __restrict__ Nsrc, cpx<T>* __restrict__ Ndst, int Nx, int Ny,
    size_t nk, const DevParams* __restrict__ prm, const double* __restrict__ n0src,
    double* __restrict__ n0dst, GenTables<T> gy) {
    extern __shared__ unsigned char smem_cs[];
    cpx<T>* a = (cpx<T>*)smem_cs;
    cpx<T>* b = a + gen_lds_elems(gy.p) / 2;
    const int kx = blockIdx.x, f = blockIdx.y;
    const size_t col = (size_t)kx * Ny;
    cpx<T>* Jf = J + f * nk + col;
    for (int j = threadIdx.x; j < Ny; j += blockDim.x) a[j] = Jf[j];
    cpx<T>* r = gen_fft<-1, T>(a, b, gy);
    cpx<T>* o = (r == a) ? b : a;
    const bool herm = sizeof(T) == 4 && (kx == 0 || (Nx % 2 == 0 && 2 * kx == Nx));
    for (int j = threadIdx.x; j < Ny; j += blockDim.x) {
        const size_t idx = col + j;
        cpx<T> N[U];
#pragma unroll
        for (int v = 0; v < U; v++) N[v] = Nsrc[v * nk + idx];
        cpx<T> x;
        if (sizeof(T) == 4 && idx == 0) {
            cpx<double> N0[U];
#pragma unroll
            for (int v = 0; v < U; v++) N0[v] = {n0src[2 * v], n0src[2 * v + 1]};
            const cpx<double> x0 = kupdate_one<U, double>(
                prm, f, 0.0, 1.0, cpx<double>{(double)r[j].x, (double)r[j].y}, N0);
            n0dst[2 * f] = x0.x;
            n0dst[2 * f + 1] = x0.y;
            x = {(T)x0.x, (T)x0.y};
        } else {
            const double k2 = gen_k2(idx, Ny, prm);
            double G = exp(prm->sig_a * k2);
            G *= G;
            x = kupdate_one<U, T>(prm, f, (T)k2, (T)G, r[j], N);
        }
        r[j] = x;
        if (!herm) Ndst[f * nk + idx] = x;
    }
    if (herm) {
        __syncthreads();
        for (int j = threadIdx.x; j < Ny; j += blockDim.x)
            o[j] = cscale(cadd(r[j], cconj(r[j == 0 ? 0 : Ny - j])), (T)0.5);
        __syncthreads();
        for (int j = threadIdx.x; j < Ny; j += blockDim.x) {
            r[j] = o[j];
            Ndst[f * nk + col + j] = o[j];
        }
    }
    const cpx<T>* q = gen_fft<+1, T>(r, o, gy);
    for (int j = threadIdx.x; j < Ny; j += blockDim.x) Jf[j] = q[j];
}


// Evaluate's k-space outputs (src/pfc-relax.c:1048-1077) from the refreshed
// state: O = [P_u N_u (u < U), G N_u (u < U), (c2_uv - k^2)^2 N_v (u < v)]
template <typename T>
__global__ __launch_bounds__(256) void k_gen_eval_mult(const cpx<T>* __restrict__ Nh, cpx<T>* __restrict__ O, size_t nk,
                                int Ny, const DevParams* __restrict__ prm) {
    const int U = prm->U;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nk;
         idx += (size_t)gridDim.x * blockDim.x) {
        const double k2 = gen_k2(idx, Ny, prm);
        int oi = 0;
        for (int u = 0; u < U; u++, oi++)
            O[oi * nk + idx] = cscale(Nh[u * nk + idx], (T)eval_mult(prm, 0, u, 0.0, k2));
        const T g = (T)eval_mult(prm, 1, 0, 0.0, k2);
        for (int u = 0; u < U; u++, oi++) O[oi * nk + idx] = cscale(Nh[u * nk + idx], g);
        for (int u = 0; u < U; u++)
            for (int v = u + 1; v < U; v++, oi++)
                O[oi * nk + idx] = cscale(Nh[v * nk + idx], (T)eval_mult(prm, 2, u, prm->qc2[u][v], k2));
    }
}

// smooth: the spectrum times exp(-k^2 / (2 sigma^2)) / (Nx Ny)
template <typename T>
__global__ __launch_bounds__(256) void k_gen_smooth_mult(cpx<T>* __restrict__ S, size_t nk, int Ny,
                                  const DevParams* __restrict__ prm, double c) {
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nk;
         idx += (size_t)gridDim.x * blockDim.x)
        S[idx] = cscale(S[idx], (T)(exp(-c * gen_k2(idx, Ny, prm)) * prm->inv_NxNy));
}

// ---------------------------------------------------------------- launchers

dim3 grid1(size_t n) {
    size_t b = (n + 255) / 256;
    return dim3((unsigned)(b < 4096 ? (b ? b : 1) : 4096));
}

template <typename T>
struct Gen {
    using C = cpx<T>;
    static size_t lds(const GenDim& g) { return (size_t)gen_lds_elems(g.p) * sizeof(C); }

    static void row_r2c(pfc_ctx* c, const void* src, void* dst, int nbuf) {
        const GenDim& g = ((GenCtx*)c->gen)->x;
        hipLaunchKernelGGL(k_gen_row_r2c<T>, dim3(c->Ny, nbuf), dim3(kGenThreads), lds(g), c->stream,
                           (const T*)src, (C*)dst, c->Nx, c->Ny, c->nr, c->nkn, tables<T>(g));
    }
    static void row_c2r(pfc_ctx* c, const void* src, void* dst, int nbuf) {
        const GenDim& g = ((GenCtx*)c->gen)->x;
        hipLaunchKernelGGL(k_gen_row_c2r<T>, dim3(c->Ny, nbuf), dim3(kGenThreads), lds(g), c->stream,
                           (const C*)src, (T*)dst, c->Nx, c->Ny, c->nkn, c->nr, tables<T>(g));
    }
    static void col_fft(pfc_ctx* c, const void* src, void* dst, int nbuf, int dir, double scale) {
        const GenDim& g = ((GenCtx*)c->gen)->y;
        const dim3 grid(c->Nx / 2 + 1, nbuf);
        if (dir < 0)
            hipLaunchKernelGGL((k_gen_col_fft<T, -1>), grid, dim3(kGenThreads), lds(g), c->stream,
                               (const C*)src, (C*)dst, c->Ny, c->nkn, scale, tables<T>(g));
        else
            hipLaunchKernelGGL((k_gen_col_fft<T, +1>), grid, dim3(kGenThreads), lds(g), c->stream,
                               (const C*)src, (C*)dst, c->Ny, c->nkn, scale, tables<T>(g));
    }
    template <int U>
    static void step_u(pfc_ctx* c, int t) {
        const GenDim& gxd = ((GenCtx*)c->gen)->x;
        hipLaunchKernelGGL((k_gen_row_r2c_nl<T, U>), dim3(c->Ny, U), dim3(kGenThreads), lds(gxd),
                           c->stream, (const T*)c->mreal, (C*)c->J, c->Nx, c->Ny, c->nr, c->nkn,
                           c->dprm, c->nan_flag, t, tables<T>(gxd));
        // the column pass per field into the second state buffer, which then
        // becomes the state (and the mixed form's new k = 0 state is copied
        // over the old one)
        const GenDim& g = ((GenCtx*)c->gen)->y;
        hipLaunchKernelGGL((k_gen_col_step<T, U>), dim3(c->Nx / 2 + 1, U), dim3(kGenThreads), lds(g),
                           c->stream, (C*)c->J, (const C*)c->Nh, (C*)c->Nh2, c->Nx, c->Ny, c->nkn,
                           c->dprm, (const double*)c->n0, c->n0 + 2 * U, tables<T>(g));
        if (sizeof(T) == 4)
            (void)hipMemcpyAsync(c->n0, c->n0 + 2 * U, 2 * U * sizeof(double), hipMemcpyDeviceToDevice,
                                 c->stream);
        void* s = c->Nh;
        c->Nh = c->Nh2;
        c->Nh2 = s;
        row_c2r(c, c->J, c->mreal, c->U);
    }
    static void step(pfc_ctx* c, int t) {
        if (t % 100 == 0) refresh(c);
        switch (c->U) {
            case 1: step_u<1>(c, t); break;
            case 2: step_u<2>(c, t); break;
            case 3: step_u<3>(c, t); break;
            case 4: step_u<4>(c, t); break;
            case 5: step_u<5>(c, t); break;
            case 6: step_u<6>(c, t); break;
            case 7: step_u<7>(c, t); break;
            default: step_u<8>(c, t); break;
        }
    }
    // Nh = FFT2(m) / (Nx Ny) (src/pfc-relax.c:1178-1184, :1124-1128)
    static void refresh(pfc_ctx* c) {
        row_r2c(c, c->mreal, c->J, c->U);
        col_fft(c, c->J, c->Nh, c->U, -1, c->hp.inv_NxNy);
    }
    template <int U>
    static void energy_u(pfc_ctx* c) {
        hipLaunchKernelGGL((k_energy<T, U, kEnergyBS>), dim3(kEnergyBlocks), dim3(kEnergyBS), 0, c->stream,
                           (const T*)c->mreal, (const T*)c->R, c->f, c->nr, c->partials, c->dprm);
    }
    static void evaluate(pfc_ctx* c) {
        refresh(c);
        hipLaunchKernelGGL(k_gen_eval_mult<T>, grid1(c->nkn), dim3(256), 0, c->stream, (const C*)c->Nh,
                           (C*)c->O, c->nkn, c->Ny, c->dprm);
        col_fft(c, c->O, c->O, c->nO, +1, 1.0);
        row_c2r(c, c->O, c->R, c->nO);
        switch (c->U) {
            case 1: energy_u<1>(c); break;
            case 2: energy_u<2>(c); break;
            case 3: energy_u<3>(c); break;
            case 4: energy_u<4>(c); break;
            case 5: energy_u<5>(c); break;
            case 6: energy_u<6>(c); break;
            case 7: energy_u<7>(c); break;
            default: energy_u<8>(c); break;
        }
    }
    template <int U>
    static void sums_u(pfc_ctx* c) {
        launch_energy_sums<T, U>((const T*)c->mreal, (const C*)c->Nh, c->nr, c->nkn, c->Nx, c->Ny,
                                 c->partials, c->partials + c->poff, kEnergyBlocks, c->dprm,
                                 c->stream);
    }
    // the local and spectral energy sums from mreal and Nh
    static void energy_sums(pfc_ctx* c) {
        switch (c->U) {
            case 1: sums_u<1>(c); break;
            case 2: sums_u<2>(c); break;
            case 3: sums_u<3>(c); break;
            case 4: sums_u<4>(c); break;
            case 5: sums_u<5>(c); break;
            case 6: sums_u<6>(c); break;
            case 7: sums_u<7>(c); break;
            default: sums_u<8>(c); break;
        }
    }
    // out = smoothed in (both real, in mreal's layout) via J
    static void smooth(pfc_ctx* c, const void* in, void* out, double sigma) {
        row_r2c(c, in, c->J, 1);
        col_fft(c, c->J, c->J, 1, -1, 1.0);
        hipLaunchKernelGGL(k_gen_smooth_mult<T>, grid1(c->nkn), dim3(256), 0, c->stream, (C*)c->J, c->nkn,
                           c->Ny, c->dprm, 0.5 / sigma / sigma);
        col_fft(c, c->J, c->J, 1, +1, 1.0);
        row_c2r(c, c->J, out, 1);
    }
};

template <typename T>
void set_lds_limits() {
    // dynamic LDS beyond the default per-launch limit
    (void)hipFuncSetAttribute((const void*)k_gen_row_r2c<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kGenMaxLds);
    (void)hipFuncSetAttribute((const void*)k_gen_row_c2r<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kGenMaxLds);
    (void)hipFuncSetAttribute((const void*)k_gen_col_fft<T, -1>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGenMaxLds);
    (void)hipFuncSetAttribute((const void*)k_gen_col_fft<T, +1>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGenMaxLds);
    const void* cs[] = {(const void*)k_gen_col_step<T, 1>, (const void*)k_gen_col_step<T, 2>,
                        (const void*)k_gen_col_step<T, 3>, (const void*)k_gen_col_step<T, 4>,
                        (const void*)k_gen_col_step<T, 5>, (const void*)k_gen_col_step<T, 6>,
                        (const void*)k_gen_col_step<T, 7>, (const void*)k_gen_col_step<T, 8>,
                        (const void*)k_gen_row_r2c_nl<T, 1>, (const void*)k_gen_row_r2c_nl<T, 2>,
                        (const void*)k_gen_row_r2c_nl<T, 3>, (const void*)k_gen_row_r2c_nl<T, 4>,
                        (const void*)k_gen_row_r2c_nl<T, 5>, (const void*)k_gen_row_r2c_nl<T, 6>,
                        (const void*)k_gen_row_r2c_nl<T, 7>, (const void*)k_gen_row_r2c_nl<T, 8>};
    for (const void* k : cs)
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGenMaxLds);
}

}  // namespace

int gen_create(pfc_ctx* c) {
    GenCtx* g = new (std::nothrow) GenCtx();
    if (!g) return PFC_ENOMEM;
    c->gen = g;
    int rc;
    if (c->prec == PFC_FP64) {
        rc = gen_dim_create<double>(g->x, c->Nx);
        if (!rc) rc = gen_dim_create<double>(g->y, c->Ny);
        set_lds_limits<double>();
    } else {
        rc = gen_dim_create<float>(g->x, c->Nx);
        if (!rc) rc = gen_dim_create<float>(g->y, c->Ny);
        set_lds_limits<float>();
    }
    return rc;
}

void gen_destroy(pfc_ctx* c) {
    GenCtx* g = (GenCtx*)c->gen;
    if (!g) return;
    gen_dim_free(g->x);
    gen_dim_free(g->y);
    delete g;
    c->gen = nullptr;
}

void gen_refresh(pfc_ctx* c) {
    if (c->prec == PFC_FP64) Gen<double>::refresh(c);
    else Gen<float>::refresh(c);
}
void gen_step(pfc_ctx* c, int t) {
    if (c->prec == PFC_FP64) Gen<double>::step(c, t);
    else Gen<float>::step(c, t);
}
void gen_evaluate(pfc_ctx* c) {
    if (c->prec == PFC_FP64) Gen<double>::evaluate(c);
    else Gen<float>::evaluate(c);
}
void gen_energy_sums(pfc_ctx* c) {
    if (c->prec == PFC_FP64) Gen<double>::energy_sums(c);
    else Gen<float>::energy_sums(c);
}
void gen_smooth(pfc_ctx* c, const void* in, void* out, double sigma) {
    if (c->prec == PFC_FP64) Gen<double>::smooth(c, in, out, sigma);
    else Gen<float>::smooth(c, in, out, sigma);
}

}  // namespace pfc

// pfc_kernels.hpp -- gfx950 kernels of the pfc-relax step and evaluation.
//
// Spectral-side layout (all device buffers of one field are H x Ny complex,
// H = Nx/2, stored column-major: element (kx, y) at kx*Ny + y):
//   I  -- row transforms of the density, R2C_x(n)(kx, y), the step state
//   J  -- row transforms of the nonlinear term
//   Nh -- the spectral state N(kx, ky)
// The half spectrum has H+1 columns; the two real-along-y columns kx = 0 and
// kx = H are packed into column 0 as (DC) + i (Nyquist), so every buffer and
// every column transform is a power of two.  In k space column 0 therefore
// holds P(ky) = N0(ky) + i NH(ky) for the two Hermitian spectra N0, NH.
//
// One step (reference step(), src/pfc-relax.c:1133-1247) is two kernels:
//   k_row_step  per row y:   I -> C2R_x -> n -> nonlinear -> R2C_x -> J
//   k_col_step  per column:  J -> FFT_y -> M; update with Nh -> Nh'; IFFT_y -> I
// so each field costs 16 B (row) + 32 B (column) of HBM traffic per point in
// FP64, half that in FP32.
#pragma once
#include "fft_device.hpp"

namespace pfc {

constexpr int kMaxU = 4;   // fields supported by the fused kernels

// Occupancy targets (waves per SIMD) handed to the register allocator.
#define PFC_WAVES(n) __attribute__((amdgpu_waves_per_eu(n)))
#ifndef PFC_ROW_WAVES
#define PFC_ROW_WAVES 2
#endif
#ifndef PFC_COL_WAVES
#define PFC_COL_WAVES 2
#endif

// Parameters in device memory, read with uniform (scalar) loads.
struct DevParams {
    int U;
    int W[8];
    int energy_compat;
    double c1;        // 1 if conserved dynamics else 0
    double dt;
    double sig_a;     // -0.5/sigma/sigma
    double NxNy;      // (double)Nx*Ny
    double inv_NxNy;  // 1.0/Nx/Ny
    double dkx, dky;  // 2 pi/(dx Nx), 2 pi/(dy Ny)
    double lambda[8], delta[8];
    double alpha[8][8], beta[8][8], gamma[8][8], epsilon[8][8], mu2[8][8];
    double qc2[8][8];  // coefficient of the energy's beta_uv gradient term
    double nu2[8][8], omega[8][8], phi[8][8];
};

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Remap so each XCD walks a contiguous range of logical indices: neighbouring
// rows then share that XCD's L2, which merges their partial cache lines.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    if (n % 8 != 0) return b;
    return (b % 8) * (n / 8) + b / 8;
}

__device__ __forceinline__ double ky_of(int j, int Ny, double dky) {
    return (j < Ny / 2) ? j * dky : (j - Ny) * dky;
}

// ---------------------------------------------------------------- k space

// A and B of update_AB (src/pfc-relax.c:940-953) for field u at k2.
__device__ __forceinline__ void op_AB(const DevParams* __restrict__ p, int u, double k2,
                                      double& A, double& B) {
    double P = 1.0;
    for (int w = 0; w < p->W[u]; w++) {
        const double pw = p->nu2[u][w] - k2;
        P *= p->omega[u][w] * pw * pw + p->phi[u][w];
    }
    const double a = p->alpha[u][u] + p->beta[u][u] * P;
    const double o = p->c1 - 1.0 - p->c1 * k2;
    const double ex = exp(o * a * p->dt);
    A = ex;
    const double b = (a == 0.0) ? o * p->dt : (ex - 1.0) / a;
    B = b * p->inv_NxNy;
}

// FP32 variant: same operator, B from expm1 so that the long-wavelength
// modes, where exp(x)-1 cancels to a few bits in FP32, keep full precision.
__device__ __forceinline__ void op_AB(const DevParams* __restrict__ p, int u, float k2,
                                      float& A, float& B) {
    float P = 1.0f;
    for (int w = 0; w < p->W[u]; w++) {
        const float pw = (float)p->nu2[u][w] - k2;
        P *= (float)p->omega[u][w] * pw * pw + (float)p->phi[u][w];
    }
    const float a = (float)p->alpha[u][u] + (float)p->beta[u][u] * P;
    const float c1 = (float)p->c1;
    const float o = c1 - 1.0f - c1 * k2;
    const float x = o * a * (float)p->dt;
    const float em1 = expm1f(x);
    A = 1.0f + em1;
    const float b = (a == 0.0f) ? o * (float)p->dt : em1 / a;
    B = b * (float)p->inv_NxNy;
}

// The k-space update of step() (src/pfc-relax.c:1214-1242) at one k point:
// explicit couplings into M, then N <- A N + B M.  On return N holds the new
// state (the reference also copies it into M).
template <int U, typename T>
__device__ __forceinline__ void kupdate(const DevParams* __restrict__ p, T k2, cpx<T> (&M)[U],
                                        cpx<T> (&N)[U]) {
    const T NxNy = (T)p->NxNy;
    if constexpr (U > 1) {
        T G2 = exp((T)p->sig_a * k2);
        G2 *= G2;
#pragma unroll
        for (int u = 0; u < U - 1; u++) {
#pragma unroll
            for (int v = u + 1; v < U; v++) {
                T d2 = (T)p->mu2[u][v] - k2;
                d2 *= d2;
                const T bd = (T)p->beta[u][v] * d2;
                const T eg = (T)p->epsilon[u][v] * G2;
                M[u].x += (bd * N[v].x + eg * N[v].x) * NxNy;
                M[u].y += (bd * N[v].y + eg * N[v].y) * NxNy;
                M[v].x += (bd * N[u].x + eg * N[u].x) * NxNy;
                M[v].y += (bd * N[u].y + eg * N[u].y) * NxNy;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        T A, B;
        op_AB(p, u, k2, A, B);
        N[u].x = A * N[u].x + B * M[u].x;
        N[u].y = A * N[u].y + B * M[u].y;
    }
}

// ------------------------------------------------------- real-data helpers

// C2R pre-processing of one row held in natural order (thread t holds
// X[t + q*TPB], X[0] packed as (DC, Nyquist)): produces Z with
// IFFT_H(Z)[n] = x[2n] + i x[2n+1], x the unnormalised c2r of X.
template <typename T, int H, int TPB, int LS>
__device__ __forceinline__ void c2r_pre(cpx<T> (&v)[H / TPB], cpx<T>* lds,
                                        const cpx<T>* __restrict__ twx, int t) {
    constexpr int P = H / TPB;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < P; q++) lds[(t + q * TPB) * LS] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < P; q++) {
        const int k = t + q * TPB;
        if (k == 0) {
            const T x0 = v[q].x, xh = v[q].y;
            v[q] = {x0 + xh, x0 - xh};
        } else {
            const cpx<T> a = v[q];
            const cpx<T> b = cconj(lds[(H - k) * LS]);
            const cpx<T> xe = cadd(a, b);
            const cpx<T> xo = cmulc(csub(a, b), twx[k]);   // * e^{+2 pi i k/Nx}
            v[q] = {xe.x - xo.y, xe.y + xo.x};              // xe + i xo
        }
    }
}

// R2C post-processing: v holds Z = FFT_H(z), z[n] = x[2n] + i x[2n+1], in
// natural order; produces X[k] (k < H) with X[0] packed as (DC, Nyquist).
template <typename T, int H, int TPB, int LS>
__device__ __forceinline__ void r2c_post(cpx<T> (&v)[H / TPB], cpx<T>* lds,
                                         const cpx<T>* __restrict__ twx, int t) {
    constexpr int P = H / TPB;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < P; q++) lds[(t + q * TPB) * LS] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < P; q++) {
        const int k = t + q * TPB;
        if (k == 0) {
            const T re = v[q].x, im = v[q].y;
            v[q] = {re + im, re - im};
        } else {
            const cpx<T> a = v[q];
            const cpx<T> b = cconj(lds[(H - k) * LS]);
            const cpx<T> s = cadd(a, b);   // 2 Xe
            const cpx<T> d = csub(a, b);   // 2i Xo
            const cpx<T> xo2 = {d.y, -d.x};   // 2 Xo
            const cpx<T> r = cadd(s, cmul(xo2, twx[k]));
            v[q] = cscale(r, (T)0.5);
        }
    }
}

// -------------------------------------------------------------- nonlinear

// Pointwise nonlinear term of step() (src/pfc-relax.c:1187-1208); all mc
// terms use the incoming values.
template <int U, typename T>
__device__ __forceinline__ void nonlinear(const DevParams* __restrict__ p, T (&m)[U]) {
    T mc[U];
#pragma unroll
    for (int u = 0; u < U; u++) mc[u] = (T)0;
#pragma unroll
    for (int u = 0; u < U - 1; u++) {
#pragma unroll
        for (int v = u + 1; v < U; v++) {
            const T a = (T)p->alpha[u][v], g = (T)p->gamma[u][v];
            mc[u] += a * m[v] + (T)0.5 * g * ((T)2.0 * m[u] * m[v] + m[v] * m[v]);
            mc[v] += a * m[u] + (T)0.5 * g * ((T)2.0 * m[v] * m[u] + m[u] * m[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const T m2 = m[u] * m[u];
        m[u] = (T)p->lambda[u] + (T)p->gamma[u][u] * m2 + (T)p->delta[u] * m[u] * m2 + mc[u];
    }
}

// ================================================================ kernels

// Row pass of a step.  Block = TR rows, lane = (fft thread t, row r) with the
// row index fastest, so the TR rows' elements of one kx are adjacent lanes
// and adjacent addresses in the column-major buffers.
template <typename T, int NX, int TPB, int TR, int U, bool REFRESH>
__global__ __launch_bounds__(TPB* TR) PFC_WAVES(PFC_ROW_WAVES) void k_row_step(
    const cpx<T>* __restrict__ I, cpx<T>* __restrict__ J, cpx<T>* __restrict__ K,
    size_t fstride, int Ny, const cpx<T>* __restrict__ twx, const DevParams* __restrict__ prm,
    int* __restrict__ nan_step, int tstep) {
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    using FFTF = BlockFFT<T, H, TPB, -1, 2, TR>;
    using FFTI = BlockFFT<T, H, TPB, +1, 2, TR>;
    __shared__ cpx<T> lds[H * TR];
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    const int y = xcd_remap(blockIdx.x, gridDim.x) * TR + r;
    cpx<T>* L = lds + r;

    cpx<T> m[U][P];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const cpx<T>* src = I + u * fstride + y;
#pragma unroll
        for (int q = 0; q < P; q++) m[u][q] = src[(size_t)(t + q * TPB) * Ny];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        c2r_pre<T, H, TPB, TR>(m[u], L, twx, t);
        FFTI::run(m[u], L, twx, t);
    }
    // NaN check on field 0 at (0,0) (src/pfc-relax.c:1175)
    if (y == 0 && t == 0) {
        const T v0 = m[0][0].x;
        if (v0 != v0) atomicMin(nan_step, tstep);
    }
    if constexpr (REFRESH) {
        // R2C of the density itself, for N = FFT(n)/NxNy (src/pfc-relax.c:1178-1184)
#pragma unroll
        for (int u = 0; u < U; u++) {
            cpx<T> w[P];
#pragma unroll
            for (int q = 0; q < P; q++) w[q] = m[u][q];
            FFTF::run(w, L, twx, t);
            r2c_post<T, H, TPB, TR>(w, L, twx, t);
            cpx<T>* dst = K + u * fstride + y;
#pragma unroll
            for (int q = 0; q < P; q++) dst[(size_t)(t + q * TPB) * Ny] = w[q];
        }
    }
#pragma unroll
    for (int q = 0; q < P; q++) {
        T a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            a[u] = m[u][q].x;
            b[u] = m[u][q].y;
        }
        nonlinear<U, T>(prm, a);
        nonlinear<U, T>(prm, b);
#pragma unroll
        for (int u = 0; u < U; u++) m[u][q] = {a[u], b[u]};
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        FFTF::run(m[u], L, twx, t);
        r2c_post<T, H, TPB, TR>(m[u], L, twx, t);
        cpx<T>* dst = J + u * fstride + y;
#pragma unroll
        for (int q = 0; q < P; q++) dst[(size_t)(t + q * TPB) * Ny] = m[u][q];
    }
}

// Unpacks the two Hermitian spectra of packed column 0 at ky index j:
// P(j) = A(j) + i B(j); pj is P at the partner index (Ny - j) mod Ny.
template <typename T>
__device__ __forceinline__ void unpack0(cpx<T> pv, cpx<T> pj, cpx<T>& a, cpx<T>& b) {
    const cpx<T> c = cconj(pj);
    a = cscale(cadd(pv, c), (T)0.5);
    const cpx<T> d = csub(pv, c);      // 2i B
    b = {(T)0.5 * d.y, (T)-0.5 * d.x};
}

// Column pass of a step for columns kx >= 1: TC columns per block (lane =
// column-major within the block), each column all Ny rows.  Lanes of the
// packed column 0 (block 0, c = 0) only take part in the barriers; that
// column is handled by k_col0_step.
template <typename T, int NY, int TPB, int TC, int U, bool REFRESH>
__global__ __launch_bounds__(TPB* TC) PFC_WAVES(PFC_COL_WAVES) void k_col_step(
    const cpx<T>* __restrict__ J, const cpx<T>* __restrict__ K, cpx<T>* __restrict__ Nh,
    cpx<T>* __restrict__ I, size_t fstride, const cpx<T>* __restrict__ twy,
    const DevParams* __restrict__ prm) {
    constexpr int P = NY / TPB;
    using FFTF = BlockFFT<T, NY, TPB, -1, 1, 1>;
    using FFTI = BlockFFT<T, NY, TPB, +1, 1, 1>;
    __shared__ cpx<T> lds[NY * TC];
    const int c = threadIdx.x / TPB;
    const int t = threadIdx.x % TPB;
    const int kx = blockIdx.x * TC + c;
    const bool live = (kx != 0);
    cpx<T>* L = lds + c * NY;
    const size_t col = (size_t)kx * NY;

    cpx<T> M[U][P];
#pragma unroll
    for (int u = 0; u < U; u++) {
#pragma unroll
        for (int q = 0; q < P; q++) M[u][q] = J[u * fstride + col + t + q * TPB];
        FFTF::run(M[u], L, twy, t);
    }
    if constexpr (REFRESH) {
        // N = FFT(n)/NxNy from the density's row transforms
        const T s = (T)prm->inv_NxNy;
#pragma unroll
        for (int u = 0; u < U; u++) {
            cpx<T> w[P];
#pragma unroll
            for (int q = 0; q < P; q++) w[q] = K[u * fstride + col + t + q * TPB];
            FFTF::run(w, L, twy, t);
            if (live) {
#pragma unroll
                for (int q = 0; q < P; q++) Nh[u * fstride + col + t + q * TPB] = cscale(w[q], s);
            }
        }
    }
    if (live) {
        const double dky = prm->dky;
        const double kxd = kx * prm->dkx;
        const double kx2d = kxd * kxd;
#pragma unroll
        for (int q = 0; q < P; q++) {
            const int j = t + q * TPB;
            const double ky = ky_of(j, NY, dky);
            const T k2 = (T)(kx2d + ky * ky);
            cpx<T> Nk[U];
#pragma unroll
            for (int u = 0; u < U; u++) Nk[u] = Nh[u * fstride + col + j];
            cpx<T> Mk[U];
#pragma unroll
            for (int u = 0; u < U; u++) Mk[u] = M[u][q];
            kupdate<U, T>(prm, k2, Mk, Nk);
#pragma unroll
            for (int u = 0; u < U; u++) {
                Nh[u * fstride + col + j] = Nk[u];
                M[u][q] = Nk[u];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        FFTI::run(M[u], L, twy, t);
        if (live) {
#pragma unroll
            for (int q = 0; q < P; q++) I[u * fstride + col + t + q * TPB] = M[u][q];
        }
    }
}

// Column 0 of a step: the packed pair of Hermitian columns kx = 0 and
// kx = H, P(ky) = N0(ky) + i NH(ky).  One block.  The forward transforms are
// staged in the scratch S (2*U*Ny elements) so partners (Ny - ky) can be read.
template <typename T, int NY, int TPB, int U, bool REFRESH>
__global__ __launch_bounds__(TPB) void k_col0_step(const cpx<T>* __restrict__ J,
                                                   const cpx<T>* __restrict__ K,
                                                   cpx<T>* __restrict__ Nh,
                                                   cpx<T>* __restrict__ I, cpx<T>* __restrict__ S,
                                                   size_t fstride, int H,
                                                   const cpx<T>* __restrict__ twy,
                                                   const DevParams* __restrict__ prm) {
    constexpr int P = NY / TPB;
    using FFTF = BlockFFT<T, NY, TPB, -1, 1, 1>;
    using FFTI = BlockFFT<T, NY, TPB, +1, 1, 1>;
    __shared__ cpx<T> lds[NY];
    const int t = threadIdx.x;
    cpx<T>* Ms = S;             // packed M of every field
    cpx<T>* Ps = S + U * NY;    // packed new state of every field
    for (int u = 0; u < U; u++) {
        cpx<T> w[P];
#pragma unroll
        for (int q = 0; q < P; q++) w[q] = J[u * fstride + t + q * TPB];
        FFTF::run(w, lds, twy, t);
#pragma unroll
        for (int q = 0; q < P; q++) Ms[u * NY + t + q * TPB] = w[q];
        if constexpr (REFRESH) {
#pragma unroll
            for (int q = 0; q < P; q++) w[q] = K[u * fstride + t + q * TPB];
            FFTF::run(w, lds, twy, t);
            const T s = (T)prm->inv_NxNy;
#pragma unroll
            for (int q = 0; q < P; q++) Nh[u * fstride + t + q * TPB] = cscale(w[q], s);
        }
    }
    __syncthreads();
    const double dky = prm->dky;
    const double kxh = (double)H * prm->dkx;
    const double kxh2 = kxh * kxh;
#pragma unroll
    for (int q = 0; q < P; q++) {
        const int j = t + q * TPB;
        const int jp = (NY - j) & (NY - 1);
        const double ky = ky_of(j, NY, dky);
        cpx<T> M0[U], MH[U], N0[U], NH[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            unpack0(Ms[u * NY + j], Ms[u * NY + jp], M0[u], MH[u]);
            unpack0(Nh[u * fstride + j], Nh[u * fstride + jp], N0[u], NH[u]);
        }
        kupdate<U, T>(prm, (T)(ky * ky), M0, N0);
        kupdate<U, T>(prm, (T)(kxh2 + ky * ky), MH, NH);
#pragma unroll
        for (int u = 0; u < U; u++) Ps[u * NY + j] = {N0[u].x - NH[u].y, N0[u].y + NH[u].x};
    }
    __syncthreads();   // all reads of the old state are done
    for (int u = 0; u < U; u++) {
        cpx<T> w[P];
#pragma unroll
        for (int q = 0; q < P; q++) {
            w[q] = Ps[u * NY + t + q * TPB];
            Nh[u * fstride + t + q * TPB] = w[q];
        }
        FFTI::run(w, lds, twy, t);
#pragma unroll
        for (int q = 0; q < P; q++) I[u * fstride + t + q * TPB] = w[q];
    }
}

// ------------------------------------------------ single-buffer row passes

// Row C2R of nbuf buffers (blockIdx.y selects the buffer): src spectral ->
// dst real rows (Ny x Nx, row-major).  If K is given, also writes R2C of the
// result (the density's row transform, used for N = FFT(n)/NxNy).
template <typename T, int NX, int TPB, int TR>
__global__ __launch_bounds__(TPB* TR) void k_row_c2r(const cpx<T>* __restrict__ src,
                                                     T* __restrict__ dst, cpx<T>* __restrict__ K,
                                                     size_t fstride, size_t rstride, int Ny,
                                                     const cpx<T>* __restrict__ twx) {
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    __shared__ cpx<T> lds[H * TR];
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    const int y = xcd_remap(blockIdx.x, gridDim.x) * TR + r;
    const int b = blockIdx.y;
    cpx<T>* L = lds + r;
    cpx<T> v[P];
    const cpx<T>* s = src + b * fstride + y;
#pragma unroll
    for (int q = 0; q < P; q++) v[q] = s[(size_t)(t + q * TPB) * Ny];
    c2r_pre<T, H, TPB, TR>(v, L, twx, t);
    BlockFFT<T, H, TPB, +1, 2, TR>::run(v, L, twx, t);
    cpx<T>* d = reinterpret_cast<cpx<T>*>(dst + b * rstride + (size_t)y * NX);
#pragma unroll
    for (int q = 0; q < P; q++) d[t + q * TPB] = v[q];
    if (K) {
        BlockFFT<T, H, TPB, -1, 2, TR>::run(v, L, twx, t);
        r2c_post<T, H, TPB, TR>(v, L, twx, t);
        cpx<T>* k = K + b * fstride + y;
#pragma unroll
        for (int q = 0; q < P; q++) k[(size_t)(t + q * TPB) * Ny] = v[q];
    }
}

// Row R2C of nbuf buffers: real rows -> spectral, times scale (1/Nx gives
// the row state I, whose C2R is the density itself).
template <typename T, int NX, int TPB, int TR>
__global__ __launch_bounds__(TPB* TR) void k_row_r2c(const T* __restrict__ src,
                                                     cpx<T>* __restrict__ dst, size_t fstride,
                                                     size_t rstride, int Ny,
                                                     const cpx<T>* __restrict__ twx, T scale) {
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    __shared__ cpx<T> lds[H * TR];
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    const int y = xcd_remap(blockIdx.x, gridDim.x) * TR + r;
    const int b = blockIdx.y;
    cpx<T>* L = lds + r;
    cpx<T> v[P];
    const cpx<T>* s = reinterpret_cast<const cpx<T>*>(src + b * rstride + (size_t)y * NX);
#pragma unroll
    for (int q = 0; q < P; q++) v[q] = s[t + q * TPB];
    BlockFFT<T, H, TPB, -1, 2, TR>::run(v, L, twx, t);
    r2c_post<T, H, TPB, TR>(v, L, twx, t);
    cpx<T>* d = dst + b * fstride + y;
#pragma unroll
    for (int q = 0; q < P; q++) d[(size_t)(t + q * TPB) * Ny] = cscale(v[q], scale);
}

// Column forward FFT of nbuf buffers with a scale: dst = scale * FFT_y(src).
template <typename T, int NY, int TPB, int TC>
__global__ __launch_bounds__(TPB* TC) void k_col_fwd(const cpx<T>* __restrict__ src,
                                                     cpx<T>* __restrict__ dst, size_t fstride,
                                                     const cpx<T>* __restrict__ twy, double scale) {
    constexpr int P = NY / TPB;
    __shared__ cpx<T> lds[NY * TC];
    const int c = threadIdx.x / TPB;
    const int t = threadIdx.x % TPB;
    const int kx = blockIdx.x * TC + c;
    const int b = blockIdx.y;
    cpx<T>* L = lds + c * NY;
    const size_t col = b * fstride + (size_t)kx * NY;
    cpx<T> v[P];
#pragma unroll
    for (int q = 0; q < P; q++) v[q] = src[col + t + q * TPB];
    BlockFFT<T, NY, TPB, -1, 1, 1>::run(v, L, twy, t);
#pragma unroll
    for (int q = 0; q < P; q++) dst[col + t + q * TPB] = cscale(v[q], (T)scale);
}

// Multiplier of one evaluate output at k2 (src/pfc-relax.c:1048-1077):
// kind 0: P_u(k) (the multimode operator), 1: G(k) (Gaussian),
// 2: (c2 - k2)^2 (gradient coupling, c2 = mu2 or the reference's nu2 alias).
__device__ __forceinline__ double eval_mult(const DevParams* __restrict__ p, int kind, int u,
                                            double c2, double k2) {
    if (kind == 0) {
        double P = 1.0;
        for (int w = 0; w < p->W[u]; w++) {
            const double pw = p->nu2[u][w] - k2;
            P *= p->omega[u][w] * pw * pw + p->phi[u][w];
        }
        return P;
    } else if (kind == 1) {
        return exp(p->sig_a * k2);
    } else {
        const double d = c2 - k2;
        return d * d;
    }
}

// Evaluate's k-space pass for source field s = blockIdx.y: M = FFT_y(K_s)/NxNy
// (also stored as the refreshed state Nh_s, src/pfc-relax.c:1124-1128), then
// every output that multiplies M_s -- P_s M_s, G M_s and (c2_us - k2)^2 M_s
// for u < s -- is transformed back along y into O.
// O buffer order: [P_0..P_{U-1}, G_0..G_{U-1}, q_uv for u<v in row order].
template <typename T, int NY, int TPB, int TC>
__global__ __launch_bounds__(TPB* TC) void k_col_eval(const cpx<T>* __restrict__ K,
                                                      cpx<T>* __restrict__ Nh,
                                                      cpx<T>* __restrict__ O, size_t fstride,
                                                      int H, const cpx<T>* __restrict__ twy,
                                                      const DevParams* __restrict__ prm) {
    constexpr int P = NY / TPB;
    __shared__ cpx<T> lds[NY * TC];
    const int c = threadIdx.x / TPB;
    const int t = threadIdx.x % TPB;
    const int kx = blockIdx.x * TC + c;
    const int s = blockIdx.y;
    const int U = prm->U;
    cpx<T>* L = lds + c * NY;
    const size_t col = (size_t)kx * NY;
    const T scale = (T)prm->inv_NxNy;
    cpx<T> M[P];
#pragma unroll
    for (int q = 0; q < P; q++) M[q] = K[s * fstride + col + t + q * TPB];
    BlockFFT<T, NY, TPB, -1, 1, 1>::run(M, L, twy, t);
#pragma unroll
    for (int q = 0; q < P; q++) {
        M[q] = cscale(M[q], scale);
        Nh[s * fstride + col + t + q * TPB] = M[q];
    }
    // partner values for the packed column 0 (block 0 only; all lanes take part)
    cpx<T> Mp[P];
    if (blockIdx.x == 0) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < P; q++) L[t + q * TPB] = M[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < P; q++) Mp[q] = L[(NY - (t + q * TPB)) & (NY - 1)];
    }
    const double dky = prm->dky;
    const double kxd = kx * prm->dkx;
    const double kx2 = kxd * kxd;
    const double kxh = (double)H * prm->dkx;
    const double kxh2 = kxh * kxh;
    // outputs with source s: (kind, field index for the multiplier, c2, out index)
    const int nout = 2 + s;
    for (int o = 0; o < nout; o++) {
        int kind, fu, oi;
        double c2 = 0.0;
        if (o == 0) { kind = 0; fu = s; oi = s; }
        else if (o == 1) { kind = 1; fu = s; oi = U + s; }
        else {
            const int u = o - 2;   // pair (u, s), u < s
            kind = 2; fu = u;
            c2 = prm->qc2[u][s];
            oi = 2 * U + u * U - u * (u + 1) / 2 + (s - u - 1);
        }
        cpx<T> X[P];
#pragma unroll
        for (int q = 0; q < P; q++) {
            const int j = t + q * TPB;
            const double ky = ky_of(j, NY, dky);
            if (kx == 0) {
                cpx<T> a, b;
                unpack0(M[q], Mp[q], a, b);
                const T f0 = (T)eval_mult(prm, kind, fu, c2, ky * ky);
                const T fh = (T)eval_mult(prm, kind, fu, c2, kxh2 + ky * ky);
                a = cscale(a, f0);
                b = cscale(b, fh);
                X[q] = {a.x - b.y, a.y + b.x};
            } else {
                X[q] = cscale(M[q], (T)eval_mult(prm, kind, fu, c2, kx2 + ky * ky));
            }
        }
        BlockFFT<T, NY, TPB, +1, 1, 1>::run(X, L, twy, t);
#pragma unroll
        for (int q = 0; q < P; q++) O[oi * fstride + col + t + q * TPB] = X[q];
    }
}

// ------------------------------------------------------------- reductions

// Compensated accumulation (Neumaier); pairs combine with TwoSum.
struct KSum {
    double s, c;
};
__device__ __forceinline__ void kadd(KSum& k, double x) {
    const double t = k.s + x;
    if (fabs(k.s) >= fabs(x)) k.c += (k.s - t) + x;
    else k.c += (x - t) + k.s;
    k.s = t;
}
__device__ __forceinline__ KSum kmerge(KSum a, KSum b) {
    KSum r;
    r.s = a.s + b.s;
    const double bb = r.s - a.s;
    const double e = (a.s - (r.s - bb)) + (b.s - bb);
    r.c = a.c + b.c + e;
    return r;
}
__device__ __forceinline__ KSum kshfl_down(KSum a, int off) {
    KSum r;
    r.s = __shfl_down(a.s, off, 64);
    r.c = __shfl_down(a.c, off, 64);
    return r;
}

// Block reduction of NV compensated sums (wavefront shuffles, then an LDS
// tree over the block's waves); thread 0 returns the totals.
template <int NV, int BS>
__device__ __forceinline__ void block_ksum(KSum (&v)[NV]) {
    __shared__ KSum sh[BS / 64][NV];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[i] = kmerge(v[i], kshfl_down(v[i], off));
        if (lane == 0) sh[wv][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) {
            KSum a = sh[0][i];
            for (int w = 1; w < BS / 64; w++) a = kmerge(a, sh[w][i]);
            v[i] = a;
        }
    }
}

// Pointwise free-energy density (src/pfc-relax.c:1093-1117) with per-block
// compensated partial sums of f and of every n_u.  R holds the real-space
// evaluate outputs in O order; n the densities.
template <typename T, int U, int BS>
__global__ __launch_bounds__(BS) void k_energy(const T* __restrict__ n, const T* __restrict__ R,
                                               double* __restrict__ f, size_t nr,
                                               double* __restrict__ partials,
                                               const DevParams* __restrict__ p) {
    KSum acc[1 + U];
#pragma unroll
    for (int i = 0; i <= U; i++) acc[i] = {0.0, 0.0};
    const double _3 = 1.0 / 3.0;
    for (size_t ij = (size_t)blockIdx.x * BS + threadIdx.x; ij < nr; ij += (size_t)gridDim.x * BS) {
        double nv[U], e[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            nv[u] = (double)n[u * nr + ij];
            e[u] = (double)R[(U + u) * nr + ij];
        }
        double fij = 0.0;
        int pq = 2 * U;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const double nu = nv[u];
            const double nunu = nu * nu;
            const double Pn = (double)R[u * nr + ij];
            fij += p->lambda[u] * nu + 0.5 * p->alpha[u][u] * nunu + 0.5 * p->beta[u][u] * nu * Pn +
                   _3 * p->gamma[u][u] * nu * nunu + 0.25 * p->delta[u] * nunu * nunu;
#pragma unroll
            for (int v = u + 1; v < U; v++) {
                const double nvv = nv[v];
                const double nvnv = nvv * nvv;
                const double q = (double)R[pq * nr + ij];
                pq++;
                fij += p->alpha[u][v] * nu * nvv + p->beta[u][v] * nu * q +
                       0.5 * p->gamma[u][v] * (nunu * nvv + nu * nvnv) +
                       p->epsilon[u][v] * e[u] * e[v];
            }
            kadd(acc[1 + u], nu);
        }
        f[ij] = fij;
        kadd(acc[0], fij);
    }
    block_ksum<1 + U, BS>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i <= U; i++) {
            partials[(size_t)blockIdx.x * 2 * (1 + U) + 2 * i] = acc[i].s;
            partials[(size_t)blockIdx.x * 2 * (1 + U) + 2 * i + 1] = acc[i].c;
        }
    }
}

// Second stage: one block folds nblk partial pairs per quantity.
template <int BS>
__global__ __launch_bounds__(BS) void k_reduce_final(const double* __restrict__ partials,
                                                     int nblk, int nq, double* __restrict__ out,
                                                     double scale) {
    for (int i = 0; i < nq; i++) {
        KSum a[1] = {{0.0, 0.0}};
        for (int b = threadIdx.x; b < nblk; b += BS) {
            KSum x = {partials[(size_t)b * 2 * nq + 2 * i], partials[(size_t)b * 2 * nq + 2 * i + 1]};
            a[0] = kmerge(a[0], x);
        }
        block_ksum<1, BS>(a);
        if (threadIdx.x == 0) out[i] = (a[0].s + a[0].c) * scale;
        __syncthreads();
    }
}

}  // namespace pfc

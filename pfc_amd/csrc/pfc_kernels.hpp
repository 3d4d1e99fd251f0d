// pfc_kernels.hpp -- gfx950 kernels of the pfc-relax step and evaluation.
//
// Spectral-side layout.  Row-spectral buffers of one field (I, J, K, O) hold
// H = Nx/2 columns of Ny complex values, column-major (element (kx, y) at
// kx*Ny + y); the R2C of a real row has real kx = 0 and kx = H entries, so
// column 0 carries them as (DC, Nyquist) = (re, im).  The spectral state Nh
// holds the half spectrum's H+1 columns unpacked (kx = 0..H), the transposed
// layout of the reference's FFTW_MPI_TRANSPOSED_OUT transforms
// (src/pfc-relax.c:227-235):
//   I  -- row transforms of the density, R2C_x(n)(kx, y)/Nx, the step state
//   J  -- row transforms of the nonlinear term (unnormalised)
//   Nh -- the spectral state N(kx, ky)
// One step (reference step(), src/pfc-relax.c:1133-1247) is two kernels:
//   k_row_step  per row y:   I -> C2R_x -> n -> nonlinear -> R2C_x -> J
//   k_col_step  per column:  J -> FFT_y -> M; update with Nh -> Nh'; IFFT_y -> I
// so each field costs 16 B (row) + 32 B (column) of HBM traffic per point in
// FP64, half that in FP32.  Block 0 of a column pass runs the kx = 0 and
// kx = H columns one after the other from the packed column 0, keeping only
// the real parts of their inverse transforms, which is what a 1D c2r uses.
#pragma once
#include "fft_device.hpp"

namespace pfc {

constexpr int kMaxU = 4;   // fields supported by the fused kernels

// Occupancy targets (waves per SIMD) handed to the register allocator, per
// element precision.
#define PFC_WAVES(n) __attribute__((amdgpu_waves_per_eu(n)))
#ifndef PFC_PROBE
#define PFC_PROBE 0
#endif
// Cache policy of the step passes' stores (0, or 2 = non-temporal).  The
// column pass writes whole columns that no later block of the pass reads, so
// its stores stream past L2 (1-2 % faster step).  The row pass's 32-byte
// pieces of J must stay cacheable: L2 merges them with the neighbouring row
// groups' pieces of the same lines; non-temporal they are 1.6x (mixed) to
// 2.4x (FP64) slower.
#ifndef PFC_ROWJ_AUX
#define PFC_ROWJ_AUX 0
#endif
#ifndef PFC_COL_AUX
#define PFC_COL_AUX 2
#endif
#ifndef PFC_STEP_PADC
#define PFC_STEP_PADC 3
#endif
#ifndef PFC_STEP_PADC64
#define PFC_STEP_PADC64 0
#endif
#ifndef PFC_ROW_WAVES_F64
#define PFC_ROW_WAVES_F64 2
#endif
#ifndef PFC_ROW_WAVES_F32
#define PFC_ROW_WAVES_F32 2
#endif
#ifndef PFC_COL_WAVES_F64
#define PFC_COL_WAVES_F64 2
#endif
#ifndef PFC_COL_WAVES_F32
#define PFC_COL_WAVES_F32 2
#endif
// FP32 rows of 4096 (2048, 8192) points run 2 (4, 1) rows per block at 4 waves per
// SIMD, so that two blocks share a CU and one block's loads and stores
// overlap the other's transforms (the 4096-point kernel needs 123 VGPRs;
// mixed step 383 -> 363 us, though 2 rows' 16-byte runs copy slower than 4
// rows' 32-byte runs; the batched-relaxation step 1.38 -> 1.35 ms)
#ifndef PFC_ROW_WAVES_F32_4096
#define PFC_ROW_WAVES_F32_4096 4
#endif
#ifndef PFC_ROW_WAVES_F32_2048
#define PFC_ROW_WAVES_F32_2048 4
#endif
#ifndef PFC_ROW_WAVES_F32_8192
#define PFC_ROW_WAVES_F32_8192 4
#endif
#ifndef PFC_ROW_WAVES_F64_4096
#define PFC_ROW_WAVES_F64_4096 2
#endif
template <typename T, int NX = 0>
constexpr int row_waves() {
    return sizeof(T) == 8 ? (NX == 4096 ? PFC_ROW_WAVES_F64_4096 : PFC_ROW_WAVES_F64)
                          : (NX == 4096   ? PFC_ROW_WAVES_F32_4096
                             : NX == 2048 ? PFC_ROW_WAVES_F32_2048
                             : NX == 8192 ? PFC_ROW_WAVES_F32_8192
                                          : PFC_ROW_WAVES_F32);
}
template <typename T>
constexpr int col_waves() { return sizeof(T) == 8 ? PFC_COL_WAVES_F64 : PFC_COL_WAVES_F32; }

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
#if PFC_PROBE
    unsigned long long* probe;   // per-block phase timestamps (diagnostic builds)
#endif
};

#if PFC_PROBE
// Diagnostic builds: lane 0 of each block records the 100 MHz wall clock at
// phase boundaries (after a barrier) and its hardware ids into probe.
#define PFC_PROBE_AT(base, k)                                                      \
    do {                                                                           \
        __syncthreads();                                                           \
        if (threadIdx.x == 0) {                                                    \
            prm->probe[(base) + (size_t)blockIdx.x * 8 + (k)] = wall_clock64();    \
            if ((k) == 0)                                                          \
                prm->probe[(base) + (size_t)blockIdx.x * 8 + 7] =                  \
                    ((unsigned long long)__builtin_amdgcn_s_getreg(63508) << 32) | \
                    (unsigned)__builtin_amdgcn_s_getreg(63492);                    \
        }                                                                          \
    } while (0)
#else
#define PFC_PROBE_AT(base, k) \
    do {                      \
    } while (0)
#endif

// Gaussian factor table of the column pass (per precision, set by
// pfc_set_box): exp(-kx^2/sigma^2) for kx = 0..H at [0, H], then
// exp(-ky^2/sigma^2) for the Ny row indices from ktab_gy(H), 16-byte aligned.
__host__ __device__ constexpr int ktab_gy(int H) { return (H + 1 + 3) & ~3; }

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Remap so each XCD walks a contiguous range of logical indices: neighbouring
// rows then share that XCD's L2, which merges their partial cache lines.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    if (n % 8 != 0) return b;
    return (b % 8) * (n / 8) + b / 8;
}

// The same for any n: XCD x takes the n/8 (+1 for x < n % 8) consecutive
// indices from x (n/8) + min(x, n % 8).
__device__ __forceinline__ int xcd_remap_any(int b, int n) {
    const int x = b % 8, q = n / 8, r = n % 8;
    return x * q + (x < r ? x : r) + b / 8;
}

__device__ __forceinline__ double ky_of(int j, int Ny, double dky) {
    return (j < Ny / 2) ? j * dky : (j - Ny) * dky;
}

// Row-spectral layouts.  A buffer of H columns by Ny rows is stored in
// y-blocks of YB = 2^s rows, and within a y-block in groups of XB = 2^x
// neighbouring columns: element (kx, y) at
//   (y >> s) Q + (kx >> x) P + (y mod YB) XB + (kx mod XB),
// with group pitch P and block pitch Q.  A layout code is s + 32 x.
// s = log2(Ny), x = 0 is the column-major layout kx P + y (whole columns
// contiguous); s = log2(TR) makes the TR rows of a row-pass block one
// contiguous run of TR elements per kx; x = 1 interleaves the columns in
// pairs, so that the row pass's TR rows of two neighbouring kx are one run
// of 2 TR elements while a column pair stays contiguous.
// Long strides are padded by PFC_CPAD bytes: P = YB XB (+ pad when YB >= 1024),
// Q = H YB (+ pad when YB < 1024).  Unpadded, the lanes of the pass that cuts
// across the layout (consecutive kx in the row pass on the column-major
// layout) sit a power of two apart (64 KiB for 4096 FP64 rows) and fall into
// the same L2 sets and HBM channels, so partial lines are evicted before
// their other half is read; an odd multiple of 128 bytes spreads them.
#ifndef PFC_CPAD
#define PFC_CPAD 384
#endif
// ceil(log2 n), as the host-side ilog2
__host__ __device__ constexpr int ilog2c(int n) {
    int s = 0;
    while ((1 << s) < n) s++;
    return s;
}
__host__ __device__ constexpr int lay_s(int code) { return code & 31; }
__host__ __device__ constexpr int lay_x(int code) { return code >> 5; }
__host__ __device__ constexpr unsigned lay_pitch(int code, unsigned E) {
    return (1u << (lay_s(code) + lay_x(code))) + (lay_s(code) >= 10 ? (unsigned)PFC_CPAD / E : 0u);
}
__host__ __device__ constexpr unsigned lay_bpitch(int code, int H, unsigned E) {
    return (unsigned)H * (1u << lay_s(code)) + (lay_s(code) < 10 ? (unsigned)PFC_CPAD / E : 0u);
}
// Row-pass lane (kx = t + q TPB, row y): byte offset of q = 0 and the step in
// q (TPB a multiple of XB).
__device__ __forceinline__ unsigned lay_row_off(int code, int H, int t, int y, unsigned E) {
    const int s = lay_s(code), x = lay_x(code);
    return ((unsigned)(y >> s) * lay_bpitch(code, H, E) + (unsigned)(t >> x) * lay_pitch(code, E) +
            ((unsigned)(y & ((1 << s) - 1)) << x) + (unsigned)(t & ((1 << x) - 1))) * E;
}
__device__ __forceinline__ unsigned lay_row_step(int code, int TPB, unsigned E) {
    return ((unsigned)(TPB >> lay_x(code)) * lay_pitch(code, E)) * E;
}
// Column-pass lane (column kx, y = t + q TPB); needs TPB >= YB or YB >= Ny.
__device__ __forceinline__ unsigned lay_col_off(int code, int H, int kx, int t, unsigned E) {
    const int s = lay_s(code), x = lay_x(code);
    return ((unsigned)(t >> s) * lay_bpitch(code, H, E) + (unsigned)(kx >> x) * lay_pitch(code, E) +
            ((unsigned)(t & ((1 << s) - 1)) << x) + (unsigned)(kx & ((1 << x) - 1))) * E;
}
__device__ __forceinline__ unsigned lay_col_step(int code, int H, int TPB, unsigned E) {
    const int s = lay_s(code), x = lay_x(code);
    return (TPB >= (1 << s)) ? (unsigned)(TPB >> s) * lay_bpitch(code, H, E) * E
                             : ((unsigned)TPB << x) * E;
}
// Elements per field of a buffer in layout code (one past the last element).
inline size_t lay_elems(int code, int H, int Ny, unsigned E) {
    const int s = lay_s(code), x = lay_x(code);
    return (size_t)(((Ny + (1 << s) - 1) >> s) - 1) * lay_bpitch(code, H, E) +
           (size_t)(H >> x) * lay_pitch(code, E);
}

// ------------------------------------------------ elementary functions
//
// exp and expm1 for the operators A and B, written for this kernel: Cody-Waite
// argument reduction x = n ln2 + r with |r| <= ln2/2 (ln2 split so that n*hi
// is exact), the Taylor series of e^r (resp. e^r - 1), and a scale by 2^n.
// No special-case branches: the argument is clamped to where the result
// underflows to 0 or overflows to inf (NaN passes through the clamps).

// FP64 exp: degree-13 Taylor series (truncation < 5e-18 relative on |r| <=
// 0.3466), within about 1 ulp of the correctly rounded value.
__device__ __forceinline__ double exp_cw(double x) {
    x = x < -746.0 ? -746.0 : x;
    x = x > 710.0 ? 710.0 : x;
    const double n = rint(x * 1.4426950408889634074);
    double r = fma(-n, 6.93147180369123816490e-01, x);   // ln2 high part (21 trailing zero bits)
    r = fma(-n, 1.90821492927058770002e-10, r);           // ln2 low part
    double q = 1.0 / 6227020800.0;                        // 1/13!
    q = fma(q, r, 1.0 / 479001600.0);
    q = fma(q, r, 1.0 / 39916800.0);
    q = fma(q, r, 1.0 / 3628800.0);
    q = fma(q, r, 1.0 / 362880.0);
    q = fma(q, r, 1.0 / 40320.0);
    q = fma(q, r, 1.0 / 5040.0);
    q = fma(q, r, 1.0 / 720.0);
    q = fma(q, r, 1.0 / 120.0);
    q = fma(q, r, 1.0 / 24.0);
    q = fma(q, r, 1.0 / 6.0);
    q = fma(q, r, 0.5);
    q = fma(q, r, 1.0);
    q = fma(q, r, 1.0);
    return ldexp(q, (int)n);
}

// FP32 expm1: degree-8 series of e^r - 1 (truncation < 2e-8 relative), then
// e^x - 1 = 2^n (e^r - 1) + (2^n - 1), which has no cancellation for n != 0.
__device__ __forceinline__ float expm1_cw(float x) {
    x = x < -104.0f ? -104.0f : x;
    x = x > 89.0f ? 89.0f : x;
    const float n = rintf(x * 1.44269504f);
    float r = fmaf(-n, 0.693359375f, x);   // ln2 = 0.693359375 - 2.12194440e-4
    r = fmaf(n, 2.12194440e-4f, r);
    float q = 1.0f / 40320.0f;
    q = fmaf(q, r, 1.0f / 5040.0f);
    q = fmaf(q, r, 1.0f / 720.0f);
    q = fmaf(q, r, 1.0f / 120.0f);
    q = fmaf(q, r, 1.0f / 24.0f);
    q = fmaf(q, r, 1.0f / 6.0f);
    q = fmaf(q, r, 0.5f);
    q = fmaf(q * r, r, r);                 // e^r - 1
    const float s = ldexpf(1.0f, (int)n);
    return fmaf(s, q, s - 1.0f);
}

// 1/a: hardware reciprocal estimate refined by two Newton steps (FP64) or
// the hardware reciprocal as is (FP32, 1 ulp).
__device__ __forceinline__ double rcp_nr(double a) {
    double r = __builtin_amdgcn_rcp(a);
    double e = fma(-a, r, 1.0);
    r = fma(r, e, r);
    e = fma(-a, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float rcp_nr(float a) { return __builtin_amdgcn_rcpf(a); }

// ---------------------------------------------------------------- k space

// A and B of update_AB (src/pfc-relax.c:940-953) for field u at k2:
// A = exp(o a dt), B = (A - 1)/a/(Nx Ny), or o dt/(Nx Ny) where a == 0.
__device__ __forceinline__ void op_AB(const DevParams* __restrict__ p, int u, double k2,
                                      double& A, double& B) {
    double P = 1.0;
    for (int w = 0; w < p->W[u]; w++) {
        const double pw = p->nu2[u][w] - k2;
        P *= p->omega[u][w] * pw * pw + p->phi[u][w];
    }
    const double a = p->alpha[u][u] + p->beta[u][u] * P;
    const double o = p->c1 - 1.0 - p->c1 * k2;
    const double ex = exp_cw(o * a * p->dt);
    A = ex;
    const double b = (a == 0.0) ? o * p->dt : (ex - 1.0) * rcp_nr(a);
    B = b * p->inv_NxNy;
}

// FP32 variant: same operator, with e^x - 1 from expm1 so that the
// long-wavelength modes, where exp(x) - 1 cancels to a few bits in FP32, keep
// full precision.
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
    const float em1 = expm1_cw(o * a * (float)p->dt);
    A = 1.0f + em1;
    const float b = (a == 0.0f) ? o * (float)p->dt : em1 * rcp_nr(a);
    B = b * (float)p->inv_NxNy;
}

// The k-space update of step() (src/pfc-relax.c:1214-1242) at one k point:
// explicit couplings into M, then N <- A N + B M.  On return N holds the new
// state (the reference also copies it into M).  G2 = exp(-k2/(2 sigma^2))^2,
// passed in (it is separable in kx and ky and comes from tables).
template <int U, typename T>
__device__ __forceinline__ void kupdate(const DevParams* __restrict__ p, T k2, T G2,
                                        cpx<T> (&M)[U], cpx<T> (&N)[U]) {
    const T NxNy = (T)p->NxNy;
    if constexpr (U > 1) {
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
    } else {
        (void)G2;
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

// The pre/post twiddle e^{-2 pi i t/Nx} of lane t (see c2r_pre).  row_wt
// loads it through an opaque zero offset, so that a kernel calling the
// helpers several times reloads it rather than keeping it (and the products
// with the roots) live throughout.
template <typename T>
__device__ __forceinline__ cpx<T> row_wt(GBuf twx, int t) {
    unsigned z = 0;
    asm volatile("" : "+s"(z));
    return gload<T>(twx, (unsigned)t * sizeof(cpx<T>), z);
}

// Index H - k of the Hermitian partner of k = t + q TPB (0 < k < H), and 0
// for k = 0 (q = t = 0), where the caller selects the packed DC/Nyquist value
// instead.  For q >= 1, k >= TPB > 0.
template <int H, int TPB>
__device__ __forceinline__ int row_partner(int t, int q) {
    return q == 0 ? (t == 0 ? 0 : H - t) : H - t - q * TPB;
}

// wt is either the twiddle or a pointer to it in LDS, read after the barrier
// (so that a table written before the call is complete).
template <typename T>
__device__ __forceinline__ cpx<T> wt_value(cpx<T> w) { return w; }
template <typename T>
__device__ __forceinline__ cpx<T> wt_value(const cpx<T>* w) {
    int z = 0;   // opaque: re-read per call rather than kept live between transforms
    asm volatile("" : "+s"(z));
    return w[z];
}

// C2R pre-processing of one row held in natural order (thread t holds
// X[t + q*TPB], X[0] packed as (DC, Nyquist)) in exchange buffer L:
// produces Z with IFFT_H(Z)[n] = x[2n] + i x[2n+1], x the unnormalised c2r
// of X.  wt = e^{-2 pi i t/Nx}: e^{-2 pi i k/Nx} for k = t + q*TPB is
// wt e^{-2 pi i q/(2P)}.
template <typename T, int H, int TPB, int LS, typename WT>
__device__ __forceinline__ void c2r_pre(cpx<T> (&v)[H / TPB], cpx<T>* L, WT wtv, int t) {
    constexpr int P = H / TPB;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < P; q++) L[(t + q * TPB) * LS] = v[q];
    __syncthreads();
    const cpx<T> wt = wt_value<T>(wtv);
#pragma unroll
    for (int q = 0; q < P; q++) {
        // k = t + q TPB is 0 only for q = t = 0 (packed DC and Nyquist): the
        // general value is computed for every lane (partner index 0 there)
        // and replaced by a select, so there is no branch
        const cpx<T> a = v[q];
        const cpx<T> b = cconj(L[row_partner<H, TPB>(t, q) * LS]);
        const cpx<T> xe = cadd(a, b);
        const cpx<T> w = cmul(wt, root<T>(q, 2 * P, -1));   // e^{-2 pi i k/Nx}
        const cpx<T> xo = cmulc(csub(a, b), w);   // * e^{+2 pi i k/Nx}
        v[q] = {xe.x - xo.y, xe.y + xo.x};              // xe + i xo
        if (q == 0 && t == 0) v[q] = {a.x + a.y, a.x - a.y};
    }
}
// R2C post-processing: v holds Z = FFT_H(z), z[n] = x[2n] + i x[2n+1], in
// natural order; produces X[k] (k < H) with X[0] packed as (DC, Nyquist).
template <typename T, int H, int TPB, int LS, typename WT>
__device__ __forceinline__ void r2c_post(cpx<T> (&v)[H / TPB], cpx<T>* L, WT wtv, int t) {
    constexpr int P = H / TPB;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < P; q++) L[(t + q * TPB) * LS] = v[q];
    __syncthreads();
    const cpx<T> wt = wt_value<T>(wtv);
#pragma unroll
    for (int q = 0; q < P; q++) {
        // k = 0 only for q = t = 0; selected as in c2r_pre
        const cpx<T> a = v[q];
        const cpx<T> b = cconj(L[row_partner<H, TPB>(t, q) * LS]);
        const cpx<T> s = cadd(a, b);   // 2 Xe
        const cpx<T> d = csub(a, b);   // 2i Xo
        const cpx<T> xo2 = {d.y, -d.x};   // 2 Xo
        const cpx<T> w = cmul(wt, root<T>(q, 2 * P, -1));   // e^{-2 pi i k/Nx}
        const cpx<T> r = cadd(s, cmul(xo2, w));
        v[q] = cscale(r, (T)0.5);
        if (q == 0 && t == 0) v[q] = {a.x + a.y, a.x - a.y};   // packed (DC, Nyquist)
    }
}
// Copies NBYTES (a multiple of 16) from a global buffer at byte offset src
// into LDS with asynchronous buffer-to-LDS loads, 16 bytes per lane, no
// registers involved.  Completion is counted like a global load, so issue
// the copy before the field loads; it is complete once those are waited
// for, and a barrier then makes it visible to the block.  Reading tables
// from LDS afterwards never waits behind global loads and stores, which
// complete in issue order under one counter.
template <int NBYTES, int NT>
__device__ __forceinline__ void lds_copy16(GBuf g, unsigned src, void* dst, int tid) {
    static_assert(NBYTES % 16 == 0, "16-byte granules");
    constexpr int WAVES = NT / 64;
    const int wave = tid / 64, lane = tid % 64;
#pragma unroll
    for (int base = 0; base < NBYTES; base += WAVES * 1024) {
        const int off = base + wave * 1024;   // this wave's 64 x 16-byte chunk
        if (off < NBYTES && off + lane * 16 < NBYTES) {
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                g.r, (__attribute__((address_space(3))) void*)((char*)dst + off), 16,
                (unsigned)(src + off + lane * 16), 0, 0, 0);
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

// The spectral energy terms of one half-spectrum point (Parseval; see
// k_energy_spec): 1/2 beta_uu P_u |N_u|^2 + (beta_uv (c2 - k2)^2 + eps_uv G^2)
// Re(N_u conj N_v), counted twice for the columns 0 < kx < Nx/2 that the
// C2R along x takes as 2 Re(.).
template <int U, typename T>
__device__ __forceinline__ double spec_energy(const DevParams* __restrict__ p, int kx, int Nx,
                                              double k2, const cpx<T> (&N)[U]) {
    double re[U], im[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        re[u] = (double)N[u].x;
        im[u] = (double)N[u].y;
    }
    const double G = exp_cw(p->sig_a * k2);   // eval_mult's G(k), this file's exp
    const double G2 = G * G;
    double e = 0.0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        e += 0.5 * p->beta[u][u] * eval_mult(p, 0, u, 0.0, k2) * (re[u] * re[u] + im[u] * im[u]);
#pragma unroll
        for (int v = u + 1; v < U; v++) {
            const double m = p->beta[u][v] * eval_mult(p, 2, u, p->qc2[u][v], k2) + p->epsilon[u][v] * G2;
            e += m * (re[u] * re[v] + im[u] * im[v]);
        }
    }
    const bool once = (kx == 0) || (2 * kx == Nx);
    return once ? e : 2.0 * e;
}

// The local terms of the free energy at one point (src/pfc-relax.c:1093-1117
// without the filtered fields): lambda n + alpha n^2/2 + gamma n^3/3 + delta
// n^4/4 per field and the pair products, accumulated into acc[0], each n_u
// into acc[1 + u].
template <int U>
__device__ __forceinline__ void local_energy(const DevParams* __restrict__ p, const double (&nv)[U],
                                             KSum* acc) {
    const double _3 = 1.0 / 3.0;
    double fij = 0.0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        const double nu = nv[u];
        const double nunu = nu * nu;
        fij += p->lambda[u] * nu + 0.5 * p->alpha[u][u] * nunu + _3 * p->gamma[u][u] * nu * nunu +
               0.25 * p->delta[u] * nunu * nunu;
#pragma unroll
        for (int v = u + 1; v < U; v++) {
            const double nvv = nv[v];
            fij += p->alpha[u][v] * nu * nvv + 0.5 * p->gamma[u][v] * (nunu * nvv + nu * nvv * nvv);
        }
        kadd(acc[1 + u], nu);
    }
    kadd(acc[0], fij);
}

// ================================================================ kernels

// LDS index padding of the step kernels' FP32 transforms (BlockFFT PADC;
// FP64 transforms are swizzled instead), and the exchange buffers' sizes.
template <typename T>
constexpr int step_padc() { return sizeof(T) == 4 ? PFC_STEP_PADC : PFC_STEP_PADC64; }
template <typename T, int NX, int TPB, int TR>
constexpr int row_lds_elems() {
    return BlockFFT<T, NX / 2, TPB, -1, 2, TR, step_padc<T>()>::lds_len * TR;
}
template <typename T, int NY, int TPB>
constexpr int col_lds_len() {
    return BlockFFT<T, NY, TPB, -1, 1, 1, step_padc<T>()>::lds_len;
}

// Entries of the row pass's LDS twiddle table for L = 2^k row transforms:
// the transform's compact stage table, then e^{-2 pi i t/Nx} for t < TPB
// (the real-FFT pre/post twiddle), padded to a multiple of the block's
// TPB*TR lanes so that every lane gathers and writes without a branch;
// 1 (unused) for 3 * 2^k.
template <typename T, int NX, int TPB>
constexpr int row_tab_used() {
    using F = BlockFFT<T, NX / 2, TPB, -1, 2, 1>;
    return F::POW2 ? F::kTab + TPB : 0;
}
template <typename T, int NX, int TPB, int TR>
constexpr int row_tab_size() {
    constexpr int n = row_tab_used<T, NX, TPB>(), nt = TPB * TR;
    return n ? (n + nt - 1) / nt * nt : 1;
}

// The row pass's table gather (L = 2^k): lane tid loads entries tid + i*NT
// of the table from the Nx-point table (index clamped, no branch).
template <typename T, int NX, int TPB, int TR, int NI>
__device__ __forceinline__ void row_tab_gather(cpx<T> (&tv)[NI], GBuf twx, int tid) {
    using F = BlockFFT<T, NX / 2, TPB, -1, 2, TR>;
    constexpr int NT = TPB * TR, NTAB = row_tab_used<T, NX, TPB>();
    constexpr unsigned E = sizeof(cpx<T>);
    static_assert(NI * NT == row_tab_size<T, NX, TPB, TR>(), "one entry per lane and round");
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int e = tid + i * NT;
        const int m = e < F::kTab ? 2 * F::tab_src(e) : (e < NTAB ? e - F::kTab : 0);
        tv[i] = gload<T>(twx, (unsigned)m * E, 0u);
    }
}

// Row pass of a step.  Block = TR rows, lane = (fft thread t, row r) with the
// row index fastest, so the TR rows' elements of one kx are adjacent lanes
// and adjacent addresses in the column-major buffers (TR*sizeof(E) bytes
// contiguous per kx).
//
// Core of the row pass once the U fields of row y are in registers (m[u][q]
// = I_u(kx = t + q TPB, y)): C2R, NaN check, refresh R2C into K, nonlinear
// term, R2C into J.  L = lds + r (H*TR elements); ltab: the twiddle table
// (L = 2^k; complete before the first barrier of c2r_pre).
//
// For L = 2^k the transforms take their twiddles from ltab: global loads and
// stores complete in issue order under one counter, so twiddle loads issued
// during the transforms would make the C2R of field 0 wait for every field's
// loads and the R2C of field u + 1 wait until field u's stores have drained.
// With EVAL (a refresh step that also evaluates the free energy of the
// incoming state, pfc_step_evaluate), the local energy terms of the
// densities are added to eacc[0..U] as k_row_c2r_eloc adds them.
template <typename T, int NX, int TPB, int TR, int U, bool REFRESH, bool EVAL = false>
__device__ __forceinline__ void row_step_core(
    cpx<T> (&m)[U][NX / 2 / TPB], cpx<T>* L, const cpx<T>* ltab, GBuf twx, int t, int y,
    cpx<T>* __restrict__ J, cpx<T>* __restrict__ K, size_t fstride, int sJ,
    const DevParams* __restrict__ prm, int* __restrict__ nan_step, int tstep,
    KSum* eacc = nullptr) {
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    constexpr unsigned E = sizeof(cpx<T>);
    using FFTF = BlockFFT<T, H, TPB, -1, 2, TR, step_padc<T>()>;
    using FFTI = BlockFFT<T, H, TPB, +1, 2, TR, step_padc<T>()>;
    constexpr bool STAB = FFTF::POW2;
    const unsigned voff = lay_row_off(sJ, H, t, y, E), qstep = lay_row_step(sJ, TPB, E);
    // the transforms' twiddles and the pre/post twiddle of lane t (ltab is
    // read only after the first barrier inside c2r_pre)
    const StageTab<T> stab{ltab};
    const cpx<T>* wtp = ltab + FFTF::kTab + t;
    auto fwd = [&](cpx<T>(&v)[P]) {
        if constexpr (STAB) {
            FFTF::run(v, L, stab, t);
            r2c_post<T, H, TPB, TR>(v, L, wtp, t);
        } else {
            FFTF::run(v, L, twx, t);
            r2c_post<T, H, TPB, TR>(v, L, row_wt<T>(twx, t), t);
        }
    };
#pragma unroll
    for (int u = 0; u < U; u++) {
        if constexpr (STAB) {
            c2r_pre<T, H, TPB, TR>(m[u], L, wtp, t);
            FFTI::run(m[u], L, stab, t);
        } else {
            c2r_pre<T, H, TPB, TR>(m[u], L, row_wt<T>(twx, t), t);
            FFTI::run(m[u], L, twx, t);
        }
        if (u == 0) PFC_PROBE_AT(0, 1);
    }
    PFC_PROBE_AT(0, 2);
    // NaN check on field 0 at (0,0) (src/pfc-relax.c:1175)
    if (y == 0 && t == 0) {
        const T v0 = m[0][0].x;
        if (v0 != v0) atomicMin(nan_step, tstep);
    }
    if constexpr (EVAL) {
#pragma unroll
        for (int q = 0; q < P; q++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                double nv[U];
#pragma unroll
                for (int u = 0; u < U; u++) nv[u] = (double)(h ? m[u][q].y : m[u][q].x);
                local_energy<U>(prm, nv, eacc);
            }
        }
    }
    if constexpr (REFRESH) {
        // R2C of the density itself, for N = FFT(n)/NxNy (src/pfc-relax.c:1178-1184)
#pragma unroll
        for (int u = 0; u < U; u++) {
            cpx<T> w[P];
#pragma unroll
            for (int q = 0; q < P; q++) w[q] = m[u][q];
            fwd(w);
            const GBuf dst = gbuf(K + u * fstride);
#pragma unroll
            for (int q = 0; q < P; q++) gstore<T>(dst, voff, q * qstep, w[q]);
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
    PFC_PROBE_AT(0, 3);
#pragma unroll
    for (int u = 0; u < U; u++) {
        fwd(m[u]);
        const GBuf dst = gbuf(J + u * fstride);
#pragma unroll
        for (int q = 0; q < P; q++) gstore<T, PFC_ROWJ_AUX>(dst, voff, q * qstep, m[u][q]);
    }
    PFC_PROBE_AT(0, 4);
}

// Body of the row pass for row group g (rows g*TR ... g*TR + TR - 1) on a
// block of TPB*TR lanes; lds holds H*TR elements, ltab row_tab_size()
// (16-byte aligned).  The table gather is issued before the field loads and
// written to LDS after them, so that the write waits for the gather only.
// Shared by the one-system kernel below and the batch kernels.
template <typename T, int NX, int TPB, int TR, int U, bool REFRESH, bool EVAL = false>
__device__ __forceinline__ void row_step_body(
    cpx<T>* lds, cpx<T>* ltab, int g, const cpx<T>* __restrict__ I, cpx<T>* __restrict__ J,
    cpx<T>* __restrict__ K, size_t fstride, int sI, int sJ, const cpx<T>* __restrict__ twx_,
    const DevParams* __restrict__ prm, int* __restrict__ nan_step, int tstep,
    KSum* eacc = nullptr) {
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    constexpr int NT = TPB * TR;
    constexpr unsigned E = sizeof(cpx<T>);
    constexpr bool STAB = BlockFFT<T, H, TPB, -1, 2, TR>::POW2;
    constexpr int NI = row_tab_size<T, NX, TPB, TR>() / NT;
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    const int y = g * TR + r;
    const GBuf twx = gbuf(twx_);
    const unsigned ioff = lay_row_off(sI, H, t, y, E), istep = lay_row_step(sI, TPB, E);

    PFC_PROBE_AT(0, 0);
    cpx<T> tv[NI > 0 ? NI : 1];
    if constexpr (STAB) {
        row_tab_gather<T, NX, TPB, TR>(tv, twx, tid);
        __builtin_amdgcn_sched_barrier(0);   // issued before the field loads
    }
    cpx<T> m[U][P];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const GBuf src = gbuf(I + u * fstride);
#pragma unroll
        for (int q = 0; q < P; q++) m[u][q] = gload<T>(src, ioff, q * istep);
    }
    if constexpr (STAB) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < NI; i++) ltab[tid + i * NT] = tv[i];
    }
    row_step_core<T, NX, TPB, TR, U, REFRESH, EVAL>(m, lds + r, ltab, twx, t, y, J, K, fstride, sJ,
                                                    prm, nan_step, tstep, eacc);
}

// Block totals of the EVAL variants' compensated sums: nv sums per block at
// part[block * 2 nv].
template <int NV, int BS>
__device__ __forceinline__ void eval_partials(KSum (&acc)[NV], double* __restrict__ part) {
    block_ksum<NV, BS>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) {
            part[(size_t)blockIdx.x * 2 * NV + 2 * i] = acc[i].s;
            part[(size_t)blockIdx.x * 2 * NV + 2 * i + 1] = acc[i].c;
        }
    }
}

// The EVAL variants (one step per evaluate) run at most two waves per SIMD,
// so that the energy accumulators do not push the FP32 kernels into spills.
template <typename T, int NX, int TPB, int TR, int U, bool REFRESH, bool EVAL = false>
__global__ __launch_bounds__(TPB* TR) PFC_WAVES((EVAL ? 2 : row_waves<T, NX>())) void k_row_step(
    const cpx<T>* __restrict__ I, cpx<T>* __restrict__ J, cpx<T>* __restrict__ K,
    size_t fstride, int sI, int sJ, const cpx<T>* __restrict__ twx_,
    const DevParams* __restrict__ prm, int* __restrict__ nan_step, int tstep,
    double* __restrict__ epart) {
    __shared__ cpx<T> lds[row_lds_elems<T, NX, TPB, TR>()];
    __shared__ __attribute__((aligned(16))) cpx<T> ltab[row_tab_size<T, NX, TPB, TR>()];
#ifndef PFC_ROW_XCD
#define PFC_ROW_XCD 1
#endif
    const int g = PFC_ROW_XCD ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    KSum acc[EVAL ? 1 + U : 1];
    if constexpr (EVAL) {
#pragma unroll
        for (int i = 0; i <= U; i++) acc[i] = {0.0, 0.0};
    }
    row_step_body<T, NX, TPB, TR, U, REFRESH, EVAL>(lds, ltab, g, I, J, K, fstride, sI, sJ, twx_,
                                                    prm, nan_step, tstep, acc);
    if constexpr (EVAL) eval_partials<1 + U, TPB * TR>(acc, epart);
}

// ------------------------------------------------------- persistent passes
//
// The FP64 row pass runs as a persistent launch: one grid of resident blocks
// (the CUs times the blocks one CU holds) walks the row groups instead of one
// block per group.  A block keeps its twiddle table in LDS for all its groups
// and its next group's loads start while the CU's waves finish the previous
// one.  Measured at 4096^2 with three fields: FP64 row pass 391-401 -> 363-365
// us.  The FP32 row pass (two blocks per CU) went 176 -> 190 us and both
// column passes lost 14-40 % in the same form, so they stay one block per
// work item.

// Row group of block b (of nb) at round it, for n row groups, or n when the
// block is done.  XCD x owns the contiguous range [x n/8, (x+1) n/8) and its
// blocks j = b/8 take consecutive groups of it in every round, so the CUs of
// one XCD work on neighbouring rows at any time and their partial lines of
// the column-major buffers merge in that XCD's L2.
__device__ __forceinline__ int persist_row_group(int b, int nb, int it, int n) {
    if (n % 8 != 0 || nb % 8 != 0) {
        const int g = b + it * nb;
        return g < n ? g : n;
    }
    const int lg = it * (nb / 8) + b / 8;
    return lg < n / 8 ? (b % 8) * (n / 8) + lg : n;
}

// The row pass's LDS twiddle table (L = 2^k), gathered and written once.
template <typename T, int NX, int TPB, int TR>
__device__ __forceinline__ void row_tab_load(cpx<T>* ltab, const cpx<T>* __restrict__ twx_) {
    constexpr int NT = TPB * TR;
    constexpr int NI = row_tab_size<T, NX, TPB, TR>() / NT;
    if constexpr (BlockFFT<T, NX / 2, TPB, -1, 2, TR>::POW2) {
        cpx<T> tv[NI];
        row_tab_gather<T, NX, TPB, TR>(tv, gbuf(twx_), threadIdx.x);
#pragma unroll
        for (int i = 0; i < NI; i++) ltab[threadIdx.x + i * NT] = tv[i];
    }
}

// Persistent row pass over ngroups row groups of TR rows; the table written
// before the loop is complete at the first barrier of c2r_pre, which also
// separates each group's LDS use from the previous group's.
template <typename T, int NX, int TPB, int TR, int U, bool REFRESH, bool EVAL = false>
__global__ __launch_bounds__(TPB* TR) PFC_WAVES((EVAL ? 2 : row_waves<T, NX>())) void k_row_step_p(
    const cpx<T>* __restrict__ I, cpx<T>* __restrict__ J, cpx<T>* __restrict__ K,
    size_t fstride, int sI, int sJ, const cpx<T>* __restrict__ twx_,
    const DevParams* __restrict__ prm, int* __restrict__ nan_step, int tstep, int ngroups,
    double* __restrict__ epart) {
    __shared__ cpx<T> lds[row_lds_elems<T, NX, TPB, TR>()];
    __shared__ __attribute__((aligned(16))) cpx<T> ltab[row_tab_size<T, NX, TPB, TR>()];
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    constexpr unsigned E = sizeof(cpx<T>);
    row_tab_load<T, NX, TPB, TR>(ltab, twx_);
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    const GBuf twx = gbuf(twx_);
    const unsigned istep = lay_row_step(sI, TPB, E);
    KSum acc[EVAL ? 1 + U : 1];
    if constexpr (EVAL) {
#pragma unroll
        for (int i = 0; i <= U; i++) acc[i] = {0.0, 0.0};
    }
    for (int it = 0;; it++) {
        const int g = persist_row_group(blockIdx.x, gridDim.x, it, ngroups);
        if (g >= ngroups) break;
        const int y = g * TR + r;
        const unsigned ioff = lay_row_off(sI, H, t, y, E);
        cpx<T> m[U][P];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const GBuf src = gbuf(I + u * fstride);
#pragma unroll
            for (int q = 0; q < P; q++) m[u][q] = gload<T>(src, ioff, q * istep);
        }
        row_step_core<T, NX, TPB, TR, U, REFRESH, EVAL>(m, lds + r, ltab, twx, t, y, J, K, fstride,
                                                        sJ, prm, nan_step, tstep, acc);
    }
    if constexpr (EVAL) eval_partials<1 + U, TPB * TR>(acc, epart);
}

// Column addressing modes of the column passes.  mode 0: an ordinary column
// kx of the row-spectral buffers; mode 1: the kx = 0 column, whose real input
// is the .x half of packed column 0; mode 2: the kx = H column from the .y
// half.  Modes 1 and 2 write back only the real parts of their inverse
// transforms, into the same halves.
template <typename T>
__device__ __forceinline__ cpx<T> col_in(GBuf b, unsigned voff, unsigned soff, int mode) {
    const cpx<T> v = gload<T>(b, voff, soff);
    if (mode == 0) return v;
    return {mode == 1 ? v.x : v.y, (T)0};
}
template <typename T, int AUX = 0>
__device__ __forceinline__ void col_out(GBuf b, unsigned voff, unsigned soff, int mode, cpx<T> v) {
    if (mode == 0) gstore<T, AUX>(b, voff, soff, v);
    else gstore_real<T, AUX>(b, voff + (mode == 2 ? (unsigned)sizeof(T) : 0u), soff, v.x);
}

// Which column lane group c of block b handles.  The grid has H/TC + 1
// blocks: group 0 of block 0 takes kx = 0 and group 0 of the extra last
// block kx = H (both from packed column 0, writing disjoint halves of it);
// the other groups of the last block only join its barriers.
struct ColJob {
    int kx, mode;
    bool live;
};
__device__ __forceinline__ ColJob col_job(int b, int c, int TC, int H) {
    if (b == H / TC) return {H, 2, c == 0};
    if (b == 0 && c == 0) return {0, 1, true};
    return {b * TC + c, 0, true};
}

// Hermitian projection of the kx = 0 and kx = H columns (modes 1, 2) of a
// block-0 pass, N(ky) <- (N(ky) + conj N(-ky))/2, through LDS.  The FP32 form
// applies it to the state after every update: FP32 round-off in the
// anti-Hermitian part would otherwise be amplified by the unstable modes
// between refreshes and leak into the real parts.
template <typename T, int NY, int TPB, int U>
__device__ __forceinline__ void hermitian_cols(cpx<T> (&M)[U][NY / TPB], cpx<T>* L, int t, int mode) {
    constexpr int P = NY / TPB;
#pragma unroll
    for (int u = 0; u < U; u++) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < P; q++) L[t + q * TPB] = M[u][q];
        __syncthreads();
        if (mode != 0) {
#pragma unroll
            for (int q = 0; q < P; q++) {
                const int jj = t + q * TPB;
                const cpx<T> pj = L[jj == 0 ? 0 : NY - jj];
                M[u][q] = cscale(cadd(M[u][q], cconj(pj)), (T)0.5);
            }
        }
    }
}

// Entries of the column pass's LDS twiddle table: for L = 2^k the compact
// stage table (BlockFFT::kTab, read by stage-local index, conflict-free),
// padded to a multiple of the block's TPB*TC lanes so that every lane
// gathers and writes without a branch; for 3 * 2^k the full Ny-point table.
template <typename T, int NY, int TPB, int TC, bool LTAB>
constexpr int col_tw_entries() {
    using F = BlockFFT<T, NY, TPB, -1, 1, 1>;
    if (!LTAB) return 1;
    if (!F::POW2) return NY;
    return (F::kTab + TPB * TC - 1) / (TPB * TC) * (TPB * TC);
}

// Column pass of a step: TC columns per block (lane = column-major within the
// block), each column all Ny rows; the extra last block runs kx = H.
// fstride: row-spectral field stride (H*Ny), nstride: state stride ((H+1)*Ny).
// ktab: the Gaussian factors (ktab_gy); G2 is the product of the kx and the
// ky factor.
#ifndef PFC_COL_QB_F64
#define PFC_COL_QB_F64 4
#endif
#ifndef PFC_COL_QB_F32
#define PFC_COL_QB_F32 2
#endif
// Body of the column pass for column block cb (H/TC + 1 per system) on a
// block of TPB*TC lanes.  lds holds col_lds_len * TC elements; with LTAB, ltw
// (col_tw_entries, 16-byte aligned) holds the twiddle table, and with LGY,
// lgy (NY reals) the ky factors (read from global memory otherwise).  Shared
// by the one-system kernel below and the batch kernels.
// With EVAL (refresh steps only) the spectral energy terms of the refreshed
// state are added to eacc[0] before the update, as k_energy_spec adds them.
template <typename T, int NY, int TPB, int TC, int U, bool REFRESH, bool LTAB, bool LGY,
          bool EVAL = false>
__device__ __forceinline__ void col_step_body(
    cpx<T>* lds, cpx<T>* ltw, T* lgy, int cb, const cpx<T>* __restrict__ J,
    const cpx<T>* __restrict__ K, cpx<T>* __restrict__ Nh, cpx<T>* __restrict__ I, size_t fstride,
    size_t nstride, int H, int sI, int sJ, const cpx<T>* __restrict__ twy_,
    const T* __restrict__ ktab_, const DevParams* __restrict__ prm, double* __restrict__ n0,
    KSum* eacc = nullptr) {
    static_assert(!EVAL || REFRESH, "the evaluate reads the refreshed state");
    constexpr int P = NY / TPB;
    constexpr unsigned E = sizeof(cpx<T>);
    using FFTF = BlockFFT<T, NY, TPB, -1, 1, 1, step_padc<T>()>;
    using FFTI = BlockFFT<T, NY, TPB, +1, 1, 1, step_padc<T>()>;
    const int tid = threadIdx.x;
    const int c = tid / TPB;
    const int t = tid % TPB;
    cpx<T>* L = lds + c * FFTF::lds_len;
    const GBuf twy = gbuf(twy_);
    const GBuf ktab = gbuf(ktab_);
    PFC_PROBE_AT(65536, 0);
    {
        const ColJob jb = col_job(cb, c, TC, H);
        const int kx = jb.kx;
        // byte offsets of this lane's first element in a field's buffers
        const unsigned vr = lay_col_off(sJ, H, jb.mode ? 0 : kx, t, E);
        const unsigned rstep = lay_col_step(sJ, H, TPB, E);
        const unsigned vn = ((unsigned)kx * NY + (unsigned)t) * E;
        const unsigned vi = lay_col_off(sI, H, jb.mode ? 0 : kx, t, E);
        const unsigned istep = lay_col_step(sI, H, TPB, E);

        // tables first, then all fields' J.  For L = 2^k the twiddles are
        // the compact stage table, gathered into registers ahead of the field
        // loads and written to LDS after them (so that the write waits for
        // the gather only); for 3 * 2^k the full table is copied.
        constexpr bool CT = LTAB && FFTF::POW2;
        constexpr int NT = TPB * TC;
        constexpr int NI = CT ? col_tw_entries<T, NY, TPB, TC, LTAB>() / NT : 1;
        cpx<T> tv[NI];
        if constexpr (LTAB) {
            if constexpr (!CT) lds_copy16<NY * (int)E, NT>(twy, 0u, ltw, tid);
            if constexpr (LGY)
                lds_copy16<NY * (int)sizeof(T), NT>(ktab, (unsigned)ktab_gy(H) * sizeof(T), lgy, tid);
        }
        if constexpr (CT) {
#pragma unroll
            for (int i = 0; i < NI; i++) {
                const int e = tid + i * NT;
                tv[i] = gload<T>(twy, (unsigned)(e < FFTF::kTab ? FFTF::tab_src(e) : 0) * E, 0u);
            }
        }
        const T gx = gload_real<T>(ktab, (unsigned)kx * sizeof(T));
        __builtin_amdgcn_sched_barrier(0);
        cpx<T> M[U][P];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const GBuf src = gbuf(J + u * fstride);
#pragma unroll
            for (int q = 0; q < P; q++) M[u][q] = col_in<T>(src, vr, q * rstep, jb.mode);
        }
        if constexpr (CT) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < NI; i++) ltw[tid + i * NT] = tv[i];
        }
        __syncthreads();   // the table copies are complete once the loads before them are
        auto fwd = [&](cpx<T>(&v)[P]) {
            if constexpr (CT) FFTF::run(v, L, StageTab<T>{ltw}, t);
            else if constexpr (LTAB) FFTF::run(v, L, (const cpx<T>*)ltw, t);
            else FFTF::run(v, L, twy, t);
        };
        auto inv = [&](cpx<T>(&v)[P]) {
            if constexpr (CT) FFTI::run(v, L, StageTab<T>{ltw}, t);
            else if constexpr (LTAB) FFTI::run(v, L, (const cpx<T>*)ltw, t);
            else FFTI::run(v, L, twy, t);
        };
#pragma unroll
        for (int u = 0; u < U; u++) {
            fwd(M[u]);
            if (u == 0) PFC_PROBE_AT(65536, 1);
        }
        PFC_PROBE_AT(65536, 2);
        if constexpr (REFRESH) {
            // N = FFT(n)/NxNy from the density's row transforms
            // (src/pfc-relax.c:1178-1184), written to Nh and read back below
            // by the same lanes
            const T sc = (T)prm->inv_NxNy;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const GBuf src = gbuf(K + u * fstride);
                const GBuf dst = gbuf(Nh + u * nstride);
                cpx<T> w[P];
#pragma unroll
                for (int q = 0; q < P; q++) w[q] = col_in<T>(src, vr, q * rstep, jb.mode);
                fwd(w);
                if (jb.live) {
#pragma unroll
                    for (int q = 0; q < P; q++) gstore<T>(dst, vn, q * TPB * E, cscale(w[q], sc));
                }
            }
        }
        if (jb.live) {
            // k2 = kx^2 + ky^2 rounded as in the reference (src/pfc-relax.c:1217-1222)
            const double kxd = kx * prm->dkx;
            const double kx2 = kxd * kxd;
            // State loads in batches of QB k points, each batch's loads issued
            // together (a load issued after a store completes after it).
            constexpr int QB0 = sizeof(T) == 8 ? PFC_COL_QB_F64 : PFC_COL_QB_F32;
            // the largest of QB0, 2, 1 that divides P (P = 6 takes 2)
            constexpr int QB = (P % QB0 == 0) ? QB0 : (P % 2 == 0 ? 2 : 1);
            static_assert(P % QB == 0, "state batches must tile the lane's points");
#pragma unroll
            for (int q0 = 0; q0 < P; q0 += QB) {
                cpx<T> Nb[QB][U];
#pragma unroll
                for (int i = 0; i < QB; i++) {
#pragma unroll
                    for (int u = 0; u < U; u++)
                        Nb[i][u] = gload<T>(gbuf(Nh + u * nstride), vn, (q0 + i) * TPB * E);
                }
#pragma unroll
                for (int i = 0; i < QB; i++) {
                    const int q = q0 + i;
                    const int j = t + q * TPB;
                    const double ky = ky_of(j, NY, prm->dky);
                    const T k2 = (T)(kx2 + ky * ky);
                    if constexpr (EVAL) kadd(eacc[0], spec_energy<U, T>(prm, kx, 2 * H, kx2 + ky * ky, Nb[i]));
                    T gy;
                    if constexpr (LGY) gy = lgy[j];
                    else gy = gload_real<T>(ktab, (unsigned)(ktab_gy(H) + j) * sizeof(T));
                    const T G2 = gx * gy;
                    cpx<T> Mk[U];
#pragma unroll
                    for (int u = 0; u < U; u++) Mk[u] = M[u][q];
                    bool k0 = false;
                    if constexpr (sizeof(T) == 4) k0 = (jb.mode == 1 && j == 0);
                    if (k0) {
                        // k = 0 (the mean density) is carried in FP64: its
                        // update runs in FP64 from the FP64 state n0, and the
                        // FP32 state takes the rounded result.  Refreshes from
                        // the FP32 density leave n0 alone.
                        cpx<double> M0[U], N0[U];
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            M0[u] = {(double)Mk[u].x, (double)Mk[u].y};
                            N0[u] = {n0[2 * u], n0[2 * u + 1]};
                        }
                        kupdate<U, double>(prm, 0.0, 1.0, M0, N0);
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            n0[2 * u] = N0[u].x;
                            n0[2 * u + 1] = N0[u].y;
                            Nb[i][u] = {(T)N0[u].x, (T)N0[u].y};
                        }
                    } else {
                        kupdate<U, T>(prm, k2, G2, Mk, Nb[i]);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        M[u][q] = Nb[i][u];
                        gstore<T, PFC_COL_AUX>(gbuf(Nh + u * nstride), vn, q * TPB * E, M[u][q]);
                    }
                }
            }
        }
        PFC_PROBE_AT(65536, 3);
        if constexpr (sizeof(T) == 4) {
            // kx = 0 and kx = H: replace the stored state by its Hermitian part
            if (cb == 0 || cb == H / TC) {
                hermitian_cols<T, NY, TPB, U>(M, L, t, jb.live ? jb.mode : 0);
                if (jb.live && jb.mode != 0) {
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const GBuf nb = gbuf(Nh + u * nstride);
#pragma unroll
                        for (int q = 0; q < P; q++) gstore<T>(nb, vn, q * TPB * E, M[u][q]);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            inv(M[u]);
            if (jb.live) {
                const GBuf dst = gbuf(I + u * fstride);
#pragma unroll
                for (int q = 0; q < P; q++)
                    col_out<T, PFC_COL_AUX>(dst, vi, q * istep, jb.mode, M[u][q]);
            }
        }
    }
    PFC_PROBE_AT(65536, 4);
}

// twiddle table in LDS where it fits beside the exchange buffer (all sizes),
// and the ky factors too where they also fit (all but FP64 Ny = 8192, where
// the 128 KB exchange buffer and the 32 KB compact table fill the 160 KB)
template <typename T, int NY, int TPB, int TC>
constexpr bool col_ltab() {
    return col_lds_len<T, NY, TPB>() * TC * sizeof(cpx<T>) +
               col_tw_entries<T, NY, TPB, TC, true>() * sizeof(cpx<T>) <= 160 * 1024;
}
template <typename T, int NY, int TPB, int TC>
constexpr bool col_lgy() {
    return col_ltab<T, NY, TPB, TC>() &&
           col_lds_len<T, NY, TPB>() * TC * sizeof(cpx<T>) +
                   col_tw_entries<T, NY, TPB, TC, true>() * sizeof(cpx<T>) + NY * sizeof(T) <= 160 * 1024;
}

template <typename T, int NY, int TPB, int TC, int U, bool REFRESH, bool EVAL = false>
__global__ __launch_bounds__(TPB* TC) PFC_WAVES(col_waves<T>()) void k_col_step(
    const cpx<T>* __restrict__ J, const cpx<T>* __restrict__ K, cpx<T>* __restrict__ Nh,
    cpx<T>* __restrict__ I, size_t fstride, size_t nstride, int H, int sI, int sJ,
    const cpx<T>* __restrict__ twy_, const T* __restrict__ ktab_,
    const DevParams* __restrict__ prm, double* __restrict__ n0, double* __restrict__ epart) {
    constexpr bool LTAB = col_ltab<T, NY, TPB, TC>();
    constexpr bool LGY = col_lgy<T, NY, TPB, TC>();
    __shared__ cpx<T> lds[col_lds_len<T, NY, TPB>() * TC];
    __shared__ __attribute__((aligned(16))) cpx<T> ltw[col_tw_entries<T, NY, TPB, TC, LTAB>()];
    // with J in column groups, neighbouring column blocks on one XCD (they
    // read parts of the same lines)
    const int cb = lay_x(sJ) >= 1 ? xcd_remap_any(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    KSum acc[1] = {{0.0, 0.0}};
    if constexpr (LGY) {
        __shared__ __attribute__((aligned(16))) T lgy[NY];   // exp(-ky^2/sigma^2)
        col_step_body<T, NY, TPB, TC, U, REFRESH, LTAB, LGY, EVAL>(lds, ltw, lgy, cb, J, K, Nh, I,
                                                                   fstride, nstride, H, sI, sJ,
                                                                   twy_, ktab_, prm, n0, acc);
    } else {
        col_step_body<T, NY, TPB, TC, U, REFRESH, LTAB, LGY, EVAL>(lds, ltw, nullptr, cb, J, K, Nh,
                                                                   I, fstride, nstride, H, sI, sJ,
                                                                   twy_, ktab_, prm, n0, acc);
    }
    if constexpr (EVAL) eval_partials<1, TPB * TC>(acc, epart);
}

// ------------------------------------------------ per-field column pass
//
// The k-space update of field f alone needs the transform of f's own
// nonlinear term and the old states of all fields at the same k (the
// explicit couplings of src/pfc-relax.c:1214-1236 read N_v, not M_v).  So
// the column pass splits into one block per (column, field): each block
// transforms one field, reads every field's old state from Nsrc and writes
// its field's new state to Ndst (a second state buffer: the other fields'
// blocks still read the old one).  A block then holds one field's column in
// registers instead of U, which lets two FP64 blocks share a CU and overlap
// one block's memory traffic with the other's transforms.  The U blocks of a
// column run on one XCD at about the same time (xcd_remap_any over the
// items), so the repeated state reads are served by its L2.

// kupdate for field f alone: the same operations in the same order as
// kupdate's for M[f], N[f] (couplings by ascending v != f, then A, B).
template <int U, typename T>
__device__ __forceinline__ cpx<T> kupdate_one(const DevParams* __restrict__ p, int f, T k2, T G2,
                                             cpx<T> M, const cpx<T> (&N)[U]) {
    const T NxNy = (T)p->NxNy;
    cpx<T> Nf = N[0];
#pragma unroll
    for (int v = 0; v < U; v++) {
        if (v == f) {
            Nf = N[v];
            continue;
        }
        const int a = v < f ? v : f, b = v < f ? f : v;
        T d2 = (T)p->mu2[a][b] - k2;
        d2 *= d2;
        const T bd = (T)p->beta[a][b] * d2;
        const T eg = (T)p->epsilon[a][b] * G2;
        M.x += (bd * N[v].x + eg * N[v].x) * NxNy;
        M.y += (bd * N[v].y + eg * N[v].y) * NxNy;
    }
    T A, B;
    op_AB(p, f, k2, A, B);
    return {A * Nf.x + B * M.x, A * Nf.y + B * M.y};
}

// Compact twiddle table of the per-field pass: exactly kTab entries (no
// padding to the block's lanes, so that two FP64 blocks of 4096 fit one
// CU's LDS beside their exchange buffers).
template <typename T, int NY, int TPB>
constexpr int colf_tw_entries() {
    return BlockFFT<T, NY, TPB, -1, 1, 1>::kTab;
}
// state loads of the per-field pass in batches of QB k points (U loads each;
// 1 and 4 measured slower or spilled)
#ifndef PFC_COLF_QB
#define PFC_COLF_QB 2
#endif
#ifndef PFC_COLF_WAVES
#define PFC_COLF_WAVES 4
#endif
template <typename T, int NY>
constexpr int colf_waves() { return (sizeof(T) == 8 && NY == 8192) ? 2 : PFC_COLF_WAVES; }

// Per-field column pass of a non-refresh step (L = 2^k columns).  Grid:
// (H/TC + 1) * U blocks; item = xcd_remap(block), column block item / U,
// field item % U.  Columns kx = 0 and kx = H as in col_job.
template <typename T, int NY, int TPB, int TC, int U>
__global__ __launch_bounds__(TPB* TC) PFC_WAVES((colf_waves<T, NY>())) void k_col_step_f(
    const cpx<T>* __restrict__ J, const cpx<T>* __restrict__ Nsrc, cpx<T>* __restrict__ Ndst,
    cpx<T>* __restrict__ I, size_t fstride, size_t nstride, int H, int sI, int sJ,
    const cpx<T>* __restrict__ twy_, const T* __restrict__ ktab_,
    const DevParams* __restrict__ prm, const double* __restrict__ n0src,
    double* __restrict__ n0dst) {
    constexpr int P = NY / TPB;
    constexpr unsigned E = sizeof(cpx<T>);
    using FFTF = BlockFFT<T, NY, TPB, -1, 1, 1, step_padc<T>()>;
    using FFTI = BlockFFT<T, NY, TPB, +1, 1, 1, step_padc<T>()>;
    static_assert(FFTF::POW2, "the per-field pass takes 2^k columns");
    constexpr int NTAB = colf_tw_entries<T, NY, TPB>();
    constexpr int NT = TPB * TC;
    constexpr int NI = (NTAB + NT - 1) / NT;
    __shared__ cpx<T> lds[col_lds_len<T, NY, TPB>() * TC];
    __shared__ __attribute__((aligned(16))) cpx<T> ltw[NTAB];
    const int item = xcd_remap_any(blockIdx.x, gridDim.x);
    // items: (column block, field), field fastest; with J in column pairs
    // (layout x = 1, one column per block) the two columns of a pair are
    // consecutive items for each field, so that the halves of J's lines
    // that the two blocks read are fetched once into their XCD's L2
    int cb, f;
    const int xb = 1 << lay_x(sJ);
    if (TC == 1 && xb > 1 && item < H * U) {
        const int rem = item % (xb * U);
        cb = xb * (item / (xb * U)) + rem % xb;
        f = rem / xb;
    } else {
        cb = item / U;
        f = item % U;
    }
    const int tid = threadIdx.x;
    const int c = tid / TPB;
    const int t = tid % TPB;
    cpx<T>* L = lds + c * FFTF::lds_len;
    const GBuf twy = gbuf(twy_);
    const GBuf ktab = gbuf(ktab_);
    const ColJob jb = col_job(cb, c, TC, H);
    const int kx = jb.kx;
    const unsigned vr = lay_col_off(sJ, H, jb.mode ? 0 : kx, t, E);
    const unsigned rstep = lay_col_step(sJ, H, TPB, E);
    const unsigned vn = ((unsigned)kx * NY + (unsigned)t) * E;
    const unsigned vi = lay_col_off(sI, H, jb.mode ? 0 : kx, t, E);
    const unsigned istep = lay_col_step(sI, H, TPB, E);

    // FP64 columns of 4096: the waves issue their loads (table, J, each
    // batch of states) at a raised priority, so that on a CU shared by two
    // blocks the memory requests of one go out before the other block's
    // transform work (4096^2, three fields: 394 -> 390 us).  Columns of 8192
    // lost with it (FP64 1.746 -> 1.762 ms, FP32 0.832 -> 0.857 ms).
#ifndef PFC_COLF_PRIO
#define PFC_COLF_PRIO 2
#endif
    constexpr int PRIO = (sizeof(T) == 8 && NY == 4096) ? PFC_COLF_PRIO : 0;
    if (PRIO) __builtin_amdgcn_s_setprio(PRIO);
    // the table gather first, then the field's J (the LDS write of the
    // table waits for the gather only)
    cpx<T> tv[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int e = tid + i * NT;
        tv[i] = gload<T>(twy, (unsigned)(e < NTAB ? FFTF::tab_src(e) : 0) * E, 0u);
    }
    const T gx = gload_real<T>(ktab, (unsigned)kx * sizeof(T));
    __builtin_amdgcn_sched_barrier(0);
    cpx<T> M[P];
    {
        const GBuf src = gbuf(J + f * fstride);
#pragma unroll
        for (int q = 0; q < P; q++) M[q] = col_in<T>(src, vr, q * rstep, jb.mode);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (PRIO) __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int e = tid + i * NT;
        if (e < NTAB) ltw[e] = tv[i];
    }
    __syncthreads();
    FFTF::run(M, L, StageTab<T>{ltw}, t);
    if (jb.live) {
        const double kxd = kx * prm->dkx;
        const double kx2 = kxd * kxd;
        // state loads in batches of QB k points, U fields each (issuing each
        // batch one batch ahead measured 3 % slower at 4096^2)
        constexpr int QB = (P % PFC_COLF_QB == 0) ? PFC_COLF_QB : 1;
        const GBuf nd = gbuf(Ndst + f * nstride);
#pragma unroll
        for (int q0 = 0; q0 < P; q0 += QB) {
            cpx<T> Nb[QB][U];
            T gyb[QB];
            if (PRIO) __builtin_amdgcn_s_setprio(PRIO);
#pragma unroll
            for (int i = 0; i < QB; i++) {
#pragma unroll
                for (int v = 0; v < U; v++)
                    Nb[i][v] = gload<T>(gbuf(Nsrc + v * nstride), vn, (q0 + i) * TPB * E);
                gyb[i] = gload_real<T>(ktab, (unsigned)(ktab_gy(H) + t + (q0 + i) * TPB) * sizeof(T));
            }
            if (PRIO) __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int i = 0; i < QB; i++) {
                const int q = q0 + i;
                const int j = t + q * TPB;
                const double ky = ky_of(j, NY, prm->dky);
                const T k2 = (T)(kx2 + ky * ky);
                bool k0 = false;
                if constexpr (sizeof(T) == 4) k0 = (jb.mode == 1 && j == 0);
                if (k0) {
                    // k = 0 in FP64 from the FP64 states of all fields (n0src),
                    // as the in-place pass; field f's result to n0dst
                    cpx<double> N0[U];
#pragma unroll
                    for (int v = 0; v < U; v++) N0[v] = {n0src[2 * v], n0src[2 * v + 1]};
                    const cpx<double> r = kupdate_one<U, double>(
                        prm, f, 0.0, 1.0, cpx<double>{(double)M[q].x, (double)M[q].y}, N0);
                    n0dst[2 * f] = r.x;
                    n0dst[2 * f + 1] = r.y;
                    M[q] = {(T)r.x, (T)r.y};
                } else {
                    M[q] = kupdate_one<U, T>(prm, f, k2, gx * gyb[i], M[q], Nb[i]);
                }
                gstore<T, PFC_COL_AUX>(nd, vn, q * TPB * E, M[q]);
            }
        }
    }
    if constexpr (sizeof(T) == 4) {
        // kx = 0 and kx = H: the stored state replaced by its Hermitian part
        if (cb == 0 || cb == H / TC) {
            cpx<T>(&MM)[1][P] = *reinterpret_cast<cpx<T>(*)[1][P]>(&M);
            hermitian_cols<T, NY, TPB, 1>(MM, L, t, jb.live ? jb.mode : 0);
            if (jb.live && jb.mode != 0) {
                const GBuf nd = gbuf(Ndst + f * nstride);
#pragma unroll
                for (int q = 0; q < P; q++) gstore<T>(nd, vn, q * TPB * E, M[q]);
            }
        }
    }
    FFTI::run(M, L, StageTab<T>{ltw}, t);
    if (jb.live) {
        const GBuf dst = gbuf(I + f * fstride);
#pragma unroll
        for (int q = 0; q < P; q++) col_out<T, PFC_COL_AUX>(dst, vi, q * istep, jb.mode, M[q]);
    }
}

// ------------------------------------------------ single-buffer row passes

// Row C2R of nbuf buffers (blockIdx.y selects the buffer): src spectral ->
// dst real rows (Ny x Nx, row-major).  If K is given, also writes R2C of the
// result (the density's row transform, used for N = FFT(n)/NxNy).
template <typename T, int NX, int TPB, int TR>
__global__ __launch_bounds__(TPB* TR) void k_row_c2r(const cpx<T>* __restrict__ src,
                                                     T* __restrict__ dst, cpx<T>* __restrict__ K,
                                                     size_t fstride, size_t rstride, int sS,
                                                     int sK, const cpx<T>* __restrict__ twx_) {
    const GBuf twx = gbuf(twx_);
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    __shared__ cpx<T> lds[H * TR];
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    const int y = xcd_remap(blockIdx.x, gridDim.x) * TR + r;
    const int b = blockIdx.y;
    cpx<T>* L = lds + r;
    constexpr unsigned E = sizeof(cpx<T>);
    const unsigned soff = lay_row_off(sS, H, t, y, E), sstep = lay_row_step(sS, TPB, E);
    const unsigned voff = lay_row_off(sK, H, t, y, E), qstep = lay_row_step(sK, TPB, E);
    cpx<T> v[P];
    const GBuf sb = gbuf(src + b * fstride);
#pragma unroll
    for (int q = 0; q < P; q++) v[q] = gload<T>(sb, soff, q * sstep);
    const cpx<T> wt = gload<T>(twx, (unsigned)t * E, 0u);
    c2r_pre<T, H, TPB, TR>(v, L, wt, t);
    BlockFFT<T, H, TPB, +1, 2, TR>::run(v, L, twx, t);
    const GBuf db = gbuf(dst + b * rstride);
    const unsigned roff = ((unsigned)y * NX / 2 + (unsigned)t) * E;
#pragma unroll
    for (int q = 0; q < P; q++) gstore<T>(db, roff, q * TPB * E, v[q]);
    if (K) {
        BlockFFT<T, H, TPB, -1, 2, TR>::run(v, L, twx, t);
        r2c_post<T, H, TPB, TR>(v, L, wt, t);
        const GBuf kb = gbuf(K + b * fstride);
#pragma unroll
        for (int q = 0; q < P; q++) gstore<T>(kb, voff, q * qstep, v[q]);
    }
}

// Row R2C of nbuf buffers: real rows -> spectral, times scale (1/Nx gives
// the row state I, whose C2R is the density itself).
template <typename T, int NX, int TPB, int TR>
__global__ __launch_bounds__(TPB* TR) void k_row_r2c(const T* __restrict__ src,
                                                     cpx<T>* __restrict__ dst, size_t fstride,
                                                     size_t rstride, int sD,
                                                     const cpx<T>* __restrict__ twx_, T scale) {
    const GBuf twx = gbuf(twx_);
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    __shared__ cpx<T> lds[H * TR];
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    const int y = xcd_remap(blockIdx.x, gridDim.x) * TR + r;
    const int b = blockIdx.y;
    cpx<T>* L = lds + r;
    constexpr unsigned E = sizeof(cpx<T>);
    cpx<T> v[P];
    const GBuf sb = gbuf(src + b * rstride);
    const unsigned roff = ((unsigned)y * NX / 2 + (unsigned)t) * E;
#pragma unroll
    for (int q = 0; q < P; q++) v[q] = gload<T>(sb, roff, q * TPB * E);
    BlockFFT<T, H, TPB, -1, 2, TR>::run(v, L, twx, t);
    r2c_post<T, H, TPB, TR>(v, L, gload<T>(twx, (unsigned)t * E, 0u), t);
    const GBuf db = gbuf(dst + b * fstride);
    const unsigned voff = lay_row_off(sD, H, t, y, E), qstep = lay_row_step(sD, TPB, E);
#pragma unroll
    for (int q = 0; q < P; q++) gstore<T>(db, voff, q * qstep, cscale(v[q], scale));
}

// Column forward FFT of nbuf buffers with a scale, from packed row-spectral
// buffers (H columns) into unpacked state buffers (H+1 columns):
// dst = scale * FFT_y(src).
template <typename T, int NY, int TPB, int TC>
__global__ __launch_bounds__(TPB* TC) void k_col_fwd(const cpx<T>* __restrict__ src,
                                                     cpx<T>* __restrict__ dst, size_t fstride,
                                                     size_t nstride, int H, int sS,
                                                     const cpx<T>* __restrict__ twy_,
                                                     double scale) {
    const GBuf twy = gbuf(twy_);
    constexpr int P = NY / TPB;
    constexpr unsigned E = sizeof(cpx<T>);
    __shared__ cpx<T> lds[NY * TC];
    const int c = threadIdx.x / TPB;
    const int t = threadIdx.x % TPB;
    const int b = blockIdx.y;
    cpx<T>* L = lds + c * NY;
    const GBuf sb = gbuf(src + b * fstride), db = gbuf(dst + b * nstride);
    {
        const ColJob jb = col_job(blockIdx.x, c, TC, H);
        const unsigned vr = lay_col_off(sS, H, jb.mode ? 0 : jb.kx, t, E);
        const unsigned rstep = lay_col_step(sS, H, TPB, E);
        const unsigned vn = ((unsigned)jb.kx * NY + (unsigned)t) * E;
        cpx<T> v[P];
#pragma unroll
        for (int q = 0; q < P; q++) v[q] = col_in<T>(sb, vr, q * rstep, jb.mode);
        BlockFFT<T, NY, TPB, -1, 1, 1>::run(v, L, twy, t);
        if (jb.live) {
#pragma unroll
            for (int q = 0; q < P; q++) gstore<T>(db, vn, q * TPB * E, cscale(v[q], (T)scale));
        }
    }
}

// Gaussian smoothing along y for one packed row-spectral buffer (smooth.c:
// src/smooth.c:82-103 multiplies the 2D spectrum by exp(-c k^2)/(W H)): src
// holds R2C_x(f)/Nx; dst = IFFT_y(FFT_y(src) exp(-c k2)/Ny), so that the row
// C2R of dst is the smoothed field.  The kx = 0 and kx = H columns (modes 1,
// 2) keep only the real parts, as a 1D c2r does.
template <typename T, int NY, int TPB, int TC>
__global__ __launch_bounds__(TPB* TC) void k_col_smooth(const cpx<T>* __restrict__ src,
                                                        cpx<T>* __restrict__ dst, int H, int sS,
                                                        const cpx<T>* __restrict__ twy_,
                                                        double dkx, double dky, double c) {
    const GBuf twy = gbuf(twy_);
    constexpr int P = NY / TPB;
    constexpr unsigned E = sizeof(cpx<T>);
    __shared__ cpx<T> lds[NY * TC];
    const int cl = threadIdx.x / TPB;
    const int t = threadIdx.x % TPB;
    cpx<T>* L = lds + cl * NY;
    const GBuf sb = gbuf(src), db = gbuf(dst);
    const ColJob jb = col_job(blockIdx.x, cl, TC, H);
    const unsigned vr = lay_col_off(sS, H, jb.mode ? 0 : jb.kx, t, E);
    const unsigned rstep = lay_col_step(sS, H, TPB, E);
    cpx<T> v[P];
#pragma unroll
    for (int q = 0; q < P; q++) v[q] = col_in<T>(sb, vr, q * rstep, jb.mode);
    BlockFFT<T, NY, TPB, -1, 1, 1>::run(v, L, twy, t);
    const double kx = jb.kx * dkx;
    const double kx2 = kx * kx;
#pragma unroll
    for (int q = 0; q < P; q++) {
        const double ky = ky_of(t + q * TPB, NY, dky);
        v[q] = cscale(v[q], (T)(exp_cw(-c * (kx2 + ky * ky)) / NY));
    }
    BlockFFT<T, NY, TPB, +1, 1, 1>::run(v, L, twy, t);
    if (jb.live) {
#pragma unroll
        for (int q = 0; q < P; q++) col_out<T>(db, vr, q * rstep, jb.mode, v[q]);
    }
}


// Evaluate's k-space pass for source field s = blockIdx.y: M = FFT_y(K_s)/NxNy
// (also stored as the refreshed state Nh_s, src/pfc-relax.c:1124-1128), then
// every output that multiplies M_s -- P_s M_s, G M_s and (c2_us - k2)^2 M_s
// for u < s -- is transformed back along y into O (packed like K).
// O buffer order: [P_0..P_{U-1}, G_0..G_{U-1}, q_uv for u<v in row order].
template <typename T, int NY, int TPB, int TC>
__global__ __launch_bounds__(TPB* TC) void k_col_eval(const cpx<T>* __restrict__ K,
                                                      cpx<T>* __restrict__ Nh,
                                                      cpx<T>* __restrict__ O, size_t fstride,
                                                      size_t nstride, int H, int sK,
                                                      const cpx<T>* __restrict__ twy_,
                                                      const DevParams* __restrict__ prm) {
    const GBuf twy = gbuf(twy_);
    constexpr int P = NY / TPB;
    constexpr unsigned E = sizeof(cpx<T>);
    __shared__ cpx<T> lds[NY * TC];
    const int c = threadIdx.x / TPB;
    const int t = threadIdx.x % TPB;
    const int s = blockIdx.y;
    const int U = prm->U;
    cpx<T>* L = lds + c * NY;
    const T scale = (T)prm->inv_NxNy;
    const GBuf kb = gbuf(K + s * fstride), nb = gbuf(Nh + s * nstride);
    {
        const ColJob jb = col_job(blockIdx.x, c, TC, H);
        const unsigned vk = lay_col_off(sK, H, jb.mode ? 0 : jb.kx, t, E);
        const unsigned kstep = lay_col_step(sK, H, TPB, E);
        const unsigned vr = ((unsigned)(jb.mode ? 0 : jb.kx) * lay_pitch(ilog2c(NY), E) + (unsigned)t) * E;
        const unsigned vn = ((unsigned)jb.kx * NY + (unsigned)t) * E;
        cpx<T> M[P];
#pragma unroll
        for (int q = 0; q < P; q++) M[q] = col_in<T>(kb, vk, q * kstep, jb.mode);
        BlockFFT<T, NY, TPB, -1, 1, 1>::run(M, L, twy, t);
#pragma unroll
        for (int q = 0; q < P; q++) {
            M[q] = cscale(M[q], scale);
            if (jb.live) gstore<T>(nb, vn, q * TPB * E, M[q]);
        }
        const double dky = prm->dky;
        const double kxd = jb.kx * prm->dkx;
        const double kx2 = kxd * kxd;
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
                X[q] = cscale(M[q], (T)eval_mult(prm, kind, fu, c2, kx2 + ky * ky));
            }
            BlockFFT<T, NY, TPB, +1, 1, 1>::run(X, L, twy, t);
            const GBuf ob = gbuf(O + oi * fstride);
            if (jb.live) {
#pragma unroll
                for (int q = 0; q < P; q++) col_out<T>(ob, vr, q * TPB * E, jb.mode, X[q]);
            }
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

// The free energy's sums without its density field, in two parts that add up
// to k_energy's sum of f (src/pfc-relax.c:1093-1117).
//
// Local part, from the densities alone: lambda n + alpha n^2/2 + gamma n^3/3 +
// delta n^4/4 per field and alpha_uv n_u n_v + gamma_uv (n_u^2 n_v + n_u n_v^2)/2
// per pair; per-block compensated partials of that sum and of every n_u, in
// k_energy's layout.
template <typename T, int U, int BS>
__global__ __launch_bounds__(BS) void k_energy_local(const T* __restrict__ n, size_t nr,
                                                     double* __restrict__ partials,
                                                     const DevParams* __restrict__ p) {
    KSum acc[1 + U];
#pragma unroll
    for (int i = 0; i <= U; i++) acc[i] = {0.0, 0.0};
    for (size_t ij = (size_t)blockIdx.x * BS + threadIdx.x; ij < nr; ij += (size_t)gridDim.x * BS) {
        double nv[U];
#pragma unroll
        for (int u = 0; u < U; u++) nv[u] = (double)n[u * nr + ij];
        local_energy<U>(p, nv, acc);
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

// Spectral part: the terms with a filtered field (P_u n_u, the smoothed
// densities e_u = G n_u, the pair gradient term Q_uv) are bilinear, so by
// Parseval their grid mean is a sum over k of the state N = FFT2(n)/(Nx Ny):
//   mean(n_u C2R(m N_v)) = sum over the half spectrum of w m Re(N_u conj N_v)
// with w = 1 for kx = 0 and kx = Nx/2 (Nx even) and 2 for the columns between:
// the real-space path's C2R along x takes the interior columns as 2 Re(.) and
// the real part of the other two, so this holds for any Ny, also where the
// wavevector convention (j < Ny/2 positive) makes m uneven in ky (odd Ny).
// Per k: beta_uu/2 P_u |N_u|^2 and (beta_uv M_uv + epsilon_uv G^2)
// Re(N_u conj N_v) with the real-space path's multipliers (eval_mult).  Nh
// holds nkn = (Nx/2 + 1) Ny entries per field, kx major.  Per-block
// compensated partials of the one sum.
// The densities' row pass of an evaluate without the energy density: per
// row group, the C2R of every field's row state (the densities, kept in
// registers), their R2C into K for the refresh N = FFT(n)/NxNy
// (src/pfc-relax.c:1126-1128), and the local terms of the free energy with
// each field's sum, as k_energy_local computes them, without writing the
// densities to memory.  A grid of at most kEnergyBlocks blocks walks the row
// groups (XCD-contiguous, persist_row_group); each block writes one set of
// partial sums.
template <typename T, int NX, int TPB, int TR, int U>
__global__ __launch_bounds__(TPB* TR) PFC_WAVES((row_waves<T, NX>())) void k_row_c2r_eloc(
    const cpx<T>* __restrict__ I, cpx<T>* __restrict__ K, size_t fstride, int sI, int sK,
    const cpx<T>* __restrict__ twx_, const DevParams* __restrict__ p,
    double* __restrict__ partials, int ngroups) {
    constexpr int H = NX / 2;
    constexpr int P = H / TPB;
    constexpr int BS = TPB * TR;
    constexpr unsigned E = sizeof(cpx<T>);
    using FFTF = BlockFFT<T, H, TPB, -1, 2, TR>;
    using FFTI = BlockFFT<T, H, TPB, +1, 2, TR>;
    __shared__ cpx<T> lds[H * TR];
    const GBuf twx = gbuf(twx_);
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    cpx<T>* L = lds + r;
    const unsigned istep = lay_row_step(sI, TPB, E), kstep = lay_row_step(sK, TPB, E);
    KSum acc[1 + U];
#pragma unroll
    for (int i = 0; i <= U; i++) acc[i] = {0.0, 0.0};
    for (int it = 0;; it++) {
        const int g = persist_row_group(blockIdx.x, gridDim.x, it, ngroups);
        if (g >= ngroups) break;
        const int y = g * TR + r;
        const unsigned ioff = lay_row_off(sI, H, t, y, E), koff = lay_row_off(sK, H, t, y, E);
        cpx<T> m[U][P];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const GBuf src = gbuf(I + u * fstride);
#pragma unroll
            for (int q = 0; q < P; q++) m[u][q] = gload<T>(src, ioff, q * istep);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const cpx<T> wt = gload<T>(twx, (unsigned)t * E, 0u);
            c2r_pre<T, H, TPB, TR>(m[u], L, wt, t);
            FFTI::run(m[u], L, twx, t);
            cpx<T> w[P];
#pragma unroll
            for (int q = 0; q < P; q++) w[q] = m[u][q];
            FFTF::run(w, L, twx, t);
            r2c_post<T, H, TPB, TR>(w, L, wt, t);
            const GBuf kb = gbuf(K + u * fstride);
#pragma unroll
            for (int q = 0; q < P; q++) gstore<T>(kb, koff, q * kstep, w[q]);
        }
        // element q of a lane holds the densities at x = 2(t + q TPB) (re)
        // and x + 1 (im); the same expression as k_energy_local per point
#pragma unroll
        for (int q = 0; q < P; q++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                double nv[U];
#pragma unroll
                for (int u = 0; u < U; u++) nv[u] = (double)(h ? m[u][q].y : m[u][q].x);
                local_energy<U>(p, nv, acc);
            }
        }
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

template <typename T, int U, int BS>
__global__ __launch_bounds__(BS) void k_energy_spec(const cpx<T>* __restrict__ Nh, size_t nkn,
                                                    int Nx, int Ny, double* __restrict__ partials,
                                                    const DevParams* __restrict__ p) {
    KSum acc[1] = {{0.0, 0.0}};
    for (size_t idx = (size_t)blockIdx.x * BS + threadIdx.x; idx < nkn; idx += (size_t)gridDim.x * BS) {
        const int kx = (int)(idx / (unsigned)Ny);
        const int j = (int)(idx - (size_t)kx * Ny);
        const double kxd = kx * p->dkx;
        const double ky = ky_of(j, Ny, p->dky);
        const double k2 = kxd * kxd + ky * ky;
        cpx<T> z[U];
#pragma unroll
        for (int u = 0; u < U; u++) z[u] = Nh[u * nkn + idx];
        kadd(acc[0], spec_energy<U, T>(p, kx, Nx, k2, z));
    }
    block_ksum<1, BS>(acc);
    if (threadIdx.x == 0) {
        partials[2 * (size_t)blockIdx.x] = acc[0].s;
        partials[2 * (size_t)blockIdx.x + 1] = acc[0].c;
    }
}

// Launches both parts: local partials at partials[0 .. 2 (1 + U) nblk), the
// spectral ones after them.
template <typename T, int U>
void launch_energy_sums(const T* n, const cpx<T>* Nh, size_t nr, size_t nkn, int Nx, int Ny,
                        double* partials, double* spec_partials, int nblk, const DevParams* p,
                        hipStream_t s) {
    hipLaunchKernelGGL((k_energy_local<T, U, 256>), dim3(nblk), dim3(256), 0, s, n, nr, partials, p);
    hipLaunchKernelGGL((k_energy_spec<T, U, 256>), dim3(nblk), dim3(256), 0, s, Nh, nkn, Nx, Ny,
                       spec_partials, p);
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

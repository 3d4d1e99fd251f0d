// pfc_fields.hip -- orientation and deformation fields of a density on the
// GPU: the computation of the reference's fields tool (src/fields.c:104-359),
// in FP64, behind pfc_fields() of include/pfc_relax.h.
//
// fields.c chains complex 2D FFTs with pointwise work in real and in k space.
// Each 2D transform here is split like the relaxation step's: a row pass
// (length-W complex FFT along x, TR rows per block) that stores its result
// column-major -- element (kx, y) at kx*P + y, P = H padded by 384 bytes from
// H = 1024 on, as the relaxation's spectral buffers -- and a column pass
// (length-H FFT along y, whole contiguous columns).  The pointwise work between
// the halves of consecutive transforms is fused into the passes:
//
//   orientation (fields.c:104-226)             passes
//     min of n                                 k_fred<MIN> + k_fred_final
//     FFT(n - min)                             row  R_FWD_REAL
//     * exp(a (|k|-1)^2) e^{i fold atan2} / WH  col  C_ORIENT (FFT_y, multiply, IFFT_y)
//     IFFT, * (n - min), FFT                   row  R_MUL
//     * exp(-k^2 / 2 sb^2) / WH  (kept)         col  C_SMOOTH_SAVE
//     IFFT -> phi, / max |phi|                 row  R_OUT_CPX, k_fred<ABS>, k_fscale
//   deformation (fields.c:229-359)
//     i kx Z, i ky Z, IFFT                     col  C_GRAD, row R_GRAD:
//     |grad|^power, FFT                             (|zx|^2 + |zy|^2)^(power/2), FFT_x
//                                              col  C_FWD
//     per scale n = 2, 4, ...: * exp(-n^2 k^2 / 2) / WH, IFFT, real part,
//     / its max, summed                        col  C_LEVEL, row R_OUT_RE,
//                                              k_fred<MAX>, k_faccum
//
// Transforms are unnormalised like FFTW's; every 1/(W H) sits where fields.c
// puts it.  Wave numbers follow fields.c: index i < W/2 -> i dkx, else
// (i - W) dkx (the Nyquist index is negative), likewise for y.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pfc_relax.h"
#include "fft_generic.hpp"
#include "pfc_launch.hpp"

namespace pfc {
namespace {

using C = cpx<double>;
constexpr unsigned E = sizeof(C);

enum { R_FWD_REAL, R_MUL, R_OUT_CPX, R_GRAD, R_OUT_RE };
enum { C_ORIENT, C_SMOOTH_SAVE, C_GRAD, C_FWD, C_LEVEL };
enum { RED_MIN, RED_MAX, RED_ABS };

struct FieldArgs {
    const double* p;      // density, row-major (real space)
    const double* pmin;   // device scalar: its minimum
    double dkx, dky;      // 2 pi / (W dx), 2 pi / (H dy)
    double a;             // Gaussian exponent of the pass (-0.5/sk^2, -0.5/sb^2, -0.5 n^2)
    double divA;          // 1 / (W H)
    double fold;          // order of rotational symmetry
    double hpow;          // power / 2
};

__device__ __forceinline__ double wave(int i, int n, double dk) {
    return (i < n / 2 ? i : i - n) * dk;
}

// Row pass over TR rows per block: a length-W complex transform per row.
template <int W, int TPB, int TR, int MODE>
__global__ __launch_bounds__(TPB* TR) void k_frow(const C* __restrict__ src,
                                                  const C* __restrict__ src2, void* __restrict__ dst,
                                                  int Hp, FieldArgs fa, const C* __restrict__ twx_) {
    constexpr int P = W / TPB;
    __shared__ C lds[W * TR];
    const GBuf twx = gbuf(twx_);
    const int tid = threadIdx.x;
    const int r = tid % TR;
    const int t = tid / TR;
    const int y = blockIdx.x * TR + r;
    C* L = lds + r;
    // column-major spectral element (kx = t + q TPB, y); row-major element (x, y)
    const unsigned coff = ((unsigned)t * (unsigned)Hp + (unsigned)y) * E;
    const unsigned cstep = (unsigned)TPB * (unsigned)Hp * E;
    const unsigned rix = (unsigned)y * W + (unsigned)t;
    C v[P];
    if constexpr (MODE == R_FWD_REAL) {
        const GBuf pb = gbuf(fa.p);
        const double m = *fa.pmin;
#pragma unroll
        for (int q = 0; q < P; q++)
            v[q] = {gload_real<double>(pb, (rix + q * TPB) * sizeof(double)) - m, 0.0};
        BlockFFT<double, W, TPB, -1, 1, TR>::run(v, L, twx, t);
        const GBuf db = gbuf(dst);
#pragma unroll
        for (int q = 0; q < P; q++) gstore<double>(db, coff, q * cstep, v[q]);
    } else {
        const GBuf sb = gbuf(src);
#pragma unroll
        for (int q = 0; q < P; q++) v[q] = gload<double>(sb, coff, q * cstep);
        BlockFFT<double, W, TPB, +1, 1, TR>::run(v, L, twx, t);
        if constexpr (MODE == R_MUL) {
            // z *= n - min (fields.c:158-164), then forward again
            const GBuf pb = gbuf(fa.p);
            const double m = *fa.pmin;
#pragma unroll
            for (int q = 0; q < P; q++)
                v[q] = cscale(v[q], gload_real<double>(pb, (rix + q * TPB) * sizeof(double)) - m);
            BlockFFT<double, W, TPB, -1, 1, TR>::run(v, L, twx, t);
            const GBuf db = gbuf(dst);
#pragma unroll
            for (int q = 0; q < P; q++) gstore<double>(db, coff, q * cstep, v[q]);
        } else if constexpr (MODE == R_OUT_CPX) {
            const GBuf db = gbuf(dst);
#pragma unroll
            for (int q = 0; q < P; q++) gstore<double>(db, (rix + q * TPB) * E, 0u, v[q]);
        } else if constexpr (MODE == R_OUT_RE) {
            const GBuf db = gbuf(dst);
#pragma unroll
            for (int q = 0; q < P; q++)
                gstore_real<double>(db, (rix + q * TPB) * sizeof(double), 0u, v[q].x);
        } else {  // R_GRAD: |grad phi|^power from the two gradient components (fields.c:293-301)
            C w[P];
            const GBuf sb2 = gbuf(src2);
#pragma unroll
            for (int q = 0; q < P; q++) w[q] = gload<double>(sb2, coff, q * cstep);
            BlockFFT<double, W, TPB, +1, 1, TR>::run(w, L, twx, t);
#pragma unroll
            for (int q = 0; q < P; q++) {
                const double u = hypot(v[q].x, v[q].y);
                const double s = hypot(w[q].x, w[q].y);
                v[q] = {pow(u * u + s * s, fa.hpow), 0.0};
            }
            BlockFFT<double, W, TPB, -1, 1, TR>::run(v, L, twx, t);
            const GBuf db = gbuf(dst);
#pragma unroll
            for (int q = 0; q < P; q++) gstore<double>(db, coff, q * cstep, v[q]);
        }
    }
}

// Column pass over TC columns per block: a length-H transform per column kx.
template <int NH, int TPB, int TC, int MODE>
__global__ __launch_bounds__(TPB* TC) void k_fcol(const C* __restrict__ src, C* __restrict__ dst,
                                                  C* __restrict__ dst2, int W, int Hp, FieldArgs fa,
                                                  const C* __restrict__ twy_) {
    constexpr int P = NH / TPB;
    __shared__ C lds[NH * TC];
    const GBuf twy = gbuf(twy_);
    const int c = threadIdx.x / TPB;
    const int t = threadIdx.x % TPB;
    const int kxi = blockIdx.x * TC + c;
    C* L = lds + c * NH;
    const unsigned off = ((unsigned)kxi * (unsigned)Hp + (unsigned)t) * E;
    constexpr unsigned step = (unsigned)TPB * E;
    const double kx = wave(kxi, W, fa.dkx);
    const double kx2 = kx * kx;
    const GBuf sb = gbuf(src), db = gbuf(dst);
    C v[P];
#pragma unroll
    for (int q = 0; q < P; q++) v[q] = gload<double>(sb, off, q * step);
    if constexpr (MODE == C_ORIENT || MODE == C_SMOOTH_SAVE || MODE == C_FWD)
        BlockFFT<double, NH, TPB, -1, 1, 1>::run(v, L, twy, t);
    if constexpr (MODE == C_FWD) {
#pragma unroll
        for (int q = 0; q < P; q++) gstore<double>(db, off, q * step, v[q]);
        return;
    }
    if constexpr (MODE == C_GRAD) {
        // zx = i kx Z, zy = i ky Z (fields.c:279-291), both back along y
        C w[P];
#pragma unroll
        for (int q = 0; q < P; q++) {
            const double ky = wave(t + q * TPB, NH, fa.dky);
            w[q] = {-ky * v[q].y, ky * v[q].x};
            v[q] = {-kx * v[q].y, kx * v[q].x};
        }
        BlockFFT<double, NH, TPB, +1, 1, 1>::run(v, L, twy, t);
#pragma unroll
        for (int q = 0; q < P; q++) gstore<double>(db, off, q * step, v[q]);
        BlockFFT<double, NH, TPB, +1, 1, 1>::run(w, L, twy, t);
        const GBuf d2 = gbuf(dst2);
#pragma unroll
        for (int q = 0; q < P; q++) gstore<double>(d2, off, q * step, w[q]);
        return;
    }
#pragma unroll
    for (int q = 0; q < P; q++) {
        const double ky = wave(t + q * TPB, NH, fa.dky);
        if constexpr (MODE == C_ORIENT) {
            // exp(a q^2) e^{i fold phi} / (W H), q = |k| - 1, phi = atan2(ky, kx)
            // (fields.c:142-154)
            const double phi = atan2(ky, kx);
            const double qq = sqrt(kx2 + ky * ky) - 1.0;
            const double g = exp(fa.a * qq * qq);
            const double ang = fa.fold * phi;
            const C k = {g * cos(ang) * fa.divA, g * sin(ang) * fa.divA};
            v[q] = cmul(v[q], k);
        } else {
            // Gaussian exp(a k^2) / (W H) (fields.c:168-180, 323-337)
            v[q] = cscale(v[q], exp(fa.a * (kx2 + ky * ky)) * fa.divA);
        }
    }
    if constexpr (MODE == C_SMOOTH_SAVE) {
        const GBuf d2 = gbuf(dst2);
#pragma unroll
        for (int q = 0; q < P; q++) gstore<double>(d2, off, q * step, v[q]);
    }
    BlockFFT<double, NH, TPB, +1, 1, 1>::run(v, L, twy, t);
#pragma unroll
    for (int q = 0; q < P; q++) gstore<double>(db, off, q * step, v[q]);
}

// ------------------------------------------------------------ any size
//
// Sides the tuned transforms do not cover run the same passes on the generic
// runtime-planned transform (fft_generic.hpp): one block of kGenF lanes per
// row or column with the sequence in LDS (two buffers of gen_lds_elems / 2),
// the pointwise work in strided loops.  The layouts are those of the tuned
// passes, with column pitch Hp = H.
constexpr int kGenF = 256;

template <int MODE>
__global__ __launch_bounds__(kGenF) void k_gfrow(const C* __restrict__ src, const C* __restrict__ src2,
                                                 void* __restrict__ dst, int W, int Hp, FieldArgs fa,
                                                 GenTables<double> gx) {
    extern __shared__ unsigned char smem_gf[];
    C* a = (C*)smem_gf;
    C* b = a + gen_lds_elems(gx.p) / 2;
    const int y = blockIdx.x;
    const size_t rix = (size_t)y * W;
    C* d = (C*)dst;
    if constexpr (MODE == R_FWD_REAL) {
        const double m = *fa.pmin;
        for (int x = threadIdx.x; x < W; x += blockDim.x) a[x] = {fa.p[rix + x] - m, 0.0};
        const C* r = gen_fft<-1, double>(a, b, gx);
        for (int k = threadIdx.x; k < W; k += blockDim.x) d[(size_t)k * Hp + y] = r[k];
        return;
    }
    for (int k = threadIdx.x; k < W; k += blockDim.x) a[k] = src[(size_t)k * Hp + y];
    C* r = gen_fft<+1, double>(a, b, gx);
    C* o = (r == a) ? b : a;
    if constexpr (MODE == R_MUL) {
        // z *= n - min (fields.c:158-164), then forward again
        const double m = *fa.pmin;
        for (int x = threadIdx.x; x < W; x += blockDim.x) r[x] = cscale(r[x], fa.p[rix + x] - m);
        const C* f = gen_fft<-1, double>(r, o, gx);
        for (int k = threadIdx.x; k < W; k += blockDim.x) d[(size_t)k * Hp + y] = f[k];
    } else if constexpr (MODE == R_OUT_CPX) {
        for (int x = threadIdx.x; x < W; x += blockDim.x) d[rix + x] = r[x];
    } else if constexpr (MODE == R_OUT_RE) {
        double* dr = (double*)dst;
        for (int x = threadIdx.x; x < W; x += blockDim.x) dr[rix + x] = r[x].x;
    } else {
        // R_GRAD: |grad phi|^power (fields.c:293-301).  |zx| waits in this
        // row's own slots of dst (the block has read them; each lane reads
        // back the slots it wrote) while zy is transformed.
        for (int x = threadIdx.x; x < W; x += blockDim.x) d[(size_t)x * Hp + y] = {hypot(r[x].x, r[x].y), 0.0};
        for (int k = threadIdx.x; k < W; k += blockDim.x) a[k] = src2[(size_t)k * Hp + y];
        C* w = gen_fft<+1, double>(a, b, gx);
        C* o2 = (w == a) ? b : a;
        for (int x = threadIdx.x; x < W; x += blockDim.x) {
            const double u = d[(size_t)x * Hp + y].x;
            const double sv = hypot(w[x].x, w[x].y);
            w[x] = {pow(u * u + sv * sv, fa.hpow), 0.0};
        }
        const C* f = gen_fft<-1, double>(w, o2, gx);
        for (int k = threadIdx.x; k < W; k += blockDim.x) d[(size_t)k * Hp + y] = f[k];
    }
}

template <int MODE>
__global__ __launch_bounds__(kGenF) void k_gfcol(const C* __restrict__ src, C* __restrict__ dst,
                                                 C* __restrict__ dst2, int W, int NH, int Hp, FieldArgs fa,
                                                 GenTables<double> gy) {
    extern __shared__ unsigned char smem_gc[];
    C* a = (C*)smem_gc;
    C* b = a + gen_lds_elems(gy.p) / 2;
    const int kxi = blockIdx.x;
    const size_t off = (size_t)kxi * Hp;
    const double kx = wave(kxi, W, fa.dkx);
    const double kx2 = kx * kx;
    for (int j = threadIdx.x; j < NH; j += blockDim.x) a[j] = src[off + j];
    C* r = a;
    if constexpr (MODE == C_ORIENT || MODE == C_SMOOTH_SAVE || MODE == C_FWD)
        r = gen_fft<-1, double>(a, b, gy);
    C* o = (r == a) ? b : a;
    if constexpr (MODE == C_FWD) {
        for (int j = threadIdx.x; j < NH; j += blockDim.x) dst[off + j] = r[j];
        return;
    }
    if constexpr (MODE == C_GRAD) {
        // zx = i kx Z, zy = i ky Z (fields.c:279-291), both back along y; zy
        // waits in dst2 (each lane reads back the slots it wrote)
        for (int j = threadIdx.x; j < NH; j += blockDim.x) {
            const double ky = wave(j, NH, fa.dky);
            const C v = r[j];
            dst2[off + j] = {-ky * v.y, ky * v.x};
            r[j] = {-kx * v.y, kx * v.x};
        }
        const C* f = gen_fft<+1, double>(r, o, gy);
        for (int j = threadIdx.x; j < NH; j += blockDim.x) dst[off + j] = f[j];
        __syncthreads();   // every lane is done with f before a is rewritten
        for (int j = threadIdx.x; j < NH; j += blockDim.x) a[j] = dst2[off + j];
        const C* g = gen_fft<+1, double>(a, b, gy);
        for (int j = threadIdx.x; j < NH; j += blockDim.x) dst2[off + j] = g[j];
        return;
    }
    for (int j = threadIdx.x; j < NH; j += blockDim.x) {
        const double ky = wave(j, NH, fa.dky);
        if constexpr (MODE == C_ORIENT) {
            // exp(a q^2) e^{i fold phi} / (W H), q = |k| - 1, phi = atan2(ky, kx)
            // (fields.c:142-154)
            const double phi = atan2(ky, kx);
            const double qq = sqrt(kx2 + ky * ky) - 1.0;
            const double g = exp(fa.a * qq * qq);
            const double ang = fa.fold * phi;
            r[j] = cmul(r[j], C{g * cos(ang) * fa.divA, g * sin(ang) * fa.divA});
        } else {
            // Gaussian exp(a k^2) / (W H) (fields.c:168-180, 323-337)
            r[j] = cscale(r[j], exp(fa.a * (kx2 + ky * ky)) * fa.divA);
        }
        if constexpr (MODE == C_SMOOTH_SAVE) dst2[off + j] = r[j];
    }
    const C* f = gen_fft<+1, double>(r, o, gy);
    for (int j = threadIdx.x; j < NH; j += blockDim.x) dst[off + j] = f[j];
}

// Min of a real array, max of a real array or max modulus of a complex array,
// with fields.c's starting values and strict comparisons (a NaN never wins).
constexpr int kRedBS = 256;
constexpr int kRedBlocks = 1024;
template <int OP>
__device__ __forceinline__ double red_init() {
    return OP == RED_MIN ? 1.0e100 : -1.0e100;
}
template <int OP>
__device__ __forceinline__ double red_op(double acc, double v) {
    if constexpr (OP == RED_MIN) return v < acc ? v : acc;
    else return v > acc ? v : acc;
}
template <int OP, int BS>
__device__ __forceinline__ double red_block(double acc) {
    __shared__ double s[BS];
    s[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int h = BS / 2; h > 0; h /= 2) {
        if ((int)threadIdx.x < h) s[threadIdx.x] = red_op<OP>(s[threadIdx.x], s[threadIdx.x + h]);
        __syncthreads();
    }
    return s[0];
}
template <int OP>
__global__ __launch_bounds__(kRedBS) void k_fred(const double* __restrict__ x, size_t n,
                                                 double* __restrict__ part) {
    double acc = red_init<OP>();
    for (size_t i = (size_t)blockIdx.x * kRedBS + threadIdx.x; i < n; i += (size_t)gridDim.x * kRedBS) {
        const double v = (OP == RED_ABS) ? hypot(x[2 * i], x[2 * i + 1]) : x[i];
        acc = red_op<OP>(acc, v);
    }
    acc = red_block<OP, kRedBS>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}
template <int OP>
__global__ __launch_bounds__(1024) void k_fred_final(const double* __restrict__ part, int nb,
                                                     double* __restrict__ out) {
    double acc = red_init<OP>();
    for (int i = threadIdx.x; i < nb; i += 1024) acc = red_op<OP>(acc, part[i]);
    acc = red_block<OP, 1024>(acc);
    if (threadIdx.x == 0) *out = acc;
}

// z *= 1/max (fields.c:196-205), on n complex values
__global__ __launch_bounds__(256) void k_fscale(double* __restrict__ z, size_t n,
                                                const double* __restrict__ mx) {
    const double d = 1.0 / *mx;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        z[2 * i] *= d;
        z[2 * i + 1] *= d;
    }
}
// chi += level / max (fields.c:340-346); chi = 0 first when clear is set
__global__ __launch_bounds__(256) void k_faccum(double* __restrict__ chi,
                                                const double* __restrict__ lev, size_t n,
                                                const double* __restrict__ mx, int clear) {
    const double d = 1.0 / *mx;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        chi[i] = (clear ? 0.0 : chi[i]) + lev[i] * d;
}
__global__ __launch_bounds__(256) void k_fzero(double* __restrict__ x, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        x[i] = 0.0;
}

// ------------------------------------------------------------ launchers

struct Job {
    int W, H, Hp;
    hipStream_t s;
    const C *twx, *twy;
    // generic transforms (sides the tuned kernels do not cover)
    bool generic;
    GenTables<double> gx, gy;
    size_t ldsx, ldsy;   // dynamic LDS bytes of a row / column block
};

template <int W, int MODE>
void row_w(const Job& j, const C* src, const C* src2, void* dst, const FieldArgs& fa) {
    if constexpr (size_built(W)) {
        constexpr ColCfg g = col_cfg(W);
        hipLaunchKernelGGL((k_frow<W, g.tpb, g.tc, MODE>), dim3(j.H / g.tc), dim3(g.tpb * g.tc), 0,
                           j.s, src, src2, dst, j.Hp, fa, j.twx);
    }
}
template <int MODE>
void row(const Job& j, const C* src, const C* src2, void* dst, const FieldArgs& fa) {
    if (j.generic) {
        hipLaunchKernelGGL(k_gfrow<MODE>, dim3(j.H), dim3(kGenF), j.ldsx, j.s, src, src2, dst, j.W, j.Hp,
                           fa, j.gx);
        return;
    }
    switch (j.W) {
        case 64: row_w<64, MODE>(j, src, src2, dst, fa); break;
        case 128: row_w<128, MODE>(j, src, src2, dst, fa); break;
        case 256: row_w<256, MODE>(j, src, src2, dst, fa); break;
        case 384: row_w<384, MODE>(j, src, src2, dst, fa); break;
        case 512: row_w<512, MODE>(j, src, src2, dst, fa); break;
        case 768: row_w<768, MODE>(j, src, src2, dst, fa); break;
        case 1024: row_w<1024, MODE>(j, src, src2, dst, fa); break;
        case 1536: row_w<1536, MODE>(j, src, src2, dst, fa); break;
        case 2048: row_w<2048, MODE>(j, src, src2, dst, fa); break;
        case 4096: row_w<4096, MODE>(j, src, src2, dst, fa); break;
        case 8192: row_w<8192, MODE>(j, src, src2, dst, fa); break;
    }
}
template <int NH, int MODE>
void col_h(const Job& j, const C* src, C* dst, C* dst2, const FieldArgs& fa) {
    if constexpr (size_built(NH)) {
        constexpr ColCfg g = col_cfg(NH);
        hipLaunchKernelGGL((k_fcol<NH, g.tpb, g.tc, MODE>), dim3(j.W / g.tc), dim3(g.tpb * g.tc), 0,
                           j.s, src, dst, dst2, j.W, j.Hp, fa, j.twy);
    }
}
template <int MODE>
void col(const Job& j, const C* src, C* dst, C* dst2, const FieldArgs& fa) {
    if (j.generic) {
        hipLaunchKernelGGL(k_gfcol<MODE>, dim3(j.W), dim3(kGenF), j.ldsy, j.s, src, dst, dst2, j.W, j.H,
                           j.Hp, fa, j.gy);
        return;
    }
    switch (j.H) {
        case 64: col_h<64, MODE>(j, src, dst, dst2, fa); break;
        case 128: col_h<128, MODE>(j, src, dst, dst2, fa); break;
        case 256: col_h<256, MODE>(j, src, dst, dst2, fa); break;
        case 384: col_h<384, MODE>(j, src, dst, dst2, fa); break;
        case 512: col_h<512, MODE>(j, src, dst, dst2, fa); break;
        case 768: col_h<768, MODE>(j, src, dst, dst2, fa); break;
        case 1024: col_h<1024, MODE>(j, src, dst, dst2, fa); break;
        case 1536: col_h<1536, MODE>(j, src, dst, dst2, fa); break;
        case 2048: col_h<2048, MODE>(j, src, dst, dst2, fa); break;
        case 4096: col_h<4096, MODE>(j, src, dst, dst2, fa); break;
        case 8192: col_h<8192, MODE>(j, src, dst, dst2, fa); break;
    }
}
template <int OP>
void reduce(const Job& j, const double* x, size_t n, double* part, double* out) {
    hipLaunchKernelGGL((k_fred<OP>), dim3(kRedBlocks), dim3(kRedBS), 0, j.s, x, n, part);
    hipLaunchKernelGGL((k_fred_final<OP>), dim3(1), dim3(1024), 0, j.s, part, kRedBlocks, out);
}

void twiddles(std::vector<C>& tw, int n) {
    tw.resize(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < n; m++) {
        const long double a = -2.0L * pi * (long double)m / (long double)n;
        tw[m] = {(double)cosl(a), (double)sinl(a)};
    }
}

struct Buffers {
    std::vector<void*> all;
    int rc = PFC_OK;
    template <typename P>
    P* get(size_t bytes) {
        void* p = nullptr;
        if (rc == PFC_OK && hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            rc = PFC_ENOMEM;
        }
        if (p) all.push_back(p);
        return (P*)p;
    }
    ~Buffers() {
        for (void* p : all) (void)hipFree(p);
    }
};

}  // namespace
}  // namespace pfc

using namespace pfc;

extern "C" int pfc_fields(const double* density, int W, int H, double dx, double dy, int fold,
                          double sk, double sb, double power, double* phi, double* chi) {
    if (!density || !phi) return PFC_EINVAL;
    // the tuned transforms where they exist, else the generic ones
    const bool generic = !(size_ok(W) && size_ok(H));
    if (generic && !(gen_plannable(W, E) && gen_plannable(H, E))) return PFC_ESIZE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFC_EDEVICE;
    const size_t n = (size_t)W * H;
    const int Hp = generic ? H : (int)lay_pitch(ilog2(H), E);
    const size_t nc = (size_t)W * Hp;
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PFC_EDEVICE;
    int rc = PFC_OK;
    GenDim gdx, gdy;
    if (generic) {
        rc = gen_dim_create<double>(gdx, W);
        if (rc == PFC_OK) rc = gen_dim_create<double>(gdy, H);
        const size_t lx = (size_t)gen_lds_elems(gdx.p) * E, ly = (size_t)gen_lds_elems(gdy.p) * E;
#define PFC_GF_LDS(k, b)                                                                         \
    (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(b))
        PFC_GF_LDS(k_gfrow<R_FWD_REAL>, lx);
        PFC_GF_LDS(k_gfrow<R_MUL>, lx);
        PFC_GF_LDS(k_gfrow<R_OUT_CPX>, lx);
        PFC_GF_LDS(k_gfrow<R_GRAD>, lx);
        PFC_GF_LDS(k_gfrow<R_OUT_RE>, lx);
        PFC_GF_LDS(k_gfcol<C_ORIENT>, ly);
        PFC_GF_LDS(k_gfcol<C_SMOOTH_SAVE>, ly);
        PFC_GF_LDS(k_gfcol<C_GRAD>, ly);
        PFC_GF_LDS(k_gfcol<C_FWD>, ly);
        PFC_GF_LDS(k_gfcol<C_LEVEL>, ly);
#undef PFC_GF_LDS
    }
    if (rc == PFC_OK) {
        Buffers b;
        double* dp = b.get<double>(n * sizeof(double));
        C* z1 = b.get<C>(nc * E);
        C* z2 = b.get<C>(nc * E);
        C* zs = b.get<C>(nc * E);
        C* zphi = b.get<C>(n * E);
        double* part = b.get<double>(kRedBlocks * sizeof(double));
        double* scal = b.get<double>(4 * sizeof(double));
        C* twx = b.get<C>(W * E);
        C* twy = b.get<C>(H * E);
        double* lev = chi ? b.get<double>(n * sizeof(double)) : nullptr;
        double* dchi = chi ? b.get<double>(n * sizeof(double)) : nullptr;
        rc = b.rc;
        if (rc == PFC_OK) {
            std::vector<C> tx, ty;
            twiddles(tx, W);
            twiddles(ty, H);
            bool ok = hipMemcpyAsync(twx, tx.data(), W * E, hipMemcpyHostToDevice, s) == hipSuccess;
            ok &= hipMemcpyAsync(twy, ty.data(), H * E, hipMemcpyHostToDevice, s) == hipSuccess;
            ok &= hipMemcpyAsync(dp, density, n * sizeof(double), hipMemcpyHostToDevice, s) ==
                  hipSuccess;
            const double pi = 3.14159265358979323846264338327;
            const Job j{W, H, Hp, s, twx, twy, generic, gen_tables<double>(gdx), gen_tables<double>(gdy),
                        generic ? (size_t)gen_lds_elems(gdx.p) * E : 0,
                        generic ? (size_t)gen_lds_elems(gdy.p) * E : 0};
            FieldArgs fa{};
            fa.p = dp;
            fa.pmin = scal;
            fa.dkx = 2.0 * pi / (dx * W);
            fa.dky = 2.0 * pi / (dy * H);
            fa.divA = 1.0 / W / H;
            fa.fold = (double)fold;
            // orientation field (fields.c:104-226)
            reduce<RED_MIN>(j, dp, n, part, scal);
            row<R_FWD_REAL>(j, nullptr, nullptr, z1, fa);
            fa.a = -0.5 / (sk * sk);
            col<C_ORIENT>(j, z1, z1, nullptr, fa);
            row<R_MUL>(j, z1, nullptr, z2, fa);
            fa.a = -0.5 / (sb * sb);
            col<C_SMOOTH_SAVE>(j, z2, z2, zs, fa);
            row<R_OUT_CPX>(j, z2, nullptr, zphi, fa);
            reduce<RED_ABS>(j, (const double*)zphi, n, part, scal + 1);
            hipLaunchKernelGGL(k_fscale, dim3(1024), dim3(256), 0, s, (double*)zphi, n, scal + 1);
            ok &= hipMemcpyAsync(phi, zphi, n * E, hipMemcpyDeviceToHost, s) == hipSuccess;
            if (chi) {
                // deformation field (fields.c:229-359)
                col<C_GRAD>(j, zs, z1, z2, fa);
                fa.hpow = 0.5 * power;
                row<R_GRAD>(j, z1, z2, z1, fa);
                col<C_FWD>(j, z1, zs, nullptr, fa);
                const double l0 = 4.0 * pi / sqrt(3.0);
                double L = W * dx;
                if (H * dy < L) L = H * dy;
                bool first = true;
                for (int m = 2; l0 * m < 0.5 * L; m *= 2) {
                    const double sg = 1.0 / m;
                    fa.a = -0.5 / (sg * sg);
                    col<C_LEVEL>(j, zs, z1, nullptr, fa);
                    row<R_OUT_RE>(j, z1, nullptr, lev, fa);
                    reduce<RED_MAX>(j, lev, n, part, scal + 2);
                    hipLaunchKernelGGL(k_faccum, dim3(1024), dim3(256), 0, s, dchi, lev, n, scal + 2,
                                       first ? 1 : 0);
                    first = false;
                }
                if (first) hipLaunchKernelGGL(k_fzero, dim3(1024), dim3(256), 0, s, dchi, n);
                ok &= hipMemcpyAsync(chi, dchi, n * sizeof(double), hipMemcpyDeviceToHost, s) ==
                      hipSuccess;
            }
            ok &= hipGetLastError() == hipSuccess;
            ok &= hipStreamSynchronize(s) == hipSuccess;
            if (!ok) rc = PFC_EDEVICE;
        }
    }
    gen_dim_free(gdx);
    gen_dim_free(gdy);
    (void)hipStreamDestroy(s);
    return rc;
}

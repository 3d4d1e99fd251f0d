// fft_generic.hpp -- FFTs of any length for the generic step path.
//
// The tuned kernels (fft_device.hpp) fix the length at compile time and cover
// 2^k and 3 * 2^k.  Every other length runs here, with the length in a plan
// read at run time: one block computes one transform in LDS.
//   direct:    L = product of radices from {8, 4, 2, 3, 5, 7, 11, 13}
//              (Stockham autosort, two LDS buffers of L elements);
//   Bluestein: any other L, as a circular convolution of length M (a power of
//              two, M >= 2L - 1): X[k] = c[k] IFFT_M(FFT_M(x c) * B)[k] / M,
//              c[n] = e^{-i pi n^2 / L} and B = FFT_M(b), b[n] = conj c[|n|]
//              for |n| < L (two LDS buffers of M elements).
// The twiddle table of length L (direct) or M (Bluestein) and the chirp
// tables are computed on the host in extended precision.
#pragma once
#include "fft_device.hpp"

namespace pfc {

constexpr int kGenMaxStages = 32;

struct GenPlan {
    int L;                        // transform length
    int M;                        // Bluestein length, 0 for a direct plan
    int nst;                      // stages of the direct transform (length L, or M)
    int radix[kGenMaxStages];
};

// One dimension's plan and device tables (pfc_generic.hip): gen_dim_create
// plans length L and uploads the twiddles (and Bluestein's chirp tables);
// PFC_ESIZE if L has no plan within the LDS budget.
struct GenDim {
    GenPlan p;
    void* tw = nullptr;
    void* chirp = nullptr;
    void* bhat = nullptr;
};
template <typename T>
int gen_dim_create(GenDim& g, int L);
void gen_dim_free(GenDim& g);

// LDS elements one generic transform needs (two buffers).
__host__ __device__ inline int gen_lds_elems(const GenPlan& p) { return 2 * (p.M ? p.M : p.L); }

// Direct DFT of a runtime-selected radix R over R values v, in place:
// v[s] <- sum_r v[r] w^(r s), w = e^{DIR 2 pi i / R}; w^m is entry m * (n/R)
// of the n-point table tw (conjugated for DIR > 0).
template <int R, int DIR, typename T>
__device__ __forceinline__ void gen_dft(cpx<T> (&v)[16], const cpx<T>* __restrict__ tw, int n) {
    if constexpr (R == 2 || R == 3 || R == 4 || R == 8) {
        cpx<T> x[R];
#pragma unroll
        for (int r = 0; r < R; r++) x[r] = v[r];
        dft<R, DIR, T>(x);
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = x[r];
    } else {
        cpx<T> y[R];
        const int step = n / R;
#pragma unroll
        for (int s = 0; s < R; s++) {
            cpx<T> acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                cpx<T> w = tw[((r * s) % R) * step];
                if constexpr (DIR > 0) w.y = -w.y;
                acc = cadd(acc, cmul(v[r], w));
            }
            y[s] = acc;
        }
#pragma unroll
        for (int s = 0; s < R; s++) v[s] = y[s];
    }
}

// One Stockham stage of radix R (compile time) on an n-point sequence in src
// (LDS) into dst, with Ns the product of the earlier radices.
template <int R, int DIR, typename T>
__device__ __forceinline__ void gen_stage(const cpx<T>* src, cpx<T>* dst, int n, int Ns,
                                          const cpx<T>* __restrict__ tw) {
    const int nR = n / R;
    const int step = n / (Ns * R);   // k r step < n for k < Ns, r < R
    for (int j = threadIdx.x; j < nR; j += blockDim.x) {
        const int k = j % Ns;
        cpx<T> v[16];
#pragma unroll
        for (int r = 0; r < R; r++) {
            cpx<T> x = src[j + r * nR];
            if (r > 0 && k > 0) {
                cpx<T> w = tw[k * r * step];
                if constexpr (DIR > 0) w.y = -w.y;
                x = cmul(x, w);
            }
            v[r] = x;
        }
        gen_dft<R, DIR, T>(v, tw, n);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[base + r * Ns] = v[r];
    }
}

// n-point transform of a (LDS) with scratch b (LDS) along the radices of p;
// returns the buffer that holds the result.  Starts and ends with a barrier.
template <int DIR, typename T>
__device__ cpx<T>* gen_fft_direct(cpx<T>* a, cpx<T>* b, int n, int nst, const int* radix,
                                  const cpx<T>* __restrict__ tw) {
    cpx<T>* src = a;
    cpx<T>* dst = b;
    int Ns = 1;
    __syncthreads();
    for (int s = 0; s < nst; s++) {
        switch (radix[s]) {
            case 2: gen_stage<2, DIR, T>(src, dst, n, Ns, tw); break;
            case 3: gen_stage<3, DIR, T>(src, dst, n, Ns, tw); break;
            case 4: gen_stage<4, DIR, T>(src, dst, n, Ns, tw); break;
            case 5: gen_stage<5, DIR, T>(src, dst, n, Ns, tw); break;
            case 7: gen_stage<7, DIR, T>(src, dst, n, Ns, tw); break;
            case 8: gen_stage<8, DIR, T>(src, dst, n, Ns, tw); break;
            case 11: gen_stage<11, DIR, T>(src, dst, n, Ns, tw); break;
            default: gen_stage<13, DIR, T>(src, dst, n, Ns, tw); break;
        }
        Ns *= radix[s];
        __syncthreads();
        cpx<T>* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// The generic transform's device tables (GenTables::tw has L entries for a
// direct plan, M for Bluestein; chirp and bhat are used by Bluestein only).
template <typename T>
struct GenTables {
    GenPlan p;
    const cpx<T>* tw;
    const cpx<T>* chirp;   // c[n] = e^{-i pi n^2 / L}, n < L
    const cpx<T>* bhat;    // FFT_M(b) / M
};

template <typename T>
inline GenTables<T> gen_tables(const GenDim& g) {
    return {g.p, (const cpx<T>*)g.tw, (const cpx<T>*)g.chirp, (const cpx<T>*)g.bhat};
}

// L-point transform (unnormalised; DIR = -1 forward, +1 inverse) of the L
// values in a[0..L); a and b each hold gen_lds_elems/2 elements.  Returns the
// buffer holding X[0..L).
template <int DIR, typename T>
__device__ cpx<T>* gen_fft(cpx<T>* a, cpx<T>* b, const GenTables<T>& g) {
    const GenPlan& p = g.p;
    if (!p.M) return gen_fft_direct<DIR, T>(a, b, p.L, p.nst, p.radix, g.tw);
    // Bluestein (inverse via conjugation: IFFT(x) = conj FFT(conj x))
    const int L = p.L, M = p.M;
    __syncthreads();
    for (int n = threadIdx.x; n < M; n += blockDim.x) {
        cpx<T> v = {(T)0, (T)0};
        if (n < L) {
            v = a[n];
            if constexpr (DIR > 0) v.y = -v.y;
            v = cmul(v, g.chirp[n]);
        }
        b[n] = v;
    }
    cpx<T>* r = gen_fft_direct<-1, T>(b, a, M, p.nst, p.radix, g.tw);
    cpx<T>* o = (r == a) ? b : a;
    for (int n = threadIdx.x; n < M; n += blockDim.x) r[n] = cmul(r[n], g.bhat[n]);
    r = gen_fft_direct<+1, T>(r, o, M, p.nst, p.radix, g.tw);
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
        cpx<T> v = cmul(r[k], g.chirp[k]);
        if constexpr (DIR > 0) v.y = -v.y;
        r[k] = v;
    }
    __syncthreads();
    return r;
}

}  // namespace pfc

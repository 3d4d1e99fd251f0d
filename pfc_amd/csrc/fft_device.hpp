// fft_device.hpp -- register/LDS Stockham FFT building blocks for gfx950.
//
// A length-L transform is computed by TPB threads, each holding P = L/TPB
// points in registers.  On entry and on exit thread t holds elements
// t + q*TPB (q < P) in natural order, so the first and last stages of every
// transform read/write global memory with unit stride across a wavefront and
// only the L/P-strided middle stages go through LDS.
//
// Stage plan (L = 2^k): one leading stage of radix R0 = L / P^s (when L is
// not a power of P, done in registers with P/R0 butterflies per thread), then
// s stages of radix P (Stockham autosort).  L = 3 * 2^k uses P = 12 (or 6)
// points per thread: a radix-12 (6) stage, then radix-4 and radix-2 stages.  Twiddles come from a table tw[m] =
// exp(-2 pi i m / (L*TWS)) computed on the host in extended precision; the
// inverse uses the conjugate.
#pragma once
#include <hip/hip_runtime.h>

#ifndef PFC_TW_RECURRENCE
#define PFC_TW_RECURRENCE 1
#endif
#ifndef PFC_STORE_B128
#define PFC_STORE_B128 0
#endif

namespace pfc {

template <typename T>
struct cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.x, -a.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// a * (DIR * i)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> cmul_i(cpx<T> a) {
    if constexpr (DIR > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}

// cos/sin(2 pi k / 96), k = 0..95: every root of unity of an order dividing 96
// (2, 3, 4, 6, 8, 12, 16, 24, 32), non-recursive so it folds after unrolling
__host__ __device__ constexpr double c96(int k) {
    constexpr double t[25] = {1.0,
                              9.978589232386035067381e-1,
                              9.914448613738104111446e-1,
                              9.807852804032304491262e-1,
                              9.659258262890682867497e-1,
                              9.469301294951056642558e-1,
                              9.238795325112867561282e-1,
                              8.968727415326883038941e-1,
                              8.660254037844386467637e-1,
                              8.314696123025452370788e-1,
                              7.933533402912351645798e-1,
                              7.518398074789773964075e-1,
                              7.071067811865475244008e-1,
                              6.593458151000688684251e-1,
                              6.087614290087206394161e-1,
                              5.555702330196022247428e-1,
                              5.0e-1,
                              4.422886902190012819952e-1,
                              3.826834323650897717285e-1,
                              3.214394653031615807011e-1,
                              2.588190451025207623489e-1,
                              1.950903220161282678483e-1,
                              1.305261922200515915484e-1,
                              6.540312923014306681532e-2,
                              0.0};
    k %= 96;
    if (k < 0) k += 96;
    if (k <= 24) return t[k];
    if (k <= 48) return -t[48 - k];
    if (k <= 72) return -t[k - 48];
    return t[96 - k];
}
__host__ __device__ constexpr double s96(int k) { return c96(k - 24); }

// exp(DIR 2 pi i k / n) for n dividing 96, folded at compile time
template <typename T>
__device__ __forceinline__ cpx<T> root(int k, int n, int dir) {
    const int m = k * (96 / n);
    return {(T)c96(m), (T)(dir * s96(m))};
}

// In-register DFT of size R (1, 2, 3, 4, 6, 8, 12, 16):
// v[k] <- sum_r v[r] e^{DIR 2 pi i rk/R}
template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cpx<T>* v) {
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        // X1,2 = v0 - (v1 + v2)/2 +- DIR i sin(2 pi/3) (v1 - v2)
        const T s3 = (T)8.660254037844386467637e-1;
        const cpx<T> t1 = cadd(v[1], v[2]);
        const cpx<T> t2 = csub(v[1], v[2]);
        const cpx<T> m1 = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
        const cpx<T> m2 = cmul_i<DIR>(cpx<T>{s3 * t2.x, s3 * t2.y});
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m1, m2);
        v[2] = csub(m1, m2);
    } else {
        constexpr int H = R / 2;
        cpx<T> e[H], o[H];
#pragma unroll
        for (int i = 0; i < H; i++) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        dft<H, DIR, T>(e);
        dft<H, DIR, T>(o);
#pragma unroll
        for (int k = 0; k < H; k++) {
            cpx<T> t;
            if (k == 0) {
                t = o[k];
            } else if (4 * k == R) {
                t = cmul_i<DIR>(o[k]);
            } else {
                t = cmul(o[k], root<T>(k, R, DIR));
            }
            v[k] = cadd(e[k], t);
            v[k + H] = csub(e[k], t);
        }
    }
}

// ------------------------------------------------ buffer-resource access
//
// Global memory is addressed through buffer resources: a 128-bit descriptor
// in SGPRs, a 32-bit per-lane byte offset (one VGPR shared by every access
// of the same lane pattern) and a uniform byte offset in an SGPR.  This keeps
// 64-bit per-access address pairs out of the register file.

struct GBuf {
    __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ GBuf gbuf(const void* p) {
    // num_records = 0xffffffff (whole 4 GiB window), gfx9 word 3 = 0x00020000
    return {__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0xffffffff,
                                             0x00020000)};
}

template <typename T>
__device__ __forceinline__ cpx<T> gload(GBuf b, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ cpx<double> gload<double>(GBuf b, unsigned voff, unsigned soff) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(b.r, (int)voff, (int)soff, 0);
    return __builtin_bit_cast(cpx<double>, v);
}
template <>
__device__ __forceinline__ cpx<float> gload<float>(GBuf b, unsigned voff, unsigned soff) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)voff, (int)soff, 0);
    return __builtin_bit_cast(cpx<float>, v);
}
// Stores a complex element.  AUX is the buffer instruction's cache-policy
// operand (gfx950: bit 1 = nt, non-temporal, for data the same pass does not
// read again).
template <typename T, int AUX = 0>
__device__ __forceinline__ void gstore(GBuf b, unsigned voff, unsigned soff, cpx<T> x) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(b.r, 0, 0, 0));
    if constexpr (sizeof(T) == 8) {
#if PFC_STORE_B128 == 1
        using W = decltype(__builtin_amdgcn_raw_buffer_load_b128(b.r, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(W, x), b.r, (int)voff, (int)soff, AUX);
#elif PFC_STORE_B128 == 2
        using W = decltype(__builtin_amdgcn_raw_buffer_load_b128(b.r, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(W, x), b.r, (int)(voff + soff), 0, AUX);
#else
        // Two 64-bit stores of the same bytes instead of one 128-bit store: a
        // 128-bit buffer store with an SGPR soffset whose data VGPRs are
        // rewritten by the next VALU instructions has stored corrupted upper
        // halves on gfx950 (the compiler inserts no wait states for that
        // case).  The second half's address goes through an opaque register
        // copy so the backend cannot merge the pair back into a 128-bit store.
        unsigned v2 = voff + 8u;
        asm volatile("" : "+v"(v2));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, x.x), b.r, (int)voff, (int)soff, AUX);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, x.y), b.r, (int)v2, (int)soff, AUX);
#endif
    } else {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, x), b.r, (int)voff, (int)soff, AUX);
    }
}

// one real value
template <typename T>
__device__ __forceinline__ T gload_real(GBuf b, unsigned voff);
template <>
__device__ __forceinline__ double gload_real<double>(GBuf b, unsigned voff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)voff, 0, 0));
}
template <>
__device__ __forceinline__ float gload_real<float>(GBuf b, unsigned voff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)voff, 0, 0));
}

// one real value (the real or imaginary half of a complex element)
template <typename T, int AUX = 0>
__device__ __forceinline__ void gstore_real(GBuf b, unsigned voff, unsigned soff, T x) {
    if constexpr (sizeof(T) == 8) {
        using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(b.r, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, x), b.r, (int)voff, (int)soff, AUX);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, x), b.r, (int)voff, (int)soff, AUX);
    }
}

#ifndef PFC_LDS_SWIZZLE
#define PFC_LDS_SWIZZLE 1
#endif
// FP64 transforms interleaved up to this many per LDS buffer are swizzled
#ifndef PFC_SWZ_LS
#define PFC_SWZ_LS 2
#endif
constexpr int ilog(int b, int n) { return n < b ? 0 : 1 + ilog(b, n / b); }
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// Compact per-stage twiddle table in LDS (BlockFFT::kTab entries, filled
// from tab_src), read by stage-local index so that a wave's reads are
// contiguous or broadcast.
template <typename T>
struct StageTab {
    const cpx<T>* p;
};

// One length-L transform over TPB cooperating threads (thread index t).
// lds must hold L elements at lds[idx * LS] (LS = element stride, lets
// several interleaved transforms share a buffer without bank conflicts).
template <typename T, int L, int TPB, int DIR, int TWS, int LS, int PADC = 0>
struct BlockFFT {
    static constexpr int P = L / TPB;
    static_assert(P * TPB == L, "L must be a multiple of TPB");
    static constexpr bool POW2 = (L & (L - 1)) == 0;
    static_assert(POW2 ? (P == 2 || P == 4 || P == 8 || P == 16) : (P == 6 || P == 12),
                  "P must be 2, 4, 8 or 16 (L = 2^k) or 6, 12 (L = 3 * 2^k)");
    static constexpr int S = ilog(P, L);          // number of radix-P stages (L = 2^k)
    static constexpr int R0 = L / ipow(P, S);     // leading stage radix (1 if none)

    // LDS index swizzle of the exchanges (L = 2^k).  LDS serves 256 bytes
    // (64 banks) per cycle, i.e. 2^SWB consecutive indices of one transform
    // when LS transforms interleave.  The Stockham writes of the early stages
    // are strided (index j*R + r at Ns = 1), so consecutive lanes fall into a
    // few banks; XORing the low SWB bits of the index with c ^ (c << 3),
    // c = index >> SWB, spreads them (the gather reads, 2^SWB consecutive
    // indices, stay a permutation of one bank row).  The map is linear over
    // GF(2), so the swizzle of a lane's base index is computed once and each
    // access XORs a compile-time constant (base and offset bits are disjoint).
    // Modelled on every one-system shape: write conflicts from ~0.85 extra
    // cycles per ideal cycle to ~0.1.  But each access then needs its own
    // address register instead of an immediate offset, and measured at 4096^2
    // with three fields that pays only where registers are to spare: FP64
    // transforms, one column (column pass 447 -> 418 us) or two interleaved
    // rows (row pass -3 %, 250 of 256 VGPRs, no spills).  FP32 lost (mixed
    // row pass 229 -> 309 us from spills, FP32 column pass down to one block
    // per CU), so the swizzle is on for FP64 with LS <= 2 only.
    static constexpr int LSB = (int)sizeof(cpx<T>) * LS;
    static constexpr bool SWIZZLE =
        PFC_LDS_SWIZZLE && POW2 && sizeof(T) == 8 && LS <= PFC_SWZ_LS && PADC == 0;
    static constexpr int SWB = (!SWIZZLE || LSB >= 256) ? 0 : ilog(2, 256 / LSB);
    __device__ static __forceinline__ int sw(int i) {
        if constexpr (SWB == 0) {
            return i;
        } else {
            const int c = i >> SWB;
            return i ^ ((c ^ (c << 3)) & ((1 << SWB) - 1));
        }
    }

    // LDS index padding (L = 2^k, instead of the swizzle): index i is stored
    // at i + PADC * (i >> PSH), PSH = log2(elements per 256-byte bank row).
    // The pad is additive over the same base/offset split as the swizzle (a
    // lane's base and an access's constant never carry across a multiple of
    // 2^PSH: the constant r Ns is below Ns R with Ns | 2^PSH, or a multiple of
    // it), so an access stays base + immediate offset and needs no extra
    // register.  FP32 PADC = 3 brings the modelled write conflicts of the step
    // transforms from 1.0 (rows) and 1.67 (columns) extra cycles per
    // conflict-free cycle to 0.33 and 0.17, for 9.4 % more exchange LDS.
    static constexpr int PSH = ilog(2, 256 / (int)sizeof(cpx<T>));
    static constexpr int PC = (POW2 && SWB == 0) ? PADC : 0;
    __device__ static __forceinline__ int pd(int i) { return PC ? i + PC * (i >> PSH) : i; }
    // elements one transform's exchange occupies (times LS when interleaved)
    static constexpr int lds_len = PC ? L + PC * ((L - 1) >> PSH) : L;

    // L = 3 * 2^k: stage s has radix gcd(P, L / (radices of stages < s)), so
    // the first stage takes the factor 3 (radix 12 or 6), the rest radix 4 or 2.
    static constexpr int gcd_(int a, int b) { return b == 0 ? a : gcd_(b, a % b); }
    static constexpr int rad(int s) {
        int rem = L;
        for (int i = 0; i < s; i++) rem /= gcd_(P, rem);
        return gcd_(P, rem);
    }
    static constexpr int ns(int s) {   // product of the radices of stages < s
        int n = 1;
        for (int i = 0; i < s; i++) n *= rad(i);
        return n;
    }
    static constexpr int nstages() {
        int rem = L, n = 0;
        while (rem > 1) {
            rem /= gcd_(P, rem);
            n++;
        }
        return n;
    }

    // Stage twiddle e^{-2 pi i m/L} (conjugated for the inverse): entry m*TWS
    // of a table of L*TWS entries, in global memory or copied to LDS.  Global and LDS reads are
    // counted by different wait counters, so LDS twiddles never wait for
    // the global loads and stores in flight before them.
    __device__ static __forceinline__ cpx<T> twid(GBuf tw, int m, unsigned z) {
        cpx<T> w = gload<T>(tw, (unsigned)(m * TWS) * sizeof(cpx<T>), z);
        if constexpr (DIR > 0) w.y = -w.y;
        return w;
    }
    __device__ static __forceinline__ cpx<T> twid(const cpx<T>* ltw, int m, unsigned) {
        cpx<T> w = ltw[m * TWS];
        if constexpr (DIR > 0) w.y = -w.y;
        return w;
    }

    // Compact stage table (L = 2^k): for each stage with Ns > 1, in execution
    // order, Ns entries w1[k] = e^{-2 pi i k step/L} (step = L/(Ns R)) and,
    // in FP32, Ns entries wh[k] = e^{-2 pi i k (R/2) step/L}.
    // The stages with Ns > 1 have Ns = F, F P, F P^2, ... up to L / P, F the
    // first (R0, or P without a leading stage), so the blocks before the one
    // of Ns sum to NW (Ns - F) / (P - 1) entries.
    static constexpr int NW = (sizeof(T) == 8 && PFC_TW_RECURRENCE) ? 1 : 2;
    static constexpr int F = R0 > 1 ? R0 : P;
    __device__ static __forceinline__ int soff(int Ns) { return NW * (Ns - F) / (P - 1); }
    static constexpr int kTab = POW2 ? NW * (L - F) / (P - 1) : 0;
    // twiddle exponent m (e^{-2 pi i m/L}) of compact entry i (the stage of
    // Ns has radix R0 if Ns = 1 and a leading stage exists, else P, so every
    // stage with Ns > 1 has radix P and step L / (Ns P))
    // (selects over the compile-time stages, no branches)
    static constexpr int NSTG = POW2 ? ilog(P, L / F) : 0;
    __device__ static __forceinline__ int tab_src(int i) {
        int m = 0;
#pragma unroll
        for (int s = 0, ns = F; s < NSTG; s++, ns *= P) {
            const int step = L / (ns * P);
            m = (i >= 0 && i < ns) ? i * step : m;
            i -= ns;
            if constexpr (NW == 2) {
                m = (i >= 0 && i < ns) ? i * (P / 2) * step : m;
                i -= ns;
            }
        }
        return m;
    }
    template <typename TW>
    __device__ static __forceinline__ cpx<T> tw_w1(TW tw, int k, int, int step, unsigned z) {
        return twid(tw, k * step, z);
    }
    template <typename TW>
    __device__ static __forceinline__ cpx<T> tw_wh(TW tw, int k, int, int step, int Hh, unsigned z) {
        return twid(tw, k * Hh * step, z);
    }
    // z: the transform's opaque zero, so that the compiler re-reads the table
    // per transform instead of keeping every stage's twiddles live
    __device__ static __forceinline__ cpx<T> tw_w1(StageTab<T> st, int k, int Ns, int, unsigned z) {
        cpx<T> w = st.p[soff(Ns) + k + (int)z];
        if constexpr (DIR > 0) w.y = -w.y;
        return w;
    }
    __device__ static __forceinline__ cpx<T> tw_wh(StageTab<T> st, int k, int Ns, int, int, unsigned z) {
        cpx<T> w = st.p[soff(Ns) + Ns + k + (int)z];
        if constexpr (DIR > 0) w.y = -w.y;
        return w;
    }

    // One radix-R Stockham stage on butterflies j = t + b*TPB, inputs in in[].
    template <int R, bool LAST, typename TW>
    __device__ static __forceinline__ void stage(cpx<T> (&in)[P], int Ns, cpx<T>* lds, TW tw,
                                                 int t, unsigned z) {
        constexpr int B = P / R;
#pragma unroll
        for (int b = 0; b < B; b++) {
            const int j = t + b * TPB;
            // Ns is a power of two for L = 2^k; a compile-time constant in run_mixed
            const int k = POW2 ? (j & (Ns - 1)) : (j % Ns);
            cpx<T> x[R];
#pragma unroll
            for (int r = 0; r < R; r++) x[r] = in[b + r * B];
            if (Ns > 1) {
                const int step = L / (Ns * R);
                if constexpr (sizeof(T) == 8 && PFC_TW_RECURRENCE) {
                    // FP64: powers w^r of one table twiddle by recurrence
                    // (error grows by ~r ulp, r < 16), far fewer loads and
                    // live registers than one load per r
                    const cpx<T> w1 = tw_w1(tw, k, Ns, step, z);
                    cpx<T> w = w1;
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        x[r] = cmul(x[r], w);
                        if (r + 1 < R) w = cmul(w, w1);
                    }
                } else if constexpr (R >= 4) {
                    // FP32: two table twiddles w and w^(R/2); every other
                    // power is at most two products away (~2 ulp)
                    constexpr int Hh = R / 2;
                    const cpx<T> w1 = tw_w1(tw, k, Ns, step, z);
                    const cpx<T> wh = tw_wh(tw, k, Ns, step, Hh, z);
                    cpx<T> lo[Hh];
                    lo[0] = {(T)1, (T)0};
                    lo[1] = w1;
#pragma unroll
                    for (int r = 2; r < Hh; r++) lo[r] = cmul(lo[r - 1], w1);
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        const cpx<T> w = (r < Hh) ? lo[r] : (r == Hh ? wh : cmul(wh, lo[r - Hh]));
                        x[r] = cmul(x[r], w);
                    }
                } else if constexpr (R == 2) {
                    x[1] = cmul(x[1], tw_w1(tw, k, Ns, step, z));
                } else {
#pragma unroll
                    for (int r = 1; r < R; r++) x[r] = cmul(x[r], twid(tw, k * r * step, z));
                }
            }
            dft<R, DIR, T>(x);
            if constexpr (LAST) {
#pragma unroll
                for (int r = 0; r < R; r++) in[b + r * B] = x[r];
            } else {
                // base = (j - k) R + k and r Ns occupy disjoint bits
                const int base = PC ? pd((j - k) * R + k) : sw((j - k) * R + k);
#pragma unroll
                for (int r = 0; r < R; r++)
                    lds[(SWB ? (base ^ sw(r * Ns)) : (base + pd(r * Ns))) * LS] = x[r];
            }
        }
    }

    // Gather the inputs of a radix-R stage from LDS.
    template <int R>
    __device__ static __forceinline__ void gather(cpx<T> (&in)[P], const cpx<T>* lds, int t) {
        constexpr int B = P / R;
#pragma unroll
        for (int b = 0; b < B; b++) {
            // j < L / R: j and r L / R occupy disjoint bits
            const int j = PC ? pd(t + b * TPB) : sw(t + b * TPB);
#pragma unroll
            for (int r = 0; r < R; r++)
                in[b + r * B] = lds[(SWB ? (j ^ sw(r * (L / R))) : (j + pd(r * (L / R)))) * LS];
        }
    }

    // v: natural-order input (thread t holds t + q*TPB); same layout on exit.
    // Begins with a barrier so lds may still be in use by the caller.
    //
    // The twiddle loads take an opaque zero offset, so that the compiler does
    // not share them (and keep them live) across the transforms of a kernel.
    template <int s, typename TW>
    __device__ static __forceinline__ void run_mixed(cpx<T> (&v)[P], cpx<T>* lds, TW tw, int t,
                                                     unsigned z) {
        constexpr int R = rad(s);
        constexpr int Ns = ns(s);
        if constexpr (s == nstages() - 1) {
            stage<R, true>(v, Ns, lds, tw, t, z);
        } else {
            if constexpr (s > 0) __syncthreads();
            stage<R, false>(v, Ns, lds, tw, t, z);
            __syncthreads();
            gather<rad(s + 1)>(v, lds, t);
            run_mixed<s + 1>(v, lds, tw, t, z);
        }
    }

    template <typename TW>
    __device__ static __forceinline__ void run(cpx<T> (&v)[P], cpx<T>* lds, TW tw, int t) {
        int Ns = 1;
        unsigned z = 0;
        asm volatile("" : "+s"(z));
        __syncthreads();
        if constexpr (!POW2) {
            run_mixed<0>(v, lds, tw, t, z);
        } else {
            if constexpr (R0 > 1) {
                // in-register leading stage: butterfly b of thread t uses v[b + r*(P/R0)]
                stage<R0, false>(v, 1, lds, tw, t, z);
                Ns = R0;
                __syncthreads();
                gather<P>(v, lds, t);
            }
#pragma unroll
            for (int s = 0; s < S; s++) {
                if (s == S - 1) {
                    stage<P, true>(v, Ns, lds, tw, t, z);
                } else {
                    if (s > 0 || R0 > 1) __syncthreads();
                    stage<P, false>(v, Ns, lds, tw, t, z);
                    Ns *= P;
                    __syncthreads();
                    gather<P>(v, lds, t);
                }
            }
        }
    }
};

}  // namespace pfc

// pfc_batch.hpp -- batched relaxation of many independent systems of uneven
// sizes (pfc_batch_* in include/pfc_relax.h).
//
// Systems are grouped into classes (FFT length, points per thread P, fields
// U).  Each pass of a step launches one kernel per chunk of systems of one
// class, with one block per work item (a row group, or a column block, of one
// of its systems), running the same row and column bodies and tiles as the
// one-system kernels.  (Persistent blocks walking the items with a grid
// stride ran the batched-relaxation workload 10 % slower.)  P sets the
// FFT radix (8, 16 or 4 for 2^k; 12 or 6 for 3 * 2^k) and with it the lanes
// per transform; the autotuner picks P per (length, U) and pass by timing
// both candidates.  (One kernel holding every class, dispatching per
// item, was measured first: register allocation over its 28 inlined bodies
// spilled 150-690 VGPRs where each body alone fits, and non-inlined bodies
// save ~150 callee-saved VGPRs to scratch per call.)
#pragma once
#include "pfc_launch.hpp"

namespace pfc {

// One system as the batch kernels see it (pointers into its context).
template <typename T>
struct BatchSys {
    cpx<T>* I;
    cpx<T>* J;
    cpx<T>* K;
    cpx<T>* Nh;
    size_t fstride, nstride;
    int H, sI, sJ, pad;
    const cpx<T>* twx;
    const cpx<T>* twy;
    const T* ktab;
    const DevParams* prm;
    int* nan;
    double* n0;
};

// a row group (row pass) or column block (column pass) of one system
struct BatchItem {
    int sys, blk;
};

// The item lists and descriptors are written before the first launch and
// only read by kernels, so the kernels read them through the constant
// address space: the compiler then keeps every loaded value in scalar
// registers.  (Read as ordinary global memory, a load that may follow a
// store of the body is not provably unclobbered, and the descriptor's
// pointers went through vector registers: k_batch_row<float, 2048, 8, 3>
// spilled 28-36 bytes per lane where the same body fed from kernel
// arguments needs 125 VGPRs and no scratch.)
template <typename X>
using KConst = const __attribute__((address_space(4))) X;
template <typename X>
__device__ __forceinline__ KConst<X>* kconst(const X* p) { return (KConst<X>*)p; }

// Kernel classes (id, length n, points per thread p, fields U, divisor d of
// the rows or columns per block).  Each (n, U) has three candidate shapes
// for the autotuner: two radices (three-field classes try fewer points per
// thread, their registers holding three fields) and the first of them with
// half the rows (columns) per block, which keeps the lanes' registers and
// halves the block (more, smaller blocks per CU, each synchronising fewer
// waves).
#define PFC_BATCH_CLASSES(X)                                                                       \
    X(0, 256, 8, 1, 1) X(1, 256, 16, 1, 1) X(2, 256, 8, 1, 2)                                      \
    X(3, 256, 8, 3, 1) X(4, 256, 4, 3, 1) X(5, 256, 8, 3, 2)                                       \
    X(6, 384, 12, 1, 1) X(7, 384, 6, 1, 1) X(8, 384, 12, 1, 2)                                     \
    X(9, 384, 6, 3, 1) X(10, 384, 12, 3, 1) X(11, 384, 6, 3, 2)                                    \
    X(12, 512, 8, 1, 1) X(13, 512, 16, 1, 1) X(14, 512, 8, 1, 2)                                   \
    X(15, 512, 8, 3, 1) X(16, 512, 4, 3, 1) X(17, 512, 8, 3, 2)                                    \
    X(18, 768, 12, 1, 1) X(19, 768, 6, 1, 1) X(20, 768, 12, 1, 2)                                  \
    X(21, 768, 6, 3, 1) X(22, 768, 12, 3, 1) X(23, 768, 6, 3, 2)                                   \
    X(24, 1024, 8, 1, 1) X(25, 1024, 16, 1, 1) X(26, 1024, 8, 1, 2)                                \
    X(27, 1024, 8, 3, 1) X(28, 1024, 4, 3, 1) X(29, 1024, 8, 3, 2)                                 \
    X(30, 1536, 12, 1, 1) X(31, 1536, 6, 1, 1) X(32, 1536, 12, 1, 2)                               \
    X(33, 1536, 6, 3, 1) X(34, 1536, 12, 3, 1) X(35, 1536, 6, 3, 2)                                \
    X(36, 2048, 8, 1, 1) X(37, 2048, 16, 1, 1) X(38, 2048, 8, 1, 2)                                \
    X(39, 2048, 8, 3, 1) X(40, 2048, 4, 3, 1) X(41, 2048, 8, 3, 2)
constexpr int kBatchClasses = 42;
constexpr int kBatchCandidates = 3;
struct BatchClass {
    int n, p, u, d;
};
#define PFC_CLASS_ENTRY(ID, N, PP, U, D) {N, PP, U, D},
constexpr BatchClass kBatchClass[kBatchClasses] = {PFC_BATCH_CLASSES(PFC_CLASS_ENTRY)};
#undef PFC_CLASS_ENTRY

// Rows per row item and columns per column item: the one-system kernels'
// tiles (row_cfg, col_cfg), trimmed to at most 1024 lanes per block when a
// shape has fewer points per thread than those kernels use, then divided by
// the class's d (at least 1).
template <typename T>
constexpr int batch_tr(int n, int p, int d) {
    const int tpb = n / 2 / p, tr = row_cfg<T>(n).tr;
    const int t = tr * tpb <= 1024 ? tr : 1024 / tpb;
    return t / d > 0 ? t / d : 1;
}
template <typename T>
constexpr int batch_tc(int n, int p, int d) {
    const int tpb = n / p, tc = col_cfg(n).tc;
    const int t = tc * tpb <= 1024 ? tc : 1024 / tpb;
    return t / d > 0 ? t / d : 1;
}
template <typename T>
constexpr int batch_tr(int c) { return batch_tr<T>(kBatchClass[c].n, kBatchClass[c].p, kBatchClass[c].d); }
template <typename T>
constexpr int batch_tc(int c) { return batch_tc<T>(kBatchClass[c].n, kBatchClass[c].p, kBatchClass[c].d); }

// Row pass of one class: block w runs item w.
template <typename T, int N, int PP, int U, int D, bool REF>
__global__ __launch_bounds__(N / 2 / PP * batch_tr<T>(N, PP, D)) PFC_WAVES((row_waves<T, N>())) void k_batch_row(
    const BatchItem* __restrict__ items, int nitems, const BatchSys<T>* __restrict__ sys,
    const int* __restrict__ tdev, int toff) {
    constexpr int TPB = N / 2 / PP, TR = batch_tr<T>(N, PP, D);
    __shared__ cpx<T> lds[row_lds_elems<T, N, TPB, TR>()];
    __shared__ __attribute__((aligned(16))) cpx<T> ltab[row_tab_size<T, N, TPB, TR>()];
    const int tstep = *tdev + toff;   // the step's index (NaN reports)
    if ((int)blockIdx.x >= nitems) return;
    const int isys = kconst(items)[blockIdx.x].sys, blk = kconst(items)[blockIdx.x].blk;
    const KConst<BatchSys<T>>& S = kconst(sys)[isys];
    row_step_body<T, N, TPB, TR, U, REF>(lds, ltab, blk, S.I, S.J, S.K, S.fstride, S.sI, S.sJ,
                                         S.twx, S.prm, S.nan, tstep);
}

// Column pass of one class: block w runs item w.
template <typename T, int N, int PP, int U, int D, bool REF>
__global__ __launch_bounds__(N / PP * batch_tc<T>(N, PP, D)) PFC_WAVES(col_waves<T>()) void k_batch_col(
    const BatchItem* __restrict__ items, int nitems, const BatchSys<T>* __restrict__ sys) {
    constexpr int TPB = N / PP, TC = batch_tc<T>(N, PP, D);
    constexpr bool LTAB = col_ltab<T, N, TPB, TC>();
    constexpr bool LGY = col_lgy<T, N, TPB, TC>();
    static_assert(LGY, "batch lengths keep their ky factors in LDS");
    __shared__ cpx<T> lds[col_lds_len<T, N, TPB>() * TC];
    __shared__ __attribute__((aligned(16))) cpx<T> ltw[col_tw_entries<T, N, TPB, TC, LTAB>()];
    __shared__ __attribute__((aligned(16))) T lgy[N];
    if ((int)blockIdx.x >= nitems) return;
    const int isys = kconst(items)[blockIdx.x].sys, blk = kconst(items)[blockIdx.x].blk;
    const KConst<BatchSys<T>>& S = kconst(sys)[isys];
    col_step_body<T, N, TPB, TC, U, REF, LTAB, LGY>(lds, ltw, lgy, blk, S.J, S.K, S.Nh, S.I,
                                                    S.fstride, S.nstride, S.H, S.sI, S.sJ, S.twy,
                                                    S.ktab, S.prm, S.n0);
}

// advances the device step counter read by the row pass by the steps just run
__global__ void k_batch_tick(int* tdev, int nsteps);

template <typename T>
struct BatchLaunch {
    // one block per item
    static void row(int cls, bool ref, const BatchItem* items, int nitems, const BatchSys<T>* sys,
                    const int* tdev, int toff, hipStream_t s);
    static void col(int cls, bool ref, const BatchItem* items, int nitems, const BatchSys<T>* sys,
                    hipStream_t s);
    // scratch (spill) bytes per lane of a class's row or column kernel for a
    // plain step, -1 if unknown
    static int row_scratch(int cls);
    static int col_scratch(int cls);
};

// Out-of-line definitions, compiled only where BatchLaunch's members are
// instantiated (pfc_batch_{f32,f64}_{row,col}.hip).
#ifdef PFC_BATCH_INSTANTIATE
template <typename T>
void BatchLaunch<T>::row(int cls, bool ref, const BatchItem* items, int nitems,
                         const BatchSys<T>* sys, const int* tdev, int toff, hipStream_t s) {
    switch (cls) {
#define PFC_LROW(ID, N, PP, U, D)                                                                  \
    case ID:                                                                                       \
        if (ref)                                                                                   \
            hipLaunchKernelGGL((k_batch_row<T, N, PP, U, D, true>), dim3(nitems),                  \
                               dim3(N / 2 / PP * batch_tr<T>(N, PP, D)), 0, s, items, nitems, sys, \
                               tdev, toff);                                                        \
        else                                                                                       \
            hipLaunchKernelGGL((k_batch_row<T, N, PP, U, D, false>), dim3(nitems),                 \
                               dim3(N / 2 / PP * batch_tr<T>(N, PP, D)), 0, s, items, nitems, sys, \
                               tdev, toff);                                                        \
        break;
        PFC_BATCH_CLASSES(PFC_LROW)
#undef PFC_LROW
    }
}
template <typename T>
void BatchLaunch<T>::col(int cls, bool ref, const BatchItem* items, int nitems,
                         const BatchSys<T>* sys, hipStream_t s) {
    switch (cls) {
#define PFC_LCOL(ID, N, PP, U, D)                                                                  \
    case ID:                                                                                       \
        if (ref)                                                                                   \
            hipLaunchKernelGGL((k_batch_col<T, N, PP, U, D, true>), dim3(nitems),                  \
                               dim3(N / PP * batch_tc<T>(N, PP, D)), 0, s, items, nitems, sys);    \
        else                                                                                       \
            hipLaunchKernelGGL((k_batch_col<T, N, PP, U, D, false>), dim3(nitems),                 \
                               dim3(N / PP * batch_tc<T>(N, PP, D)), 0, s, items, nitems, sys);    \
        break;
        PFC_BATCH_CLASSES(PFC_LCOL)
#undef PFC_LCOL
    }
}
// a kernel's scratch bytes per lane, -1 if unknown
inline int kernel_scratch(const void* f) {
    hipFuncAttributes a;
    if (!f || hipFuncGetAttributes(&a, f) != hipSuccess) return -1;
    return (int)a.localSizeBytes;
}
template <typename T>
int BatchLaunch<T>::row_scratch(int cls) {
    switch (cls) {
#define PFC_LSCR(ID, N, PP, U, D)                                                                  \
    case ID: return kernel_scratch((const void*)k_batch_row<T, N, PP, U, D, false>);
        PFC_BATCH_CLASSES(PFC_LSCR)
#undef PFC_LSCR
    }
    return -1;
}
template <typename T>
int BatchLaunch<T>::col_scratch(int cls) {
    switch (cls) {
#define PFC_LSCR(ID, N, PP, U, D)                                                                  \
    case ID: return kernel_scratch((const void*)k_batch_col<T, N, PP, U, D, false>);
        PFC_BATCH_CLASSES(PFC_LSCR)
#undef PFC_LSCR
    }
    return -1;
}
#endif

}  // namespace pfc

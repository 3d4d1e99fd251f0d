// pfc_launch.hpp -- context layout and kernel launchers, shared by the
// engine (pfc_engine.hip) and the per-precision instantiation units
// (pfc_launch_f64.hip, pfc_launch_f32.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pfc_relax.h"
#include "pfc_kernels.hpp"

namespace pfc {

constexpr int kEnergyBlocks = 1024;
constexpr int kEnergyBS = 256;

// Sizes compiled into the library (a narrower range builds faster).
#ifndef PFC_NMIN
#define PFC_NMIN 64
#endif
#ifndef PFC_NMAX
#define PFC_NMAX 8192
#endif
constexpr bool size_built(int n) { return n >= PFC_NMIN && n <= PFC_NMAX; }

#ifndef PFC_ROW_TR4096_F64
#define PFC_ROW_TR4096_F64 2
#endif
#ifndef PFC_ROW_TPB4096_F64
#define PFC_ROW_TPB4096_F64 256
#endif
#ifndef PFC_ROW_TR4096_F32
#define PFC_ROW_TR4096_F32 2
#endif
#ifndef PFC_ROW_TR2048_F32
#define PFC_ROW_TR2048_F32 4
#endif
#ifndef PFC_ROW_TR8192_F32
#define PFC_ROW_TR8192_F32 1
#endif

struct RowCfg {
    int tpb, tr;
};
// Row transforms have length H = Nx/2 and TR rows share a block of TPB*TR
// lanes.  TR is as large as the on-chip budget of a block allows (all U
// fields of TR rows stay in registers), so that TR*sizeof(E) bytes of each
// column-major buffer are read and written contiguously.
template <typename T>
constexpr RowCfg row_cfg(int nx) {
    constexpr bool f64 = sizeof(T) == 8;
    switch (nx) {
        case 64: return {8, 32};
        case 128: return {8, 32};
        case 256: return {16, 16};
        case 384: return {16, 16};   // 3 * 2^k: 12 points per thread
        case 512: return {32, 8};
        case 768: return {32, 8};
        case 1024: return {64, 4};
        case 1536: return {64, 4};
        case 2048: return {128, f64 ? 4 : PFC_ROW_TR2048_F32};
        case 4096: return {f64 ? PFC_ROW_TPB4096_F64 : 256, f64 ? PFC_ROW_TR4096_F64 : PFC_ROW_TR4096_F32};
        case 8192: return {512, f64 ? 1 : PFC_ROW_TR8192_F32};
        default: return {0, 0};
    }
}
struct ColCfg {
    int tpb, tc;
};
constexpr ColCfg col_cfg(int ny) {
    switch (ny) {
        case 64: return {8, 32};
        case 128: return {16, 16};
        case 256: return {32, 8};
        case 384: return {32, 8};    // 3 * 2^k: 12 points per thread
        case 512: return {64, 4};
        case 768: return {64, 4};
        case 1024: return {128, 2};
        case 1536: return {128, 2};
        case 2048: return {256, 1};
        case 4096: return {512, 1};
        case 8192: return {1024, 1};
        default: return {0, 0};
    }
}

// The step's column pass: col_cfg, except columns of 8192, which take
// PFC_COL_TPB8192 lanes (1024 lanes of 8 points are capped at 128 VGPRs and
// spill; 512 lanes of 16 points have 256).
#ifndef PFC_COL_TPB8192
#define PFC_COL_TPB8192 512
#endif
constexpr ColCfg col_step_cfg(int ny) { return ny == 8192 ? ColCfg{PFC_COL_TPB8192, 1} : col_cfg(ny); }

inline bool size_ok(int n) { return row_cfg<double>(n).tpb != 0 && size_built(n); }

inline int ilog2(int n) {
    int s = 0;
    while ((1 << s) < n) s++;
    return s;
}

// y-block height of a row-spectral buffer: Ny (column-major), or if blocked
// the rows of one row-pass block (one contiguous run per kx for the row
// pass), at most the column pass's lanes per column (its addressing needs
// TPB >= YB).
#ifndef PFC_I_BLOCKED
#define PFC_I_BLOCKED 0
#endif
#ifndef PFC_J_BLOCKED
#define PFC_J_BLOCKED 0
#endif
template <typename T>
inline int row_block(int nx, int ny, bool blocked) {
    if (!blocked) return ny;
    const int tr = row_cfg<T>(nx).tr, tpb = col_cfg(ny).tpb;
    const int b = tr < tpb ? tr : tpb;
    return b < ny ? b : ny;
}

}  // namespace pfc

// Device context behind the opaque pfc_ctx of the C ABI.
struct pfc_ctx {
    int Nx, Ny, U, prec, nO;
    int sI, sJ, sC;   // log2 of the y-block heights of I, of J and K, and of O
    size_t nk;    // row-spectral buffer elements per field, (Nx/2) * Ny
    size_t nkn;   // spectral-state elements per field, (Nx/2 + 1) * Ny
    size_t nr;    // real-space points, Nx * Ny
    size_t esz;  // bytes per complex element
    pfc_params p;
    double dx, dy;
    pfc::DevParams hp;
    pfc::DevParams* dprm;
    void *I, *J, *K, *Nh, *O, *mreal, *R, *twx, *twy;
    double* n0;   // mixed form: the k = 0 state of each field in FP64 (re, im)
    void* ktab;   // Gaussian factors exp(-k^2/sigma^2) per kx and per ky (ktab_gy), set by pfc_set_box
    void* hktab;  // host staging of ktab
    // partials: the energy sums' per-block partial pairs, local ones (1 + U
    // per block) from 0 and spectral ones (1 per block) from poff; sums:
    // their totals; hsums: pinned host copy for pfc_step_evaluate
    double *f, *partials, *sums, *hsums;
    size_t poff;
    hipEvent_t ev_sums;
    int* nan_flag;
    hipStream_t stream;
    int ncu;       // compute units of the device
    int persist;   // FP64 row pass as a persistent launch (PFC_PERSIST=0 turns it off)
    // column pass of non-refresh steps per field (k_col_step_f), which writes
    // the new state to the other of two state buffers: Nh2, allocated only
    // for this (PFC_COLSPLIT=0 turns it off); the mixed form's FP64 k = 0
    // state alternates likewise between n0 and n0 + 2U
    int colsplit;
    void* Nh2;
    // Nh was recomputed from the current densities (I) by an evaluate or
    // pfc_set_fields and no step has run since: a t % 100 step then skips
    // its refresh, which would recompute the same state (tuned path only)
    int nh_fresh;
    int generic;   // shape or U outside the tuned kernels: the generic path (pfc_generic.hip)
    void* gen;     // its plans and tables
    int have_fields, have_box;
    int f_stale;   // the last evaluate was pfc_evaluate_sums: f is older than it
    // pass timing (pfc_timing): 3 events per recorded step
    int timing;
    int nrec;
    int nev;
    hipEvent_t* ev;
};

namespace pfc {

// The generic path (pfc_generic.hip): any Nx, Ny whose transforms fit one
// block's LDS and any U <= PFC_MAX_FIELDS.  The density lives in mreal and
// the half spectra in the transposed (Nx/2 + 1) x Ny layout (nk = nkn).
struct GenPlan;
bool gen_plan(int L, size_t esz, GenPlan* p);     // false: L has no generic plan
bool gen_plannable(int L, size_t esz);
int gen_create(pfc_ctx* c);                       // plans and device tables
void gen_destroy(pfc_ctx* c);
void gen_refresh(pfc_ctx* c);                     // Nh = FFT2(m) / (Nx Ny)
void gen_step(pfc_ctx* c, int t);                 // one step t
void gen_evaluate(pfc_ctx* c);                    // up to the energy partial sums
void gen_energy_sums(pfc_ctx* c);                 // k_energy_local + k_energy_spec
void gen_smooth(pfc_ctx* c, const void* in, void* out, double sigma);

// Launchers for one element precision T.  Only declared here, so that a unit
// that launches (the engine) compiles no kernel; pfc_launch_impl.hpp defines
// them and pfc_launch_f64.hip / pfc_launch_f32.hip instantiate them.
template <typename T>
struct Launch {
    // a step's row and column passes; epart (refresh steps only): also the
    // free energy's partial sums of the incoming state (returns the blocks)
    static int row_step_any(pfc_ctx* c, int t, bool ref, double* epart = nullptr);
    static int col_step_any(pfc_ctx* c, bool ref, double* epart = nullptr);
    // per-field column pass of a non-refresh step: state read from src,
    // written to dst (both U * nkn elements); false if not built for Ny
    static bool col_step_split_any(pfc_ctx* c, const void* src, void* dst);
    static bool col_split_built(int Ny);
    static void row_c2r_any(pfc_ctx* c, const void* src, int sS, void* dst, void* K, int nbuf);
    static void row_r2c_any(pfc_ctx* c, const void* src, void* dst, int sD, int nbuf);
    static void col_fwd_any(pfc_ctx* c, const void* src, int sS, void* dst, int nbuf, double scale);
    static void col_smooth_any(pfc_ctx* c, const void* src, void* dst, double sigma);
    static void col_eval_any(pfc_ctx* c);
    static void energy_any(pfc_ctx* c);
    static void energy_sums_any(pfc_ctx* c);   // after the refresh: Nh and mreal current
    // the densities' row pass of pfc_evaluate_sums: K and the local energy
    // sums, without the densities (returns the partial-sum blocks)
    static int eval_local_any(pfc_ctx* c);
    static void energy_spec_any(pfc_ctx* c);   // the spectral sums from Nh
};

}  // namespace pfc

// pfc_launch_impl.hpp -- definitions of the launchers declared in
// pfc_launch.hpp: one switch over the compiled sizes per kernel family.
// Included only by the per-precision instantiation units.
#pragma once
#include "pfc_launch.hpp"

namespace pfc {

// Launchers for one element precision T; explicitly instantiated for double
// and float in their own translation units.
// Resident blocks of a kernel per CU (1 if the runtime cannot tell).
inline int occupancy(const void* kernel, int threads) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) n = 1;
    return n;
}
// Grid of a persistent launch over n work items: every resident block slot,
// a multiple of 8 (one share per XCD) where that is at least 8, at most n.
inline int persist_grid(const pfc_ctx* c, int occ, int n) {
    int nb = c->ncu * occ;
    if (nb > n) nb = n;
    if (nb >= 8) nb -= nb % 8;
    return nb > 0 ? nb : 1;
}

template <typename T>
struct LaunchImpl {
    using C = cpx<T>;

    // The step's passes; with epart (a refresh step only) the EVAL variants,
    // which also write the free energy's partial sums of the incoming state
    // there.  Return the blocks launched (the partial sums written).
    template <int NX, int U, bool REF, bool EV>
    static int row_step(pfc_ctx* c, int t, double* epart) {
        if constexpr (size_built(NX)) {
            constexpr RowCfg g = row_cfg<T>(NX);
            const int ng = c->Ny / g.tr;
            if constexpr (sizeof(T) == 8) {
                if (c->persist) {
                    auto k = k_row_step_p<T, NX, g.tpb, g.tr, U, REF, EV>;
                    static const int occ = occupancy((const void*)k, g.tpb * g.tr);
                    const int nb = persist_grid(c, occ, ng);
                    hipLaunchKernelGGL(k, dim3(nb), dim3(g.tpb * g.tr), 0, c->stream, (const C*)c->I,
                                       (C*)c->J, (C*)c->K, c->nk, c->sI, c->sJ, (const C*)c->twx,
                                       c->dprm, c->nan_flag, t, ng, epart);
                    return nb;
                }
            }
            hipLaunchKernelGGL((k_row_step<T, NX, g.tpb, g.tr, U, REF, EV>), dim3(ng),
                               dim3(g.tpb * g.tr), 0, c->stream, (const C*)c->I, (C*)c->J, (C*)c->K,
                               c->nk, c->sI, c->sJ, (const C*)c->twx, c->dprm, c->nan_flag, t, epart);
            return ng;
        }
        return 0;
    }
    template <int NX, int U>
    static int row_step_u(pfc_ctx* c, int t, bool ref, double* epart) {
        // (the FP64 form only: pfc_step_evaluate takes it for FP64 alone)
        if constexpr (sizeof(T) == 8) {
            if (epart) return row_step<NX, U, true, true>(c, t, epart);
        }
        if (ref) return row_step<NX, U, true, false>(c, t, nullptr);
        return row_step<NX, U, false, false>(c, t, nullptr);
    }
    template <int NX>
    static int row_step_n(pfc_ctx* c, int t, bool ref, double* epart) {
        switch (c->U) {
            case 1: return row_step_u<NX, 1>(c, t, ref, epart);
            case 2: return row_step_u<NX, 2>(c, t, ref, epart);
            case 3: return row_step_u<NX, 3>(c, t, ref, epart);
            case 4: return row_step_u<NX, 4>(c, t, ref, epart);
        }
        return 0;
    }
    static int row_step_any(pfc_ctx* c, int t, bool ref, double* epart) {
        switch (c->Nx) {
            case 64: return row_step_n<64>(c, t, ref, epart);
            case 128: return row_step_n<128>(c, t, ref, epart);
            case 256: return row_step_n<256>(c, t, ref, epart);
            case 384: return row_step_n<384>(c, t, ref, epart);
            case 512: return row_step_n<512>(c, t, ref, epart);
            case 768: return row_step_n<768>(c, t, ref, epart);
            case 1024: return row_step_n<1024>(c, t, ref, epart);
            case 1536: return row_step_n<1536>(c, t, ref, epart);
            case 2048: return row_step_n<2048>(c, t, ref, epart);
            case 4096: return row_step_n<4096>(c, t, ref, epart);
            case 8192: return row_step_n<8192>(c, t, ref, epart);
        }
        return 0;
    }

    template <int NY, int U, bool REF, bool EV>
    static int col_step(pfc_ctx* c, double* epart) {
        if constexpr (size_built(NY)) {
            constexpr ColCfg g = col_step_cfg(NY);
            const int H = c->Nx / 2;
            hipLaunchKernelGGL((k_col_step<T, NY, g.tpb, g.tc, U, REF, EV>), dim3(H / g.tc + 1),
                               dim3(g.tpb * g.tc), 0, c->stream, (const C*)c->J, (const C*)c->K,
                               (C*)c->Nh, (C*)c->I, c->nk, c->nkn, H, c->sI, c->sJ, (const C*)c->twy,
                               (const T*)c->ktab, c->dprm, c->n0, epart);
            return H / g.tc + 1;
        }
        return 0;
    }
    template <int NY, int U>
    static int col_step_u(pfc_ctx* c, bool ref, double* epart) {
        if constexpr (sizeof(T) == 8) {
            if (epart) return col_step<NY, U, true, true>(c, epart);
        }
        if (ref) return col_step<NY, U, true, false>(c, nullptr);
        return col_step<NY, U, false, false>(c, nullptr);
    }
    template <int NY>
    static int col_step_n(pfc_ctx* c, bool ref, double* epart) {
        switch (c->U) {
            case 1: return col_step_u<NY, 1>(c, ref, epart);
            case 2: return col_step_u<NY, 2>(c, ref, epart);
            case 3: return col_step_u<NY, 3>(c, ref, epart);
            case 4: return col_step_u<NY, 4>(c, ref, epart);
        }
        return 0;
    }
    static int col_step_any(pfc_ctx* c, bool ref, double* epart) {
        switch (c->Ny) {
            case 64: return col_step_n<64>(c, ref, epart);
            case 128: return col_step_n<128>(c, ref, epart);
            case 256: return col_step_n<256>(c, ref, epart);
            case 384: return col_step_n<384>(c, ref, epart);
            case 512: return col_step_n<512>(c, ref, epart);
            case 768: return col_step_n<768>(c, ref, epart);
            case 1024: return col_step_n<1024>(c, ref, epart);
            case 1536: return col_step_n<1536>(c, ref, epart);
            case 2048: return col_step_n<2048>(c, ref, epart);
            case 4096: return col_step_n<4096>(c, ref, epart);
            case 8192: return col_step_n<8192>(c, ref, epart);
        }
        return 0;
    }

    // per-field column pass (2^k columns); the mixed form's FP64 k = 0
    // state alternates between n0 and n0 + 2U with the spectral state
    static constexpr bool split_built(int ny) {
        return size_built(ny) && (ny & (ny - 1)) == 0;
    }
    template <int NY, int U>
    static void col_split(pfc_ctx* c, const void* src, void* dst) {
        if constexpr (split_built(NY)) {
            constexpr ColCfg g = col_step_cfg(NY);
            const int H = c->Nx / 2;
            const bool back = src == c->Nh2;
            hipLaunchKernelGGL((k_col_step_f<T, NY, g.tpb, g.tc, U>), dim3((H / g.tc + 1) * U),
                               dim3(g.tpb * g.tc), 0, c->stream, (const C*)c->J, (const C*)src,
                               (C*)dst, (C*)c->I, c->nk, c->nkn, H, c->sI, c->sJ,
                               (const C*)c->twy, (const T*)c->ktab, c->dprm,
                               (const double*)(back ? c->n0 + 2 * U : c->n0),
                               back ? c->n0 : c->n0 + 2 * U);
        }
    }
    template <int NY>
    static void col_split_n(pfc_ctx* c, const void* src, void* dst) {
        switch (c->U) {
            case 1: col_split<NY, 1>(c, src, dst); break;
            case 2: col_split<NY, 2>(c, src, dst); break;
            case 3: col_split<NY, 3>(c, src, dst); break;
            case 4: col_split<NY, 4>(c, src, dst); break;
        }
    }
    static bool col_step_split_any(pfc_ctx* c, const void* src, void* dst) {
        if (!split_built(c->Ny)) return false;
        switch (c->Ny) {
            case 64: col_split_n<64>(c, src, dst); break;
            case 128: col_split_n<128>(c, src, dst); break;
            case 256: col_split_n<256>(c, src, dst); break;
            case 512: col_split_n<512>(c, src, dst); break;
            case 1024: col_split_n<1024>(c, src, dst); break;
            case 2048: col_split_n<2048>(c, src, dst); break;
            case 4096: col_split_n<4096>(c, src, dst); break;
            case 8192: col_split_n<8192>(c, src, dst); break;
            default: return false;
        }
        return true;
    }

    template <int NX>
    static void row_c2r(pfc_ctx* c, const void* src, int sS, void* dst, void* K, int nbuf) {
        if constexpr (size_built(NX)) {
            constexpr RowCfg g = row_cfg<T>(NX);
            hipLaunchKernelGGL((k_row_c2r<T, NX, g.tpb, g.tr>), dim3(c->Ny / g.tr, nbuf),
                               dim3(g.tpb * g.tr), 0, c->stream, (const C*)src, (T*)dst, (C*)K, c->nk,
                               c->nr, sS, c->sJ, (const C*)c->twx);
        }
    }
    static void row_c2r_any(pfc_ctx* c, const void* src, int sS, void* dst, void* K, int nbuf) {
        switch (c->Nx) {
            case 64: row_c2r<64>(c, src, sS, dst, K, nbuf); break;
            case 128: row_c2r<128>(c, src, sS, dst, K, nbuf); break;
            case 256: row_c2r<256>(c, src, sS, dst, K, nbuf); break;
            case 384: row_c2r<384>(c, src, sS, dst, K, nbuf); break;
            case 512: row_c2r<512>(c, src, sS, dst, K, nbuf); break;
            case 768: row_c2r<768>(c, src, sS, dst, K, nbuf); break;
            case 1024: row_c2r<1024>(c, src, sS, dst, K, nbuf); break;
            case 1536: row_c2r<1536>(c, src, sS, dst, K, nbuf); break;
            case 2048: row_c2r<2048>(c, src, sS, dst, K, nbuf); break;
            case 4096: row_c2r<4096>(c, src, sS, dst, K, nbuf); break;
            case 8192: row_c2r<8192>(c, src, sS, dst, K, nbuf); break;
        }
    }
    template <int NX>
    static void row_r2c(pfc_ctx* c, const void* src, void* dst, int sD, int nbuf) {
        if constexpr (size_built(NX)) {
            constexpr RowCfg g = row_cfg<T>(NX);
            hipLaunchKernelGGL((k_row_r2c<T, NX, g.tpb, g.tr>), dim3(c->Ny / g.tr, nbuf),
                               dim3(g.tpb * g.tr), 0, c->stream, (const T*)src, (C*)dst, c->nk, c->nr,
                               sD, (const C*)c->twx, (T)(1.0 / NX));
        }
    }
    static void row_r2c_any(pfc_ctx* c, const void* src, void* dst, int sD, int nbuf) {
        switch (c->Nx) {
            case 64: row_r2c<64>(c, src, dst, sD, nbuf); break;
            case 128: row_r2c<128>(c, src, dst, sD, nbuf); break;
            case 256: row_r2c<256>(c, src, dst, sD, nbuf); break;
            case 384: row_r2c<384>(c, src, dst, sD, nbuf); break;
            case 512: row_r2c<512>(c, src, dst, sD, nbuf); break;
            case 768: row_r2c<768>(c, src, dst, sD, nbuf); break;
            case 1024: row_r2c<1024>(c, src, dst, sD, nbuf); break;
            case 1536: row_r2c<1536>(c, src, dst, sD, nbuf); break;
            case 2048: row_r2c<2048>(c, src, dst, sD, nbuf); break;
            case 4096: row_r2c<4096>(c, src, dst, sD, nbuf); break;
            case 8192: row_r2c<8192>(c, src, dst, sD, nbuf); break;
        }
    }
    template <int NY>
    static void col_fwd(pfc_ctx* c, const void* src, int sS, void* dst, int nbuf, double scale) {
        if constexpr (size_built(NY)) {
            constexpr ColCfg g = col_cfg(NY);
            const int H = c->Nx / 2;
            hipLaunchKernelGGL((k_col_fwd<T, NY, g.tpb, g.tc>), dim3(H / g.tc + 1, nbuf),
                               dim3(g.tpb * g.tc), 0, c->stream, (const C*)src, (C*)dst, c->nk,
                               c->nkn, H, sS, (const C*)c->twy, scale);
        }
    }
    static void col_fwd_any(pfc_ctx* c, const void* src, int sS, void* dst, int nbuf, double scale) {
        switch (c->Ny) {
            case 64: col_fwd<64>(c, src, sS, dst, nbuf, scale); break;
            case 128: col_fwd<128>(c, src, sS, dst, nbuf, scale); break;
            case 256: col_fwd<256>(c, src, sS, dst, nbuf, scale); break;
            case 384: col_fwd<384>(c, src, sS, dst, nbuf, scale); break;
            case 512: col_fwd<512>(c, src, sS, dst, nbuf, scale); break;
            case 768: col_fwd<768>(c, src, sS, dst, nbuf, scale); break;
            case 1024: col_fwd<1024>(c, src, sS, dst, nbuf, scale); break;
            case 1536: col_fwd<1536>(c, src, sS, dst, nbuf, scale); break;
            case 2048: col_fwd<2048>(c, src, sS, dst, nbuf, scale); break;
            case 4096: col_fwd<4096>(c, src, sS, dst, nbuf, scale); break;
            case 8192: col_fwd<8192>(c, src, sS, dst, nbuf, scale); break;
        }
    }
    template <int NY>
    static void col_smooth(pfc_ctx* c, const void* src, void* dst, double sigma) {
        if constexpr (size_built(NY)) {
            constexpr ColCfg g = col_cfg(NY);
            const int H = c->Nx / 2;
            hipLaunchKernelGGL((k_col_smooth<T, NY, g.tpb, g.tc>), dim3(H / g.tc + 1),
                               dim3(g.tpb * g.tc), 0, c->stream, (const C*)src, (C*)dst, H, c->sC,
                               (const C*)c->twy, c->hp.dkx, c->hp.dky, 0.5 / sigma / sigma);
        }
    }
    static void col_smooth_any(pfc_ctx* c, const void* src, void* dst, double sigma) {
        switch (c->Ny) {
            case 64: col_smooth<64>(c, src, dst, sigma); break;
            case 128: col_smooth<128>(c, src, dst, sigma); break;
            case 256: col_smooth<256>(c, src, dst, sigma); break;
            case 384: col_smooth<384>(c, src, dst, sigma); break;
            case 512: col_smooth<512>(c, src, dst, sigma); break;
            case 768: col_smooth<768>(c, src, dst, sigma); break;
            case 1024: col_smooth<1024>(c, src, dst, sigma); break;
            case 1536: col_smooth<1536>(c, src, dst, sigma); break;
            case 2048: col_smooth<2048>(c, src, dst, sigma); break;
            case 4096: col_smooth<4096>(c, src, dst, sigma); break;
            case 8192: col_smooth<8192>(c, src, dst, sigma); break;
        }
    }
    template <int NY>
    static void col_eval(pfc_ctx* c) {
        if constexpr (size_built(NY)) {
            constexpr ColCfg g = col_cfg(NY);
            const int H = c->Nx / 2;
            hipLaunchKernelGGL((k_col_eval<T, NY, g.tpb, g.tc>), dim3(H / g.tc + 1, c->U),
                               dim3(g.tpb * g.tc), 0, c->stream, (const C*)c->K, (C*)c->Nh, (C*)c->O,
                               c->nk, c->nkn, H, c->sJ, (const C*)c->twy, c->dprm);
        }
    }
    static void col_eval_any(pfc_ctx* c) {
        switch (c->Ny) {
            case 64: col_eval<64>(c); break;
            case 128: col_eval<128>(c); break;
            case 256: col_eval<256>(c); break;
            case 384: col_eval<384>(c); break;
            case 512: col_eval<512>(c); break;
            case 768: col_eval<768>(c); break;
            case 1024: col_eval<1024>(c); break;
            case 1536: col_eval<1536>(c); break;
            case 2048: col_eval<2048>(c); break;
            case 4096: col_eval<4096>(c); break;
            case 8192: col_eval<8192>(c); break;
        }
    }
    template <int U>
    static void energy(pfc_ctx* c) {
        hipLaunchKernelGGL((k_energy<T, U, kEnergyBS>), dim3(kEnergyBlocks), dim3(kEnergyBS), 0,
                           c->stream, (const T*)c->mreal, (const T*)c->R, c->f, c->nr, c->partials,
                           c->dprm);
    }
    static void energy_any(pfc_ctx* c) {
        switch (c->U) {
            case 1: energy<1>(c); break;
            case 2: energy<2>(c); break;
            case 3: energy<3>(c); break;
            case 4: energy<4>(c); break;
        }
    }
    // evaluate without the energy density: the densities' row pass with the
    // local energy sums (k_row_c2r_eloc, K for the refresh); returns its
    // number of partial-sum blocks
    template <int NX, int U>
    static int eval_local(pfc_ctx* c) {
        if constexpr (size_built(NX)) {
            constexpr RowCfg g = row_cfg<T>(NX);
            const int ng = c->Ny / g.tr;
            int nb = ng < kEnergyBlocks ? ng : kEnergyBlocks;
            if (nb >= 8) nb -= nb % 8;
            hipLaunchKernelGGL((k_row_c2r_eloc<T, NX, g.tpb, g.tr, U>), dim3(nb), dim3(g.tpb * g.tr),
                               0, c->stream, (const C*)c->I, (C*)c->K, c->nk, c->sI, c->sJ,
                               (const C*)c->twx, c->dprm, c->partials, ng);
            return nb;
        }
        return 0;
    }
    template <int NX>
    static int eval_local_n(pfc_ctx* c) {
        switch (c->U) {
            case 1: return eval_local<NX, 1>(c);
            case 2: return eval_local<NX, 2>(c);
            case 3: return eval_local<NX, 3>(c);
            case 4: return eval_local<NX, 4>(c);
        }
        return 0;
    }
    static int eval_local_any(pfc_ctx* c) {
        switch (c->Nx) {
            case 64: return eval_local_n<64>(c);
            case 128: return eval_local_n<128>(c);
            case 256: return eval_local_n<256>(c);
            case 384: return eval_local_n<384>(c);
            case 512: return eval_local_n<512>(c);
            case 768: return eval_local_n<768>(c);
            case 1024: return eval_local_n<1024>(c);
            case 1536: return eval_local_n<1536>(c);
            case 2048: return eval_local_n<2048>(c);
            case 4096: return eval_local_n<4096>(c);
            case 8192: return eval_local_n<8192>(c);
        }
        return 0;
    }
    template <int U>
    static void energy_spec(pfc_ctx* c) {
        hipLaunchKernelGGL((k_energy_spec<T, U, 256>), dim3(kEnergyBlocks), dim3(256), 0, c->stream,
                           (const C*)c->Nh, c->nkn, c->Nx, c->Ny, c->partials + c->poff, c->dprm);
    }
    static void energy_spec_any(pfc_ctx* c) {
        switch (c->U) {
            case 1: energy_spec<1>(c); break;
            case 2: energy_spec<2>(c); break;
            case 3: energy_spec<3>(c); break;
            case 4: energy_spec<4>(c); break;
        }
    }
    template <int U>
    static void energy_sums(pfc_ctx* c) {
        launch_energy_sums<T, U>((const T*)c->mreal, (const C*)c->Nh, c->nr, c->nkn, c->Nx, c->Ny,
                                 c->partials, c->partials + c->poff, kEnergyBlocks, c->dprm,
                                 c->stream);
    }
    static void energy_sums_any(pfc_ctx* c) {
        switch (c->U) {
            case 1: energy_sums<1>(c); break;
            case 2: energy_sums<2>(c); break;
            case 3: energy_sums<3>(c); break;
            case 4: energy_sums<4>(c); break;
        }
    }
};

template <typename T>
int Launch<T>::row_step_any(pfc_ctx* c, int t, bool ref, double* epart) {
    return LaunchImpl<T>::row_step_any(c, t, ref, epart);
}
template <typename T>
int Launch<T>::col_step_any(pfc_ctx* c, bool ref, double* epart) {
    return LaunchImpl<T>::col_step_any(c, ref, epart);
}
template <typename T>
bool Launch<T>::col_step_split_any(pfc_ctx* c, const void* src, void* dst) {
    return LaunchImpl<T>::col_step_split_any(c, src, dst);
}
template <typename T>
bool Launch<T>::col_split_built(int Ny) { return LaunchImpl<T>::split_built(Ny); }
template <typename T>
void Launch<T>::row_c2r_any(pfc_ctx* c, const void* src, int sS, void* dst, void* K, int nbuf) {
    LaunchImpl<T>::row_c2r_any(c, src, sS, dst, K, nbuf);
}
template <typename T>
void Launch<T>::row_r2c_any(pfc_ctx* c, const void* src, void* dst, int sD, int nbuf) {
    LaunchImpl<T>::row_r2c_any(c, src, dst, sD, nbuf);
}
template <typename T>
void Launch<T>::col_fwd_any(pfc_ctx* c, const void* src, int sS, void* dst, int nbuf, double scale) {
    LaunchImpl<T>::col_fwd_any(c, src, sS, dst, nbuf, scale);
}
template <typename T>
void Launch<T>::col_smooth_any(pfc_ctx* c, const void* src, void* dst, double sigma) {
    LaunchImpl<T>::col_smooth_any(c, src, dst, sigma);
}
template <typename T>
void Launch<T>::col_eval_any(pfc_ctx* c) { LaunchImpl<T>::col_eval_any(c); }
template <typename T>
void Launch<T>::energy_any(pfc_ctx* c) { LaunchImpl<T>::energy_any(c); }
template <typename T>
void Launch<T>::energy_sums_any(pfc_ctx* c) { LaunchImpl<T>::energy_sums_any(c); }
template <typename T>
int Launch<T>::eval_local_any(pfc_ctx* c) { return LaunchImpl<T>::eval_local_any(c); }
template <typename T>
void Launch<T>::energy_spec_any(pfc_ctx* c) { LaunchImpl<T>::energy_spec_any(c); }

}  // namespace pfc

// lbm_launch_impl.hpp -- definitions of the Launch entry points: each picks the
// kernel instantiation for the run's method, collision, LES and DDF toggles and
// the addressing width. Included only by the per-pair instantiation units.
#pragma once

#include <algorithm>

#include "lbm_launch.hpp"

namespace lbm {

template <class L, class S> struct LaunchImpl {
    using R = typename S::R;
    using T = typename S::T;

    template <int MODE, int COLL, int LES, bool DDF, bool LINKS>
    static int step_l(const Handle& h, const StepParams<R>& p, hipStream_t st) {
        const int s = h.g.tzb - h.g.wzb;  // 2^s workgroups per tile, one z slab each
        const dim3 grid((unsigned)h.g.n_tiles << s), block((unsigned)h.g.tile_cells >> s);
        if constexpr (S::id == LBM_STORE_F16C && L::D == 3 && MODE == MODE_DB && !LINKS) {
            if (h.staged_pull && !h.big && h.g.txb == 3 && h.g.tyb == 3 && h.g.wzb == 2) {
                if (h.staged_pull == 2 && h.g.Tx % 2 == 0) {
                    const dim3 grid2((unsigned)(h.g.n_tiles / 2) << s), block2(512);
                    const size_t lds2 = StagedBox<L, 2>::bytes();
                    hipLaunchKernelGGL((step_kernel_staged<L, COLL, LES, DDF, 2>), grid2, block2, lds2, st, p);
                } else {
                    const size_t lds1 = StagedBox<L, 1>::bytes();
                    hipLaunchKernelGGL((step_kernel_staged<L, COLL, LES, DDF, 1>), grid, block, lds1, st, p);
                }
                return check_launch("step_kernel_staged");
            }
        }
        if (h.big)
            hipLaunchKernelGGL((step_kernel<L, S, MODE, COLL, LES, DDF, true, LINKS>), grid, block, 0, st, p);
        else
            hipLaunchKernelGGL((step_kernel<L, S, MODE, COLL, LES, DDF, false, LINKS>), grid, block, 0, st, p);
        return check_launch("step_kernel");
    }

    template <int MODE, int COLL, int LES, bool DDF>
    static int step_k(const Handle& h, const StepParams<R>& p, hipStream_t st) {
        if (MODE == MODE_DB && p.link_nodes) return step_l<MODE, COLL, LES, DDF, true>(h, p, st);
        return step_l<MODE, COLL, LES, DDF, false>(h, p, st);
    }

    // The learned closure: 3-D lattices, 32-bit addressing, no interpolated links.
    template <int MODE, int COLL, bool DDF>
    static int step_mlp(const Handle& h, const StepParams<R>& p, hipStream_t st) {
        if constexpr (L::D == 3) {
            if (h.big || p.link_nodes)
                return set_error(LBM_ERR_UNSUPPORTED,
                                 "the learned closure runs without interpolated links and with buffers under 2 GiB");
            if (!p.mlp) return set_error(LBM_ERR_INVALID, "the learned closure needs lbm_set_mlp first");
            if (h.g.tile_cells % 64)
                return set_error(LBM_ERR_UNSUPPORTED, "the learned closure needs tiles of whole 64-lane waves");
            const int s = h.g.tzb - h.g.wzb;  // 2^s workgroups per tile, one z slab each
            hipLaunchKernelGGL((step_kernel_mlp<L, S, MODE, COLL, DDF>), dim3((unsigned)h.g.n_tiles << s),
                               dim3((unsigned)h.g.tile_cells >> s), 0, st, p);
            return check_launch("step_kernel");
        } else {
            return set_error(LBM_ERR_UNSUPPORTED, "the learned closure needs a 3-D lattice");
        }
    }

    template <int MODE, int COLL>
    static int step_coll(const Handle& h, const StepParams<R>& p, hipStream_t st) {
        const bool ddf = h.d.ddf_shift != 0;
        if (h.d.les == LBM_LES_MLP) {
            if constexpr (COLL == C_SRT || COLL == C_KBC)
                return ddf ? step_mlp<MODE, COLL, true>(h, p, st) : step_mlp<MODE, COLL, false>(h, p, st);
            return set_error(LBM_ERR_UNSUPPORTED, "the learned closure is provided with SRT and KBC collision");
        }
        if (h.d.les == LBM_LES_SMAGORINSKY)
            return ddf ? step_k<MODE, COLL, LES_SMAG, true>(h, p, st) : step_k<MODE, COLL, LES_SMAG, false>(h, p, st);
        return ddf ? step_k<MODE, COLL, LES_OFF, true>(h, p, st) : step_k<MODE, COLL, LES_OFF, false>(h, p, st);
    }

    template <int MODE>
    static int step_mode(const Handle& h, const StepParams<R>& p, hipStream_t st) {
        switch (h.d.collision) {
            case LBM_COLL_TRT: return step_coll<MODE, C_TRT>(h, p, st);
            case LBM_COLL_KBC: return step_coll<MODE, C_KBC>(h, p, st);
            case LBM_COLL_RLBM: return step_coll<MODE, C_RLBM>(h, p, st);
            default: return step_coll<MODE, C_SRT>(h, p, st);
        }
    }

    // The per-tile flag summary of (flags, bc), the tile kinds and the neighbour masks of the
    // boundary tiles: checked before a step unless static flags are promised and they were built
    // from these buffers; rebuilt when the flags or the moving-wall bytes differ from the ones
    // they were built from (the kind and mask kernels return at once otherwise).
    static int ensure_summary(const Handle& h, const uint8_t* flags, const void* bc, hipStream_t st) {
        const uint32_t n_chunks = (uint32_t)(((size_t)h.g.n_tiles * h.g.tile_cells) >> 4);
        if (!h.static_flags || h.summary_flags != (const void*)flags || h.summary_bc != bc) {
            uint8_t* mvs = h.d.include_moving_boundary ? h.mvs : nullptr;
            hipLaunchKernelGGL((tile_summary_kernel<L::D, R>), dim3((n_chunks + 255) / 256), dim3(256), 0, st, flags,
                               (const R*)bc, mvs, h.summary, h.btiles, h.flags_seen, h.mask_changed, h.g, n_chunks);
            if (int rc = check_launch("tile_summary_kernel")) return rc;
            hipLaunchKernelGGL((tile_kind_kernel<L::D>), dim3((h.g.n_tiles + 255) / 256), dim3(256), 0, st, h.summary,
                               h.tile_kind, h.btiles, (const uint32_t*)h.mask_changed, h.g);
            if (int rc = check_launch("tile_kind_kernel")) return rc;
            hipLaunchKernelGGL((node_mask_kernel<L>), dim3(std::min<uint32_t>(kMaskGrid, (uint32_t)h.g.n_tiles)),
                               dim3(h.g.tile_cells), 2 * h.lds_bytes, st, flags, (const uint8_t*)mvs, h.btiles,
                               (const uint32_t*)h.mask_changed, h.nbmask, h.g, (uint32_t)h.lds_bytes);
            if (int rc = check_launch("node_mask_kernel")) return rc;
            hipLaunchKernelGGL(mask_done_kernel<0>, dim3(1), dim3(64), 0, st, h.mask_changed, h.btiles,
                               (uint32_t)h.g.n_tiles);
            if (int rc = check_launch("mask_done_kernel")) return rc;
            h.summary_flags = flags;
            h.summary_bc = bc;
        }
        return LBM_OK;
    }

    static int step(const Handle& h, void* f_out, const void* f_in, const void* bc, const uint8_t* flags, double tau,
                    int mode, hipStream_t st) {
        if ((uintptr_t)flags & 15u) return set_error(LBM_ERR_INVALID, "flags must be 16-byte aligned");
        if (int rc = ensure_summary(h, flags, bc, st)) return rc;
        StepParams<R> p{};
        p.f_in = f_in;
        p.f_out = f_out;
        p.bc = (const R*)bc;
        p.flags = flags;
        p.tile_kind = h.tile_kind;
        p.nbmask = h.nbmask;
        p.tau = (R)tau;
        p.inv_tau = (R)1 / p.tau;
        p.cs = (R)h.d.smagorinsky_cs;
        p.moving = h.d.include_moving_boundary != 0;
        p.eqbc = (h.d.included_bcs >> LBM_FLAG_EQUILIBRIUM) & 1u;
        p.capture = (int)(h.d.included_bcs & (LBM_FLAG_CAPTURE_DENSITY | LBM_FLAG_CAPTURE_VELOCITY));
        p.f_bytes = (uint32_t)(h.big ? 0 : h.f_bytes);
        p.link_nodes = h.n_links ? h.link_nodes : nullptr;
        p.link_q = h.link_q;
        p.n_links = h.n_links;
        p.force_partial = h.force_partial;
        p.node_force = h.n_links ? h.node_force : nullptr;
        p.mlp = h.mlp;
        p.g = h.g;
        if (mode == MODE_DB) return step_mode<MODE_DB>(h, p, st);
        if (mode == MODE_EP_EVEN) return step_mode<MODE_EP_EVEN>(h, p, st);
        return step_mode<MODE_EP_ODD>(h, p, st);
    }

    template <int MODE, bool DDF, int KIND>
    static int macro_k(const Handle& h, MacroParams<R> p, hipStream_t st) {
        // LDS: flag tile + halo, then (Q-criterion) the velocity tile + halo, then the output staging
        p.su_offset = (uint32_t)((h.lds_bytes + 15) & ~(size_t)15);
        const size_t su_bytes = KIND == MK_QCRIT ? h.lds_bytes * L::D * sizeof(R) : 0;
        p.stage_offset = (uint32_t)((p.su_offset + su_bytes + 15) & ~(size_t)15);
        const size_t stage_bytes = KIND == MK_FIELDS ? (size_t)h.g.tile_cells * (L::D + 1) * sizeof(R)
                                   : KIND == MK_QCRIT ? (size_t)h.g.tile_cells * sizeof(R) : 0;
        const size_t lds = p.stage_offset + stage_bytes;
        if (lds > 65536) return set_error(LBM_ERR_UNSUPPORTED, "tile too large for the output kernel's LDS (64 KiB)");
        if constexpr (KIND == MK_FIELDS && L::D == 3 && sizeof(R) == 4) {
            // whole-line output columns (four z-adjacent tiles) on 8 x 8 x 8 tiles with Tz a multiple of 4,
            // 4-byte moments and 16-byte aligned outputs (FP64 moments take the tile kernel)
            const bool aligned = (((uintptr_t)p.rho | (uintptr_t)p.u) & 15u) == 0;
            if (aligned && p.tile_kind && p.nbmask && h.g.txb == 3 && h.g.tyb == 3 &&
                h.g.tzb == 3 && h.g.Tz % 4 == 0) {
                const dim3 grid((unsigned)(h.g.n_tiles / 4));
                if (h.big)
                    hipLaunchKernelGGL((macro_fields_col<L, S, MODE, DDF, true>), grid, dim3(512), 0, st, p);
                else
                    hipLaunchKernelGGL((macro_fields_col<L, S, MODE, DDF, false>), grid, dim3(512), 0, st, p);
                return check_launch("macro_fields_col");
            }
        }
        if constexpr (KIND == MK_QCRIT && L::D == 3 && sizeof(R) == 4) {
            // the same columns for the Q-criterion (16-byte aligned output)
            if (!((uintptr_t)p.qout & 15u) && p.tile_kind && p.nbmask && h.g.txb == 3 && h.g.tyb == 3 &&
                h.g.tzb == 3 && h.g.Tz % 4 == 0) {
                const dim3 grid((unsigned)(h.g.n_tiles / 4));
                const size_t qlds = (size_t)(kQColBoxCells * L::D + kQColStage) * sizeof(R);
                if (h.big)
                    hipLaunchKernelGGL((macro_qcrit_col<L, S, MODE, DDF, true>), grid, dim3(512), qlds, st, p);
                else
                    hipLaunchKernelGGL((macro_qcrit_col<L, S, MODE, DDF, false>), grid, dim3(512), qlds, st, p);
                return check_launch("macro_qcrit_col");
            }
        }
        if (h.big)
            hipLaunchKernelGGL((macro_kernel<L, S, MODE, DDF, KIND, true>), dim3(h.g.n_tiles), dim3(h.g.tile_cells),
                               lds, st, p);
        else
            hipLaunchKernelGGL((macro_kernel<L, S, MODE, DDF, KIND, false>), dim3(h.g.n_tiles),
                               dim3(h.g.tile_cells), lds, st, p);
        return check_launch("macro_kernel");
    }

    template <int KIND>
    static int macro_any(const Handle& h, const MacroParams<R>& p, int mode, hipStream_t st) {
        const bool ddf = h.d.ddf_shift != 0;
        if (mode == MODE_DB) return ddf ? macro_k<MODE_DB, true, KIND>(h, p, st) : macro_k<MODE_DB, false, KIND>(h, p, st);
        if (mode == MODE_EP_EVEN)
            return ddf ? macro_k<MODE_EP_EVEN, true, KIND>(h, p, st) : macro_k<MODE_EP_EVEN, false, KIND>(h, p, st);
        return ddf ? macro_k<MODE_EP_ODD, true, KIND>(h, p, st) : macro_k<MODE_EP_ODD, false, KIND>(h, p, st);
    }

    static MacroParams<R> macro_params(const Handle& h, const void* f, const void* bc, const uint8_t* flags,
                                       bool mode_is_db, hipStream_t st, int* rc) {
        MacroParams<R> p{};
        *rc = LBM_OK;
        // interior tiles skip the flags when the summary can be built (16-byte aligned flags)
        if (!((uintptr_t)flags & 15u)) {
            *rc = ensure_summary(h, flags, bc, st);
            p.tile_kind = h.tile_kind;
            p.nbmask = h.nbmask;
        }
        p.f = f;
        p.bc = (const R*)bc;
        p.flags = flags;
        p.moving = h.d.include_moving_boundary != 0;
        p.eqbc = (h.d.included_bcs >> LBM_FLAG_EQUILIBRIUM) & 1u;
        p.f_bytes = (uint32_t)(h.big ? 0 : h.f_bytes);
        p.g = h.g;
        if (mode_is_db && h.n_links) {
            p.link_nodes = h.link_nodes;
            p.link_q = h.link_q;
            p.n_links = h.n_links;
        }
        return p;
    }

    static int macro(const Handle& h, void* rho, void* u, const void* f, const void* bc, const uint8_t* flags,
                     int mode, int u_layout, hipStream_t st) {
        int prc = LBM_OK;
        MacroParams<R> p = macro_params(h, f, bc, flags, mode == MODE_DB, st, &prc);
        if (prc) return prc;
        p.rho = (R*)rho;
        p.u = (R*)u;
        p.u_layout = u_layout;
        return macro_any<MK_FIELDS>(h, p, mode, st);
    }

    static int qcrit(const Handle& h, void* q, const void* u, const void* f, const void* bc, const uint8_t* flags,
                     double tau, int mode, int u_layout, hipStream_t st) {
        int prc = LBM_OK;
        MacroParams<R> p = macro_params(h, f, bc, flags, mode == MODE_DB, st, &prc);
        if (prc) return prc;
        p.vel = (const R*)u;
        p.qout = (R*)q;
        p.tau = (R)tau;
        p.u_layout = u_layout;
        return macro_any<MK_QCRIT>(h, p, mode, st);
    }

    static int reduce(const Handle& h, double* out, const void* f, const void* bc, const uint8_t* flags, int mode,
                      hipStream_t st) {
        int prc = LBM_OK;
        MacroParams<R> p = macro_params(h, f, bc, flags, mode == MODE_DB, st, &prc);
        if (prc) return prc;
        p.partial = h.partial;
        int rc = macro_any<MK_REDUCE>(h, p, mode, st);
        if (rc) return rc;
        if (L::D == 2) hipLaunchKernelGGL((reduce_finish<3>), dim3(1), dim3(256), 0, st, h.partial, h.g.n_tiles, out);
        else hipLaunchKernelGGL((reduce_finish<4>), dim3(1), dim3(256), 0, st, h.partial, h.g.n_tiles, out);
        return check_launch("reduce_finish");
    }

    static int link_force(const Handle& h, double* out, hipStream_t st) {
        const int waves = (h.g.tile_cells + 63) / 64;
        const int n = h.g.n_tiles * waves;
        if (L::D == 2) hipLaunchKernelGGL((reduce_finish<2>), dim3(1), dim3(256), 0, st, h.force_partial, n, out);
        else hipLaunchKernelGGL((reduce_finish<3>), dim3(1), dim3(256), 0, st, h.force_partial, n, out);
        return check_launch("reduce_finish");
    }

    static int init(const Handle& h, void* f, double rho_d, const double* u_d, hipStream_t st) {
        QVals<R> v{};
        R u[3] = {0, 0, 0};
        if (u_d)
            for (int d = 0; d < L::D; ++d) u[d] = (R)u_d[d];
        const R rho = (R)rho_d;
        const R uu = dot_uu<L, R>(u);
        const bool ddf = h.d.ddf_shift != 0;
        for (int q = 0; q < L::Q; ++q) v.v[q] = ddf ? feq<L, R, true>(q, rho, u, uu) : feq<L, R, false>(q, rho, u, uu);
        const size_t n = (size_t)h.g.n_tiles * h.g.tile_cells * L::Q;
        const unsigned blocks = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL((fill_kernel<S>), dim3(blocks), dim3(256), 0, st, (T*)f, v, L::Q, (uint32_t)h.g.tile_cells, n);
        return check_launch("fill_kernel");
    }
};

template <class L, class S>
int Launch<L, S>::step(const Handle& h, void* f_out, const void* f_in, const void* bc, const uint8_t* flags,
                       double tau, int mode, hipStream_t st) {
    return LaunchImpl<L, S>::step(h, f_out, f_in, bc, flags, tau, mode, st);
}
template <class L, class S>
int Launch<L, S>::macro(const Handle& h, void* rho, void* u, const void* f, const void* bc, const uint8_t* flags,
                        int mode, int u_layout, hipStream_t st) {
    return LaunchImpl<L, S>::macro(h, rho, u, f, bc, flags, mode, u_layout, st);
}
template <class L, class S>
int Launch<L, S>::reduce(const Handle& h, double* out, const void* f, const void* bc, const uint8_t* flags, int mode,
                         hipStream_t st) {
    return LaunchImpl<L, S>::reduce(h, out, f, bc, flags, mode, st);
}
template <class L, class S>
int Launch<L, S>::qcrit(const Handle& h, void* q, const void* u, const void* f, const void* bc, const uint8_t* flags,
                        double tau, int mode, int u_layout, hipStream_t st) {
    return LaunchImpl<L, S>::qcrit(h, q, u, f, bc, flags, tau, mode, u_layout, st);
}
template <class L, class S> int Launch<L, S>::link_force(const Handle& h, double* out, hipStream_t st) {
    return LaunchImpl<L, S>::link_force(h, out, st);
}
template <class L, class S>
int Launch<L, S>::init(const Handle& h, void* f, double rho_d, const double* u_d, hipStream_t st) {
    return LaunchImpl<L, S>::init(h, f, rho_d, u_d, st);
}

}  // namespace lbm

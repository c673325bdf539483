// lbm_batch_impl.hpp -- the batched step kernel and the definitions of the
// BatchLaunch entry points. Included only by the instantiation units.
#pragma once

#include <algorithm>

#include "lbm_batch.hpp"

namespace lbm {

template <class R> struct BatchCaseRec {
    StepParams<R> p;      // f_in = the case's slice of buffer 0, f_out = of buffer 1
    uint32_t tile_begin;  // first batch tile of the case
};

template <class L, class S, int COLL, int LES, bool DDF, bool LINKS>
__global__ void __launch_bounds__(MaxThreads<L::Q>::value)
    batch_step_kernel(const BatchCaseRec<typename S::R>* __restrict__ cases, const uint32_t* __restrict__ tile_case,
                      uint32_t n_tiles, uint32_t chunk, int parity) {
    using R = typename S::R;
    const uint32_t n_slots = 8u * chunk;
    for (uint32_t b = blockIdx.x; b < n_slots; b += gridDim.x) {
        // slot b runs on XCD b % 8 (round-robin dispatch); XCD k walks tiles [k chunk, (k+1) chunk)
        const uint32_t gt = (b & 7u) * chunk + (b >> 3);
        if (gt < n_tiles) {
            const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)tile_case[gt]);
            StepParams<R> p = cases[c].p;
            if (parity) {
                const void* a = p.f_in;
                p.f_in = p.f_out;
                p.f_out = const_cast<void*>(a);
            }
            step_tile<L, S, MODE_DB, COLL, LES, DDF, false, LINKS>(p, gt - cases[c].tile_begin, gt, 0, p.g.tzb);
        }
    }
}

template <class L, class S> struct BatchLaunchImpl {
    using R = typename S::R;
    using T = typename S::T;

    template <int COLL, int LES, bool DDF>
    static int step_k(const Batch& b, int parity, hipStream_t st) {
        const uint32_t slots = 8u * b.chunk;
        const uint32_t grid = b.grid_blocks > 0 ? std::min<uint32_t>((uint32_t)b.grid_blocks, slots) : slots;
        if (b.any_links)
            hipLaunchKernelGGL((batch_step_kernel<L, S, COLL, LES, DDF, true>), dim3(grid), dim3(b.tile_cells), 0, st,
                               (const BatchCaseRec<R>*)b.recs, b.tile_case, b.n_tiles, b.chunk, parity);
        else
            hipLaunchKernelGGL((batch_step_kernel<L, S, COLL, LES, DDF, false>), dim3(grid), dim3(b.tile_cells), 0, st,
                               (const BatchCaseRec<R>*)b.recs, b.tile_case, b.n_tiles, b.chunk, parity);
        return check_launch("batch_step_kernel");
    }

    template <int COLL>
    static int step_c(const Batch& b, int parity, hipStream_t st) {
        const bool les = b.proto.les == LBM_LES_SMAGORINSKY;
        const bool ddf = b.proto.ddf_shift != 0;
        if (les) return ddf ? step_k<COLL, LES_SMAG, true>(b, parity, st) : step_k<COLL, LES_SMAG, false>(b, parity, st);
        return ddf ? step_k<COLL, LES_OFF, true>(b, parity, st) : step_k<COLL, LES_OFF, false>(b, parity, st);
    }

    static int step(const Batch& b, int parity, hipStream_t st) {
        switch (b.proto.collision) {
            case LBM_COLL_TRT: return step_c<C_TRT>(b, parity, st);
            case LBM_COLL_KBC: return step_c<C_KBC>(b, parity, st);
            case LBM_COLL_RLBM: return step_c<C_RLBM>(b, parity, st);
            default: return step_c<C_SRT>(b, parity, st);
        }
    }

    // Per-case flag summaries and the device case records.
    static int prepare(Batch& b, hipStream_t st) {
        const int D = L::D;
        std::vector<BatchCaseRec<R>> recs(b.cases.size());
        for (size_t c = 0; c < b.cases.size(); ++c) {
            const Geo& g = b.geo[c];
            const uint8_t* fl = b.flags + b.fl_off[c];
            const R* bcp = (const R*)b.bc + b.bc_off[c];
            uint32_t* sm = b.summary + (size_t)b.tile_begin[c];
            const uint32_t n_chunks = (uint32_t)(((size_t)g.n_tiles * g.tile_cells) >> 4);
            uint8_t* mv = b.proto.include_moving_boundary ? b.mvs + b.fl_off[c] : nullptr;
            hipLaunchKernelGGL((tile_summary_kernel<L::D, R>), dim3((n_chunks + 255) / 256), dim3(256), 0, st, fl, bcp,
                               mv, sm, b.btiles, (uint8_t*)nullptr, (uint32_t*)nullptr, g, n_chunks);
            if (int rc = check_launch("tile_summary_kernel")) return rc;
            uint32_t* tk = b.tile_kind + b.tile_begin[c];
            uint32_t* nm = b.nbmask + (size_t)b.tile_begin[c] * b.tile_cells;
            hipLaunchKernelGGL((tile_kind_kernel<L::D>), dim3((g.n_tiles + 255) / 256), dim3(256), 0, st, sm, tk,
                               b.btiles, (const uint32_t*)nullptr, g);
            if (int rc = check_launch("tile_kind_kernel")) return rc;
            hipLaunchKernelGGL((node_mask_kernel<L>), dim3(std::min<uint32_t>(kMaskGrid, (uint32_t)g.n_tiles)),
                               dim3(g.tile_cells), 2 * b.lds_bytes, st, fl, (const uint8_t*)mv, b.btiles,
                               (const uint32_t*)nullptr, nm, g, (uint32_t)b.lds_bytes);
            if (int rc = check_launch("node_mask_kernel")) return rc;
            StepParams<R> p{};
            p.f_in = (const T*)b.f[0] + b.f_off[c];
            p.f_out = (T*)b.f[1] + b.f_off[c];
            p.bc = bcp;
            p.flags = fl;
            p.tile_kind = tk;
            p.nbmask = nm;
            p.tau = (R)b.cases[c].tau;
            p.inv_tau = (R)1 / p.tau;
            p.cs = (R)b.proto.smagorinsky_cs;
            p.moving = b.proto.include_moving_boundary != 0;
            p.eqbc = (b.proto.included_bcs >> LBM_FLAG_EQUILIBRIUM) & 1u;
            p.f_bytes = (uint32_t)((size_t)g.n_tiles * g.tile_cells * L::Q * sizeof(T));
            const size_t nl = b.h_links[c].size();
            size_t off = 0;
            for (size_t k = 0; k < c; ++k) off += b.h_links[k].size();
            p.link_nodes = nl ? b.link_nodes + off : nullptr;
            p.link_q = nl ? b.link_q + off * (L::Q - 1) : nullptr;
            p.n_links = (uint32_t)nl;
            p.force_partial = b.force_partial;
            p.g = g;
            recs[c].p = p;
            recs[c].tile_begin = b.tile_begin[c];
            (void)D;
        }
        return hip_rc(hipMemcpyAsync(b.recs, recs.data(), recs.size() * sizeof(BatchCaseRec<R>),
                                     hipMemcpyHostToDevice, st),
                      "copy of the case records");
    }

    static size_t rec_bytes() { return sizeof(BatchCaseRec<R>); }

    static int init(const Batch& b, hipStream_t st) {
        QVals<R> v{};
        R u[3] = {0, 0, 0};
        const R uu = dot_uu<L, R>(u);
        const bool ddf = b.proto.ddf_shift != 0;
        for (int q = 0; q < L::Q; ++q)
            v.v[q] = ddf ? feq<L, R, true>(q, (R)1, u, uu) : feq<L, R, false>(q, (R)1, u, uu);
        const unsigned blocks = (unsigned)((b.f_elems + 255) / 256);
        for (int k = 0; k < 2; ++k) {
            hipLaunchKernelGGL((fill_kernel<S>), dim3(blocks), dim3(256), 0, st, (T*)b.f[k], v, L::Q,
                               (uint32_t)b.tile_cells, b.f_elems);
            if (int rc = check_launch("fill_kernel")) return rc;
        }
        return LBM_OK;
    }

    static int force(const Batch& b, double* out, hipStream_t st) {
        const int waves = (int)((b.tile_cells + 63) / 64);
        for (size_t c = 0; c < b.cases.size(); ++c) {
            const double* part = b.force_partial + (size_t)b.tile_begin[c] * waves * L::D;
            const int n = (int)(b.case_tiles[c] * (uint32_t)waves);
            if (L::D == 2) hipLaunchKernelGGL((reduce_finish<2>), dim3(1), dim3(256), 0, st, part, n, out + c * 2);
            else hipLaunchKernelGGL((reduce_finish<3>), dim3(1), dim3(256), 0, st, part, n, out + c * 3);
            if (int rc = check_launch("reduce_finish")) return rc;
        }
        return LBM_OK;
    }
};

template <class L, class S> int BatchLaunch<L, S>::step(const Batch& b, int parity, hipStream_t st) {
    return BatchLaunchImpl<L, S>::step(b, parity, st);
}
template <class L, class S> int BatchLaunch<L, S>::prepare(Batch& b, hipStream_t st) {
    return BatchLaunchImpl<L, S>::prepare(b, st);
}
template <class L, class S> size_t BatchLaunch<L, S>::rec_bytes() { return BatchLaunchImpl<L, S>::rec_bytes(); }
template <class L, class S> int BatchLaunch<L, S>::init(const Batch& b, hipStream_t st) {
    return BatchLaunchImpl<L, S>::init(b, st);
}
template <class L, class S> int BatchLaunch<L, S>::force(const Batch& b, double* out, hipStream_t st) {
    return BatchLaunchImpl<L, S>::force(b, out, st);
}

}  // namespace lbm

// lbm_batch.hip -- extern "C" batch API (include/lbm_hip.h, "Batched cases"):
// layout and staging of the cases, the step loop with its captured hipGraph,
// per-case forces and the tile-shape autotuner.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "lbm_batch.hpp"

namespace lbm {

template <class F> static int batch_dispatch(const lbm_desc& d, F&& f) {
#define LBM_CASE(QQ, SS, SID) \
    if (d.Q == QQ && d.storage == SID) return f(BatchLaunch<Lattice<QQ>, SS>{});
    LBM_CASE(9, StoreF32, LBM_STORE_F32)
    LBM_CASE(9, StoreF16C, LBM_STORE_F16C)
    LBM_CASE(9, StoreF64, LBM_STORE_F64)
    LBM_CASE(19, StoreF32, LBM_STORE_F32)
    LBM_CASE(19, StoreF16C, LBM_STORE_F16C)
    LBM_CASE(19, StoreF64, LBM_STORE_F64)
    LBM_CASE(27, StoreF32, LBM_STORE_F32)
    LBM_CASE(27, StoreF16C, LBM_STORE_F16C)
    LBM_CASE(27, StoreF64, LBM_STORE_F64)
#undef LBM_CASE
    return set_error(LBM_ERR_INVALID, "no kernel set for this lattice / storage");
}

static Batch* B(lbm_batch* b) { return reinterpret_cast<Batch*>(b); }

static int pad_to(int n, int t) { return (n + t - 1) / t * t; }

static size_t tiled_index(const Geo& g, int x, int y, int z, int K, int k) {
    const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
    const size_t loc = (size_t)((x & (tx - 1)) + ((y & (ty - 1)) << g.txb) + ((z & (tz - 1)) << (g.txb + g.tyb)));
    const size_t tl = (size_t)(x >> g.txb) + (size_t)g.Tx * ((size_t)(y >> g.tyb) + (size_t)g.Ty * (size_t)(z >> g.tzb));
    return (tl * (size_t)K + (size_t)k) * (size_t)g.tile_cells + loc;
}

static void free_graph(Batch& b) {
    if (b.graph) (void)hipGraphExecDestroy(b.graph);
    if (b.graph_def) (void)hipGraphDestroy(b.graph_def);
    b.graph = nullptr;
    b.graph_def = nullptr;
}

static void free_device(Batch& b) {
    free_graph(b);
    void* ptrs[] = {b.f[0], b.f[1], b.bc, b.flags, b.summary, b.tile_kind, b.nbmask, b.mvs, b.btiles, b.link_nodes, b.link_q,
                    b.force_partial, b.recs, b.tile_case};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b.f[0] = b.f[1] = b.bc = nullptr;
    b.flags = nullptr;
    b.summary = nullptr;
    b.tile_kind = nullptr;
    b.nbmask = nullptr;
    b.mvs = nullptr;
    b.btiles = nullptr;
    b.link_nodes = nullptr;
    b.link_q = nullptr;
    b.link_cap = 0;
    b.force_partial = nullptr;
    b.recs = nullptr;
    b.tile_case = nullptr;
    b.laid_out = b.staged = b.initialized = false;
}

// Padded per-case descriptions for a tile shape (validated), without touching the batch.
static int case_descs(const Batch& b, const int tile[3], std::vector<lbm_desc>& out) {
    out.assign(b.cases.size(), b.proto);
    for (size_t c = 0; c < b.cases.size(); ++c) {
        lbm_desc& d = out[c];
        d.tx = tile[0];
        d.ty = tile[1];
        d.tz = tile[2];
        d.nx = pad_to(b.cases[c].n[0], tile[0]);
        d.ny = pad_to(b.cases[c].n[1], tile[1]);
        d.nz = pad_to(b.cases[c].n[2], tile[2]);
        if (int rc = validate(&d)) return rc;
    }
    return LBM_OK;
}

// Allocates every device buffer for the current tile shape.
static int layout(Batch& b) {
    free_device(b);
    std::vector<lbm_desc> cd;
    if (int rc = case_descs(b, b.tile, cd)) return rc;
    const int D = b.proto.D, Q = b.proto.Q;
    b.cdesc = cd;
    b.geo.resize(cd.size());
    b.f_off.resize(cd.size());
    b.bc_off.resize(cd.size());
    b.fl_off.resize(cd.size());
    b.tile_begin.resize(cd.size());
    b.case_tiles.resize(cd.size());
    size_t f = 0, bcn = 0, fl = 0;
    uint32_t tiles = 0;
    b.real_nodes = 0;
    std::vector<uint32_t> tile_case;
    for (size_t c = 0; c < cd.size(); ++c) {
        Geo g = make_geo(&cd[c]);
        g.xcd_chunk = 0;  // the batch kernel orders the tiles itself
        b.geo[c] = g;
        const size_t cells = (size_t)g.n_tiles * g.tile_cells;
        b.f_off[c] = f;
        b.bc_off[c] = bcn;
        b.fl_off[c] = fl;
        b.tile_begin[c] = tiles;
        b.case_tiles[c] = (uint32_t)g.n_tiles;
        f += cells * Q;
        bcn += cells * (D + 1);
        fl += cells;
        tiles += (uint32_t)g.n_tiles;
        tile_case.insert(tile_case.end(), (size_t)g.n_tiles, (uint32_t)c);
        b.real_nodes += (uint64_t)b.cases[c].n[0] * b.cases[c].n[1] * b.cases[c].n[2];
        if ((double)cells * Q * (b.proto.storage == LBM_STORE_F64 ? 8 : (b.proto.storage == LBM_STORE_F16C ? 2 : 4)) >=
            2147483648.0)
            return set_error(LBM_ERR_UNSUPPORTED, "a batch case needs under 2 GiB of populations per buffer");
    }
    b.tile_cells = (uint32_t)(b.tile[0] * b.tile[1] * b.tile[2]);
    b.n_tiles = tiles;
    b.chunk = (tiles + 7) / 8;
    b.lds_bytes = (size_t)(b.tile[0] + 2) * (b.tile[1] + 2) * (D == 3 ? b.tile[2] + 2 : 1);
    b.f_elems = f;
    b.bc_elems = bcn;
    b.fl_bytes = fl;
    b.f_elem_bytes = b.proto.storage == LBM_STORE_F16C ? 2 : (b.proto.storage == LBM_STORE_F64 ? 8 : 4);
    b.r_bytes = b.proto.storage == LBM_STORE_F64 ? 8 : 4;
    const int waves = (int)((b.tile_cells + 63) / 64);
    size_t rec = 0;
    batch_dispatch(b.proto, [&](auto l) {
        rec = decltype(l)::rec_bytes();
        return 0;
    });
    bool ok = hipMalloc(&b.f[0], f * b.f_elem_bytes) == hipSuccess && hipMalloc(&b.f[1], f * b.f_elem_bytes) == hipSuccess &&
              hipMalloc(&b.bc, bcn * b.r_bytes) == hipSuccess && hipMalloc((void**)&b.flags, fl) == hipSuccess &&
              hipMalloc((void**)&b.summary, (size_t)tiles * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void**)&b.tile_kind, (size_t)tiles * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void**)&b.nbmask, (size_t)tiles * b.tile_cells * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void**)&b.mvs, fl) == hipSuccess &&
              hipMalloc((void**)&b.btiles, ((size_t)tiles + 1) * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void**)&b.force_partial, (size_t)tiles * waves * D * sizeof(double)) == hipSuccess &&
              hipMalloc(&b.recs, rec * cd.size()) == hipSuccess &&
              hipMalloc((void**)&b.tile_case, (size_t)tiles * sizeof(uint32_t)) == hipSuccess;
    if (!ok) {
        free_device(b);
        return set_error(LBM_ERR_HIP, "hipMalloc of the batch buffers failed");
    }
    if (hipMemcpy(b.tile_case, tile_case.data(), tile_case.size() * sizeof(uint32_t), hipMemcpyHostToDevice) !=
        hipSuccess) {
        free_device(b);
        return set_error(LBM_ERR_HIP, "copy of the tile table failed");
    }
    b.laid_out = true;
    return LBM_OK;
}

static void put_r(std::vector<uint8_t>& v, size_t i, size_t rb, double x) {
    if (rb == 8) std::memcpy(v.data() + i * 8, &x, 8);
    else {
        const float y = (float)x;
        std::memcpy(v.data() + i * 4, &y, 4);
    }
}

// Tiled host copies of flags, bc and links, copied to the device; summaries and case records.
static int stage(Batch& b) {
    // A captured graph holds the kernel variant (with or without links) and the link buffer
    // pointers of the staging it was recorded after; either may change here.
    free_graph(b);
    const int D = b.proto.D, Q = b.proto.Q;
    b.h_flags.assign(b.fl_bytes, (uint8_t)LBM_FLAG_SOLID);  // padding nodes are SOLID
    b.h_bc.assign(b.bc_elems * b.r_bytes, 0);
    b.h_links.assign(b.cases.size(), {});
    b.h_qd.assign(b.cases.size(), {});
    size_t total_links = 0;
    for (size_t c = 0; c < b.cases.size(); ++c) {
        const BatchCaseHost& hc = b.cases[c];
        const Geo& g = b.geo[c];
        uint8_t* fl = b.h_flags.data() + b.fl_off[c];
        for (int z = 0; z < g.nz; ++z)
            for (int y = 0; y < g.ny; ++y)
                for (int x = 0; x < g.nx; ++x) {
                    const bool real = x < hc.n[0] && y < hc.n[1] && z < hc.n[2];
                    const size_t r = real ? ((size_t)x * hc.n[1] + y) * hc.n[2] + z : 0;
                    const size_t ti = tiled_index(g, x, y, z, 1, 0);
                    if (real) fl[ti] = hc.flags.empty() ? (uint8_t)LBM_FLAG_FLUID : hc.flags[r];
                    for (int k = 0; k <= D; ++k) {
                        double v = k < D ? 0.0 : 1.0;
                        if (real && !hc.bc.empty()) v = hc.bc[r * (D + 1) + k];
                        put_r(b.h_bc, b.bc_off[c] + tiled_index(g, x, y, z, D + 1, k), b.r_bytes, v);
                    }
                }
        // links: row-major ids -> tiled ids, records sorted by tiled id
        const size_t nl = hc.links.size();
        std::vector<std::pair<uint32_t, uint32_t>> order(nl);
        for (size_t i = 0; i < nl; ++i) {
            const uint32_t r = hc.links[i];
            const int z = (int)(r % (uint32_t)hc.n[2]);
            const int y = (int)((r / (uint32_t)hc.n[2]) % (uint32_t)hc.n[1]);
            const int x = (int)(r / ((uint32_t)hc.n[2] * (uint32_t)hc.n[1]));
            order[i] = {(uint32_t)tiled_index(g, x, y, z, 1, 0), (uint32_t)i};
        }
        std::sort(order.begin(), order.end());
        b.h_links[c].resize(nl);
        b.h_qd[c].resize(nl * (Q - 1));
        for (size_t i = 0; i < nl; ++i) {
            b.h_links[c][i] = order[i].first;
            std::memcpy(&b.h_qd[c][i * (Q - 1)], &hc.qd[(size_t)order[i].second * (Q - 1)], (Q - 1) * sizeof(float));
        }
        total_links += nl;
    }
    if (total_links > b.link_cap) {
        if (b.link_nodes) (void)hipFree(b.link_nodes);
        if (b.link_q) (void)hipFree(b.link_q);
        b.link_nodes = nullptr;
        b.link_q = nullptr;
        b.link_cap = 0;
        if (hipMalloc((void**)&b.link_nodes, total_links * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc((void**)&b.link_q, total_links * (Q - 1) * sizeof(float)) != hipSuccess)
            return set_error(LBM_ERR_HIP, "hipMalloc of the batch links failed");
        b.link_cap = total_links;
    }
    std::vector<uint32_t> all_links;
    std::vector<float> all_q;
    all_links.reserve(total_links);
    all_q.reserve(total_links * (Q - 1));
    for (size_t c = 0; c < b.cases.size(); ++c) {
        all_links.insert(all_links.end(), b.h_links[c].begin(), b.h_links[c].end());
        all_q.insert(all_q.end(), b.h_qd[c].begin(), b.h_qd[c].end());
    }
    if (hipMemcpy(b.flags, b.h_flags.data(), b.fl_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.bc, b.h_bc.data(), b.h_bc.size(), hipMemcpyHostToDevice) != hipSuccess)
        return set_error(LBM_ERR_HIP, "copy of the batch flags / bc failed");
    if (total_links &&
        (hipMemcpy(b.link_nodes, all_links.data(), total_links * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(b.link_q, all_q.data(), all_q.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess))
        return set_error(LBM_ERR_HIP, "copy of the batch links failed");
    b.any_links = total_links > 0;
    int rc = batch_dispatch(b.proto, [&](auto l) { return decltype(l)::prepare(b, (hipStream_t)0); });
    if (rc) return rc;
    if (hipDeviceSynchronize() != hipSuccess) return set_error(LBM_ERR_HIP, "batch staging failed");
    b.staged = true;
    return LBM_OK;
}

static int step_once(Batch& b, int parity, hipStream_t st) {
    return batch_dispatch(b.proto, [&](auto l) { return decltype(l)::step(b, parity, st); });
}

static int capture_graph(Batch& b) {
    free_graph(b);
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        return set_error(LBM_ERR_HIP, "stream for graph capture");
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return set_error(LBM_ERR_HIP, "hipStreamBeginCapture failed");
    }
    int rc = LBM_OK;
    for (int i = 0; i < b.graph_steps && rc == LBM_OK; ++i) rc = step_once(b, i & 1, cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) return set_error(LBM_ERR_HIP, "hipStreamEndCapture failed");
    b.graph_def = g;
    if (hipGraphInstantiate(&b.graph, g, nullptr, nullptr, 0) != hipSuccess) {
        free_graph(b);
        return set_error(LBM_ERR_HIP, "hipGraphInstantiate failed");
    }
    return LBM_OK;
}

static int run(Batch& b, int64_t steps, hipStream_t st) {
    if (!b.initialized) return set_error(LBM_ERR_INVALID, "initialise the batch state first");
    if (!b.staged)
        if (int rc = stage(b)) return rc;
    while (steps > 0) {
        if (b.graph_steps > 0 && b.parity == 0 && steps >= b.graph_steps) {
            if (!b.graph)
                if (int rc = capture_graph(b)) return rc;
            if (hipGraphLaunch(b.graph, st) != hipSuccess) return set_error(LBM_ERR_HIP, "hipGraphLaunch failed");
            steps -= b.graph_steps;
        } else {
            if (int rc = step_once(b, b.parity, st)) return rc;
            b.parity ^= 1;
            --steps;
        }
        b.force_valid = b.any_links;
    }
    return LBM_OK;
}

static int init_state(Batch& b, hipStream_t st) {
    if (!b.laid_out)
        if (int rc = layout(b)) return rc;
    if (!b.staged)
        if (int rc = stage(b)) return rc;
    int rc = batch_dispatch(b.proto, [&](auto l) { return decltype(l)::init(b, st); });
    if (rc) return rc;
    b.parity = 0;
    b.initialized = true;
    b.force_valid = false;
    return LBM_OK;
}

static int set_tile(Batch& b, const int t[3]) {
    std::vector<lbm_desc> cd;
    if (int rc = case_descs(b, t, cd)) return rc;
    b.tile[0] = t[0];
    b.tile[1] = t[1];
    b.tile[2] = t[2];
    return layout(b);
}

static int time_steps(Batch& b, int steps, hipStream_t st, float* ms) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return set_error(LBM_ERR_HIP, "hipEventCreate failed");
    int rc = run(b, 10, st);
    if (!rc) (void)hipEventRecord(e0, st);
    if (!rc) rc = run(b, steps, st);
    if (!rc) {
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) rc = set_error(LBM_ERR_HIP, "timing run failed");
        else (void)hipEventElapsedTime(ms, e0, e1);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace lbm

using namespace lbm;

extern "C" {

int lbm_batch_create(const lbm_desc* proto, int32_t n_cases, const int32_t* sizes, const double* taus,
                     lbm_batch** out) {
    if (!out) return set_error(LBM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!proto || !sizes || !taus || n_cases <= 0) return set_error(LBM_ERR_INVALID, "empty or NULL case list");
    if (proto->method != LBM_DOUBLE_BUFFER)
        return set_error(LBM_ERR_UNSUPPORTED, "batched cases use the double-buffer method");
    if (proto->les == LBM_LES_MLP) return set_error(LBM_ERR_UNSUPPORTED, "batched cases take no learned closure");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(LBM_ERR_NO_DEVICE, "no HIP device visible");
    Batch* b = new Batch();
    b->proto = *proto;
    b->tile[0] = proto->tx;
    b->tile[1] = proto->ty;
    b->tile[2] = proto->tz;
    b->cases.resize(n_cases);
    for (int c = 0; c < n_cases; ++c) {
        BatchCaseHost& hc = b->cases[c];
        for (int k = 0; k < 3; ++k) hc.n[k] = sizes[3 * c + k];
        hc.tau = taus[c];
        if (hc.n[0] <= 0 || hc.n[1] <= 0 || hc.n[2] <= 0 || (proto->D == 2 && hc.n[2] != 1) || !(hc.tau > 0.5)) {
            delete b;
            return set_error(LBM_ERR_INVALID, "case sizes must be positive (nz = 1 in 2-D) and tau > 1/2");
        }
    }
    (void)hipGetDevice(&b->device);
    if (int rc = layout(*b)) {
        free_device(*b);
        delete b;
        return rc;
    }
    *out = reinterpret_cast<lbm_batch*>(b);
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch* bb) {
    if (!bb) return LBM_OK;
    free_device(*B(bb));
    delete B(bb);
    return LBM_OK;
}

int lbm_batch_set_case(lbm_batch* bb, int32_t c, const uint8_t* flags, const double* bc, const uint32_t* link_nodes,
                       const float* qdist, uint32_t n_links) {
    if (!bb) return set_error(LBM_ERR_INVALID, "batch is NULL");
    Batch& b = *B(bb);
    if (c < 0 || c >= (int)b.cases.size()) return set_error(LBM_ERR_INVALID, "case index out of range");
    BatchCaseHost& hc = b.cases[c];
    const size_t cells = (size_t)hc.n[0] * hc.n[1] * hc.n[2];
    if (n_links && (!link_nodes || !qdist)) return set_error(LBM_ERR_INVALID, "NULL link arrays");
    for (uint32_t i = 0; i < n_links; ++i) {
        if (link_nodes[i] >= cells) return set_error(LBM_ERR_INVALID, "link node index outside the case");
        if (i && link_nodes[i] <= link_nodes[i - 1])
            return set_error(LBM_ERR_INVALID, "link nodes must be strictly increasing");
    }
    const int D = b.proto.D, Q = b.proto.Q;
    hc.flags.clear();
    hc.bc.clear();
    hc.links.clear();
    hc.qd.clear();
    if (flags) hc.flags.assign(flags, flags + cells);
    if (bc) hc.bc.assign(bc, bc + cells * (D + 1));
    if (n_links) {
        hc.links.assign(link_nodes, link_nodes + n_links);
        hc.qd.assign(qdist, qdist + (size_t)n_links * (Q - 1));
    }
    b.staged = false;
    return LBM_OK;
}

int lbm_batch_set_tile(lbm_batch* bb, int32_t tx, int32_t ty, int32_t tz) {
    if (!bb) return set_error(LBM_ERR_INVALID, "batch is NULL");
    const int t[3] = {tx, ty, tz};
    return set_tile(*B(bb), t);
}

int lbm_batch_set_option(lbm_batch* bb, int option, int value) {
    if (!bb) return set_error(LBM_ERR_INVALID, "batch is NULL");
    Batch& b = *B(bb);
    if (option == LBM_BATCH_OPT_GRID_BLOCKS) {
        if (value < 0) return set_error(LBM_ERR_INVALID, "grid blocks must be >= 0");
        b.grid_blocks = (value + 7) / 8 * 8;
        free_graph(b);
        return LBM_OK;
    }
    if (option == LBM_BATCH_OPT_GRAPH_STEPS) {
        if (value < 0 || (value & 1)) return set_error(LBM_ERR_INVALID, "graph steps must be even and >= 0");
        b.graph_steps = value;
        free_graph(b);
        return LBM_OK;
    }
    return set_error(LBM_ERR_INVALID, "unknown batch option");
}

int lbm_batch_init_equilibrium(lbm_batch* bb, void* stream) {
    if (!bb) return set_error(LBM_ERR_INVALID, "batch is NULL");
    return init_state(*B(bb), (hipStream_t)stream);
}

int lbm_batch_run(lbm_batch* bb, int64_t steps, void* stream) {
    if (!bb) return set_error(LBM_ERR_INVALID, "batch is NULL");
    if (steps < 0) return set_error(LBM_ERR_INVALID, "steps must be >= 0");
    return run(*B(bb), steps, (hipStream_t)stream);
}

int lbm_batch_force(lbm_batch* bb, double* out_dev, void* stream) {
    if (!bb) return set_error(LBM_ERR_INVALID, "batch is NULL");
    if (!out_dev) return set_error(LBM_ERR_INVALID, "NULL buffer");
    Batch& b = *B(bb);
    if (!b.force_valid) return set_error(LBM_ERR_INVALID, "no step with links has run since the state was set");
    return batch_dispatch(b.proto, [&](auto l) { return decltype(l)::force(b, out_dev, (hipStream_t)stream); });
}

int lbm_batch_case_desc(lbm_batch* bb, int32_t c, lbm_desc* out) {
    if (!bb || !out) return set_error(LBM_ERR_INVALID, "NULL argument");
    Batch& b = *B(bb);
    if (c < 0 || c >= (int)b.cases.size()) return set_error(LBM_ERR_INVALID, "case index out of range");
    if (!b.laid_out)
        if (int rc = layout(b)) return rc;
    *out = b.cdesc[c];
    return LBM_OK;
}

int lbm_batch_export_case(lbm_batch* bb, int32_t c, uint8_t* flags, void* bc, uint32_t* link_nodes, float* qdist,
                          uint32_t* n_links) {
    if (!bb || !n_links) return set_error(LBM_ERR_INVALID, "NULL argument");
    Batch& b = *B(bb);
    if (c < 0 || c >= (int)b.cases.size()) return set_error(LBM_ERR_INVALID, "case index out of range");
    if (!b.laid_out)
        if (int rc = layout(b)) return rc;
    if (!b.staged)
        if (int rc = stage(b)) return rc;
    const Geo& g = b.geo[c];
    const size_t cells = (size_t)g.n_tiles * g.tile_cells;
    if (flags) std::memcpy(flags, b.h_flags.data() + b.fl_off[c], cells);
    if (bc) std::memcpy(bc, b.h_bc.data() + b.bc_off[c] * b.r_bytes, cells * (b.proto.D + 1) * b.r_bytes);
    const size_t nl = b.h_links[c].size();
    if (link_nodes && nl) std::memcpy(link_nodes, b.h_links[c].data(), nl * sizeof(uint32_t));
    if (qdist && nl) std::memcpy(qdist, b.h_qd[c].data(), b.h_qd[c].size() * sizeof(float));
    *n_links = (uint32_t)nl;
    return LBM_OK;
}

static int case_f(lbm_batch* bb, int32_t c, size_t* off, size_t* bytes) {
    if (!bb) return set_error(LBM_ERR_INVALID, "batch is NULL");
    Batch& b = *B(bb);
    if (c < 0 || c >= (int)b.cases.size()) return set_error(LBM_ERR_INVALID, "case index out of range");
    if (!b.laid_out) return set_error(LBM_ERR_INVALID, "batch has no layout");
    *off = b.f_off[c] * b.f_elem_bytes;
    *bytes = (size_t)b.case_tiles[c] * b.tile_cells * b.proto.Q * b.f_elem_bytes;
    return LBM_OK;
}

int lbm_batch_get_f(lbm_batch* bb, int32_t c, void* dst_dev, void* stream) {
    size_t off = 0, bytes = 0;
    if (int rc = case_f(bb, c, &off, &bytes)) return rc;
    if (!dst_dev) return set_error(LBM_ERR_INVALID, "NULL buffer");
    Batch& b = *B(bb);
    if (!b.initialized) return set_error(LBM_ERR_INVALID, "batch state not initialised");
    return hip_rc(hipMemcpyAsync(dst_dev, (const uint8_t*)b.f[b.parity] + off, bytes, hipMemcpyDeviceToDevice,
                                 (hipStream_t)stream),
                  "hipMemcpyAsync");
}

int lbm_batch_set_f(lbm_batch* bb, int32_t c, const void* src_dev, void* stream) {
    size_t off = 0, bytes = 0;
    if (int rc = case_f(bb, c, &off, &bytes)) return rc;
    if (!src_dev) return set_error(LBM_ERR_INVALID, "NULL buffer");
    Batch& b = *B(bb);
    if (!b.initialized) return set_error(LBM_ERR_INVALID, "initialise the batch state before replacing a case");
    for (int k = 0; k < 2; ++k)
        if (int rc = hip_rc(hipMemcpyAsync((uint8_t*)b.f[k] + off, src_dev, bytes, hipMemcpyDeviceToDevice,
                                           (hipStream_t)stream),
                            "hipMemcpyAsync"))
            return rc;
    b.force_valid = false;
    return LBM_OK;
}

int lbm_batch_autotune(lbm_batch* bb, int32_t steps, void* stream, int32_t* best) {
    if (!bb || !best) return set_error(LBM_ERR_INVALID, "NULL argument");
    if (steps <= 0) return set_error(LBM_ERR_INVALID, "steps must be > 0");
    Batch& b = *B(bb);
    hipStream_t st = (hipStream_t)stream;
    static const int cand2[][3] = {{16, 16, 1}, {32, 8, 1}, {64, 4, 1}, {8, 32, 1}, {32, 16, 1},
                                   {64, 8, 1},  {128, 4, 1}, {16, 32, 1}, {128, 8, 1}, {256, 4, 1}};
    static const int cand3[][3] = {{8, 8, 8}, {16, 8, 4}, {32, 4, 4}, {16, 4, 8}, {8, 8, 4}, {16, 16, 2}, {32, 8, 2}};
    const int n_cand = b.proto.D == 2 ? 10 : 7;
    const int saved_grid = b.grid_blocks;
    int bt[3] = {b.tile[0], b.tile[1], b.tile[2]};
    float best_ms = 1e30f;
    b.grid_blocks = 0;
    for (int i = 0; i < n_cand; ++i) {
        const int* t = b.proto.D == 2 ? cand2[i] : cand3[i];
        std::vector<lbm_desc> cd;
        if (case_descs(b, t, cd) != LBM_OK) continue;
        if (int rc = set_tile(b, t)) return rc;
        if (int rc = init_state(b, st)) return rc;
        float ms = 0;
        if (int rc = time_steps(b, steps, st, &ms)) return rc;
        if (ms < best_ms) {
            best_ms = ms;
            bt[0] = t[0];
            bt[1] = t[1];
            bt[2] = t[2];
        }
    }
    if (int rc = set_tile(b, bt)) return rc;
    int best_grid = 0;
    if (int rc = init_state(b, st)) return rc;
    float ms0 = 0;
    if (int rc = time_steps(b, steps, st, &ms0)) return rc;
    static const int grids[] = {512, 1024, 2048, 4096};
    for (int g : grids) {
        if ((uint32_t)g >= 8u * b.chunk) continue;
        b.grid_blocks = g;
        free_graph(b);
        float ms = 0;
        if (int rc = time_steps(b, steps, st, &ms)) return rc;
        if (ms < ms0) {
            ms0 = ms;
            best_grid = g;
        }
    }
    (void)saved_grid;
    b.grid_blocks = best_grid;
    free_graph(b);
    if (int rc = init_state(b, st)) return rc;
    best[0] = bt[0];
    best[1] = bt[1];
    best[2] = bt[2];
    best[3] = best_grid;
    return LBM_OK;
}

int lbm_batch_info(lbm_batch* bb, int64_t* out) {
    if (!bb || !out) return set_error(LBM_ERR_INVALID, "NULL argument");
    Batch& b = *B(bb);
    out[0] = (int64_t)b.cases.size();
    out[1] = (int64_t)b.real_nodes;
    out[2] = (int64_t)b.fl_bytes;
    out[3] = (int64_t)b.n_tiles;
    out[4] = b.tile[0];
    out[5] = b.tile[1];
    out[6] = b.tile[2];
    out[7] = b.grid_blocks;
    out[8] = b.graph_steps;
    return LBM_OK;
}

}  // extern "C"

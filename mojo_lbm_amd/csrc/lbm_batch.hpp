// lbm_batch.hpp -- many independent double-buffer cases of different sizes and
// relaxation times advanced by ONE kernel launch per step.
//
// Every case keeps its own tiled grid (padded to tile multiples with SOLID
// nodes) and its own slice of two concatenated population buffers, boundary
// data, flags, flag summaries and interpolated bounce-back links. A batch tile
// table maps the launch's tile index to (case, tile of the case); the tile list
// is walked in XCD-contiguous order, optionally by a persistent grid that loops
// over it. Flags are static between lbm_batch_set_case calls, so the per-tile
// flag summaries are built once, not per step. Runs of many steps replay a
// captured hipGraph of graph_steps launches.
#pragma once

#include <vector>

#include "lbm_launch.hpp"

namespace lbm {

// Host-side copy of one case as set by the caller (row-major (x, y, z) node order).
struct BatchCaseHost {
    int n[3] = {0, 0, 0};
    double tau = 1.0;
    std::vector<uint8_t> flags;   // empty: every node FLUID
    std::vector<double> bc;       // (D+1) per node; empty: u = 0, rho = 1
    std::vector<uint32_t> links;  // row-major node ids, increasing
    std::vector<float> qd;        // (Q-1) per link
};

struct Batch {
    lbm_desc proto{};
    int device = 0;
    std::vector<BatchCaseHost> cases;
    int tile[3] = {0, 0, 0};
    // layout of the current tile shape
    bool laid_out = false, staged = false, initialized = false;
    std::vector<lbm_desc> cdesc;  // padded per-case descriptions
    std::vector<Geo> geo;
    std::vector<size_t> f_off, bc_off, fl_off;  // element offsets of each case
    std::vector<uint32_t> tile_begin, case_tiles;
    uint32_t n_tiles = 0, chunk = 0, tile_cells = 0;
    size_t lds_bytes = 0, f_elems = 0, bc_elems = 0, fl_bytes = 0, f_elem_bytes = 4, r_bytes = 4;
    uint64_t real_nodes = 0;
    // staged (tiled) host copies, kept for lbm_batch_export_case
    std::vector<uint8_t> h_flags;
    std::vector<uint8_t> h_bc;                       // R elements
    std::vector<std::vector<uint32_t>> h_links;      // tiled node ids per case, increasing
    std::vector<std::vector<float>> h_qd;
    // device buffers
    void* f[2] = {nullptr, nullptr};
    void* bc = nullptr;
    uint8_t* flags = nullptr;
    uint32_t* summary = nullptr;
    uint32_t* tile_kind = nullptr;  // per batch tile: boundary / moving-wall bits
    uint32_t* nbmask = nullptr;     // per node of the batch: neighbour masks of boundary tiles
    uint8_t* mvs = nullptr;         // per node of the batch: moving SOLID nodes
    uint32_t* btiles = nullptr;     // boundary tile ids of one case, then their count (staging)
    uint32_t* link_nodes = nullptr;
    float* link_q = nullptr;
    size_t link_cap = 0;
    double* force_partial = nullptr;
    void* recs = nullptr;
    uint32_t* tile_case = nullptr;
    // execution
    int grid_blocks = 0;  // 0: one workgroup per tile, else a persistent grid of this many
    int graph_steps = 100;
    int parity = 0;       // buffer holding the current state
    bool force_valid = false;
    bool any_links = false;
    hipGraph_t graph_def = nullptr;
    hipGraphExec_t graph = nullptr;
};

// Entry points of one (velocity set, storage) pair, defined in lbm_batch_impl.hpp and
// instantiated in lbm_batch_inst_*.hip.
template <class L, class S> struct BatchLaunch {
    static int step(const Batch& b, int parity, hipStream_t st);
    static int prepare(Batch& b, hipStream_t st);  // per-case flag summaries, masks + device case records
    static size_t rec_bytes();
    static int init(const Batch& b, hipStream_t st);
    static int force(const Batch& b, double* out, hipStream_t st);
};

#define LBM_EXTERN_BATCH(L, S) extern template struct BatchLaunch<L, S>;
LBM_FOR_EACH_INSTANCE(LBM_EXTERN_BATCH)

}  // namespace lbm

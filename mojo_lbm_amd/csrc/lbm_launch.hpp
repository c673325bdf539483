// lbm_launch.hpp -- the handle and the declarations of the host-side launchers,
// one Launch<Lattice, Storage> per (velocity set, population storage).
#pragma once

#include <string>

#include "lbm_kernels.hpp"

namespace lbm {

struct Handle {
    lbm_desc d;
    Geo g;
    int device = 0;
    size_t lds_bytes = 0;        // flag tile + halo
    size_t f_bytes = 0;          // one population buffer
    bool big = false;            // f_bytes >= 2^31: 64-bit addressing
    double* partial = nullptr;   // per-tile reduction partials (n_tiles * (D+1))
    uint32_t* summary = nullptr; // per-tile region masks, rebuilt before every step
    uint32_t* tile_kind = nullptr;  // per tile: boundary / moving-wall bits (node_mask_kernel)
    uint32_t* nbmask = nullptr;     // per node: neighbour mask (read for boundary tiles only)
    uint8_t* mvs = nullptr;         // per node: 1 for a moving SOLID node (moving walls included)
    uint32_t* btiles = nullptr;     // boundary tile ids, then their count (n_tiles + 1 words)
    uint8_t* flags_seen = nullptr;  // per node: the flags the masks were built from
    uint32_t* mask_changed = nullptr;  // set when the flags or moving-wall bytes changed (cleared after the rebuild)
    bool static_flags = false;   // LBM_OPT_STATIC_FLAGS: rebuild only for new flags / bc buffers
    int staged_pull = 1;         // LBM_OPT_STAGED_PULL: Float16C double-buffer pulls through LDS (2: tile pairs)
    mutable const void* summary_flags = nullptr;  // buffers the summary was built from
    mutable const void* summary_bc = nullptr;
    uint32_t* link_nodes = nullptr;  // interpolated bounce-back records (handle-owned copies)
    float* link_q = nullptr;
    uint32_t n_links = 0;
    double* force_partial = nullptr; // per wave of every tile: D doubles
    double* node_force = nullptr;    // caller's per-link-record force output (lbm_set_link_node_force)
    float* mlp = nullptr;            // learned-closure parameter block (lbm_set_mlp)
    bool force_valid = false;        // a step with links has run since lbm_set_links
};

int set_error(int code, const std::string& msg);
int check_launch(const char* what);
int hip_rc(hipError_t e, const char* what);
int validate(const lbm_desc* d);  // description limits; sets the error text
Geo make_geo(const lbm_desc* d);

// Entry points of one (velocity set, storage) pair. Defined in lbm_launch_impl.hpp and
// instantiated once per pair in lbm_inst_*.hip; other units see only these declarations.
template <class L, class S> struct Launch {
    static int step(const Handle& h, void* f_out, const void* f_in, const void* bc, const uint8_t* flags, double tau,
                    int mode, hipStream_t st);
    static int macro(const Handle& h, void* rho, void* u, const void* f, const void* bc, const uint8_t* flags,
                     int mode, int u_layout, hipStream_t st);
    static int qcrit(const Handle& h, void* q, const void* u, const void* f, const void* bc, const uint8_t* flags,
                     double tau, int mode, int u_layout, hipStream_t st);
    static int reduce(const Handle& h, double* out, const void* f, const void* bc, const uint8_t* flags, int mode,
                      hipStream_t st);
    static int link_force(const Handle& h, double* out, hipStream_t st);
    static int init(const Handle& h, void* f, double rho_d, const double* u_d, hipStream_t st);
};

// Launchers are instantiated once per (lattice, storage) in lbm_inst_*.hip.
#define LBM_FOR_EACH_INSTANCE(X)        \
    X(Lattice<9>, StoreF32)             \
    X(Lattice<9>, StoreF16C)            \
    X(Lattice<9>, StoreF64)             \
    X(Lattice<19>, StoreF32)            \
    X(Lattice<19>, StoreF16C)           \
    X(Lattice<19>, StoreF64)            \
    X(Lattice<27>, StoreF32)            \
    X(Lattice<27>, StoreF16C)           \
    X(Lattice<27>, StoreF64)

#define LBM_EXTERN_INSTANCE(L, S) extern template struct Launch<L, S>;
LBM_FOR_EACH_INSTANCE(LBM_EXTERN_INSTANCE)

}  // namespace lbm

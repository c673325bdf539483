// lbm_capi.hip -- the extern "C" boundary of liblbm_hip.so (declared in include/lbm_hip.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_launch.hpp"

namespace lbm {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LBM_OK;
}

int hip_rc(hipError_t e, const char* what) {
    if (e != hipSuccess) return set_error(LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LBM_OK;
}

static bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
static int log2i(int v) {
    int b = 0;
    while ((1 << b) < v) ++b;
    return b;
}

// Validation of a run description: the reference's comptime asserts and
// host checks (grid.mojo:392-440, config.mojo:100-150, esoteric_pull.mojo:76-78,
// KBC_.mojo:220) plus this library's own limits.
int validate(const lbm_desc* d) {
    if (!d) return set_error(LBM_ERR_INVALID, "desc is NULL");
    const bool lat_ok = (d->D == 2 && d->Q == 9) || (d->D == 3 && (d->Q == 19 || d->Q == 27));
    if (!lat_ok) return set_error(LBM_ERR_INVALID, "lattice must be D2Q9, D3Q19 or D3Q27");
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return set_error(LBM_ERR_INVALID, "grid shape must be positive");
    if (d->D == 2 && (d->nz != 1 || d->tz != 1))
        return set_error(LBM_ERR_INVALID, "a 2-D grid needs nz = 1 and tz = 1");
    const int t[3] = {d->tx, d->ty, d->tz};
    for (int i = 0; i < 3; ++i) {
        if (!pow2(t[i])) return set_error(LBM_ERR_INVALID, "tile dimensions must be powers of two");
        if (i < d->D && t[i] < 2) return set_error(LBM_ERR_INVALID, "tile dimensions must be > 1 along each lattice axis");
    }
    if (d->nx % d->tx || d->ny % d->ty || d->nz % d->tz)
        return set_error(LBM_ERR_INVALID,
                         "The grid shape along nx,ny and nz should be divisible by the block (tile) shape");
    const long cells = (long)d->tx * d->ty * d->tz;
    const long maxthreads = d->D == 2 ? 1024 : 512;
    if (cells > maxthreads) return set_error(LBM_ERR_INVALID, "tile holds more nodes than one workgroup can (2-D: 1024, 3-D: 512)");
    const long halo = (long)(d->tx + 2) * (d->ty + 2) * (d->D == 3 ? d->tz + 2 : 1);
    if (cells < 32 || halo > 8 * cells)
        return set_error(LBM_ERR_INVALID, "tile too small: needs >= 32 nodes and a halo of at most 8 nodes per node");
    if (d->method != LBM_DOUBLE_BUFFER && d->method != LBM_ESOTERIC_PULL)
        return set_error(LBM_ERR_INVALID, "method must be double buffer or esoteric pull");
    if (d->collision != LBM_COLL_SRT && d->collision != LBM_COLL_TRT && d->collision != LBM_COLL_KBC &&
        d->collision != LBM_COLL_RLBM)
        return set_error(LBM_ERR_INVALID, "unknown collision operator");
    if (d->storage < LBM_STORE_F32 || d->storage > LBM_STORE_F64) return set_error(LBM_ERR_INVALID, "unknown storage");
    if (d->les != LBM_LES_NONE && d->les != LBM_LES_SMAGORINSKY && d->les != LBM_LES_MLP)
        return set_error(LBM_ERR_INVALID, "unknown LES model");
    if (d->les == LBM_LES_MLP && d->D != 3)
        return set_error(LBM_ERR_UNSUPPORTED, "the learned closure needs a 3-D lattice");
    if (!(d->smagorinsky_cs >= 0.0)) return set_error(LBM_ERR_INVALID, "smagorinsky_cs must be >= 0");
    // Node indices (flags, link records) are 32-bit; population and bc offsets are 64-bit in the
    // kernels that buffers of 2 GiB and more use, so only the node count is bounded.
    const double nodes = (double)d->nx * d->ny * d->nz;
    if (nodes >= 4294967296.0) return set_error(LBM_ERR_UNSUPPORTED, "2^32 nodes or more");
    return LBM_OK;
}

// log2 of the z layers of a step workgroup of `lanes` lanes, or -1 if that many lanes do
// not make whole z layers of the tile (a power of two, at least one wave, at most the
// tile) or the slab's halo would exceed 8 nodes per lane.
static int step_slab_bits(const Geo& g, int D, int lanes) {
    const int layer = 1 << (g.txb + g.tyb);
    if (lanes < 64 || lanes > g.tile_cells || (lanes & (lanes - 1)) || lanes % layer) return -1;
    const int wzb = log2i(lanes / layer);
    const long halo = (long)((1 << g.txb) + 2) * ((1 << g.tyb) + 2) * (D == 3 ? (1 << wzb) + 2 : 1);
    return halo > 8L * lanes ? -1 : wzb;
}

Geo make_geo(const lbm_desc* d) {
    Geo g{};
    g.nx = d->nx; g.ny = d->ny; g.nz = d->nz;
    g.txb = log2i(d->tx); g.tyb = log2i(d->ty); g.tzb = log2i(d->tz);
    g.Tx = d->nx / d->tx; g.Ty = d->ny / d->ty; g.Tz = d->nz / d->tz;
    g.tile_cells = d->tx * d->ty * d->tz;
    g.n_tiles = g.Tx * g.Ty * g.Tz;
    // XCD-contiguous tile order when the tile count allows it (measured faster on MI355X for
    // the wall-bounded and Float16C cases); lbm_set_option(LBM_OPT_XCD_REMAP) overrides it.
    g.xcd_chunk = (g.n_tiles % 8 == 0) ? g.n_tiles / 8 : 0;
    // One step workgroup per tile, except 3-D tiles of more than 256 nodes with Float16C storage
    // or the learned closure: 256-lane workgroups on z slabs (measured on MI355X with 8x8x8
    // tiles: Float16C 9-13 % faster at 256^3, the learned closure 7 % at 128^3; FP32 with
    // SRT does not gain at 256^3). LBM_OPT_STEP_LANES overrides.
    g.wzb = g.tzb;
    if ((d->storage == LBM_STORE_F16C || d->les == LBM_LES_MLP) && d->D == 3 && g.tile_cells > 256) {
        const int wzb = step_slab_bits(g, d->D, 256);
        if (wzb >= 0) g.wzb = wzb;
    }
    const uint64_t hx = (uint64_t)d->tx + 2, hxy = hx * ((uint64_t)d->ty + 2);
    g.hx_magic = (uint32_t)((1ull << 32) / hx + 1);
    g.hxy_magic = (uint32_t)((1ull << 32) / hxy + 1);
    return g;
}

// Dispatch on (Q, storage) to the matching Launch instantiation.
template <class F> static int dispatch(const lbm_desc& d, F&& f) {
#define LBM_CASE(QQ, SS, SID) \
    if (d.Q == QQ && d.storage == SID) return f(Launch<Lattice<QQ>, SS>{});
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

static int check_handle(lbm_handle* h) {
    if (!h) return set_error(LBM_ERR_INVALID, "handle is NULL");
    return LBM_OK;
}

static Handle* H(lbm_handle* h) { return reinterpret_cast<Handle*>(h); }

static size_t grid_bytes(const lbm_desc* d) { return d->storage == LBM_STORE_F64 ? 8 : 4; }

}  // namespace lbm

using namespace lbm;

extern "C" {

int lbm_version(void) { return 100; }

const char* lbm_last_error(void) { return g_last_error.c_str(); }

int lbm_device_count(int* count) {
    if (!count) return set_error(LBM_ERR_INVALID, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return LBM_OK;
}

size_t lbm_f_bytes(const lbm_desc* d) {
    if (validate(d)) return 0;
    const size_t s = d->storage == LBM_STORE_F16C ? 2 : (d->storage == LBM_STORE_F64 ? 8 : 4);
    return (size_t)d->nx * d->ny * d->nz * d->Q * s;
}
size_t lbm_bc_bytes(const lbm_desc* d) {
    if (validate(d)) return 0;
    return (size_t)d->nx * d->ny * d->nz * (d->D + 1) * grid_bytes(d);
}
size_t lbm_flags_bytes(const lbm_desc* d) {
    if (validate(d)) return 0;
    return (size_t)d->nx * d->ny * d->nz;
}

int lbm_create(const lbm_desc* desc, lbm_handle** out) {
    if (!out) return set_error(LBM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate(desc);
    if (rc) return rc;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_error(LBM_ERR_NO_DEVICE, "no HIP device visible");
    Handle* h = new Handle();
    h->d = *desc;
    h->g = make_geo(desc);
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return set_error(LBM_ERR_NO_DEVICE, "hipGetDevice failed");
    }
    const int hz = desc->D == 3 ? desc->tz + 2 : 1;
    h->lds_bytes = (size_t)(desc->tx + 2) * (desc->ty + 2) * hz;
    const size_t esz = desc->storage == LBM_STORE_F16C ? 2 : (desc->storage == LBM_STORE_F64 ? 8 : 4);
    h->f_bytes = (size_t)desc->nx * desc->ny * desc->nz * desc->Q * esz;
    h->big = h->f_bytes >= (1ull << 31);
    const size_t pb = (size_t)h->g.n_tiles * (desc->D + 1) * sizeof(double);
    if (hipMalloc((void**)&h->partial, pb) != hipSuccess) {
        delete h;
        return set_error(LBM_ERR_HIP, "hipMalloc of the reduction partials failed");
    }
    if (hipMalloc((void**)&h->summary, (size_t)h->g.n_tiles * sizeof(uint32_t)) != hipSuccess) {
        (void)hipFree(h->partial);
        delete h;
        return set_error(LBM_ERR_HIP, "hipMalloc of the tile summary failed");
    }
    // tile kinds and the per-node neighbour masks (4 bytes per node; read for boundary tiles only)
    if (hipMalloc((void**)&h->tile_kind, (size_t)h->g.n_tiles * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&h->nbmask, (size_t)h->g.n_tiles * h->g.tile_cells * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&h->mvs, (size_t)h->g.n_tiles * h->g.tile_cells) != hipSuccess ||
        hipMalloc((void**)&h->btiles, ((size_t)h->g.n_tiles + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&h->flags_seen, (size_t)h->g.n_tiles * h->g.tile_cells) != hipSuccess ||
        hipMalloc((void**)&h->mask_changed, sizeof(uint32_t)) != hipSuccess) {
        if (h->flags_seen) (void)hipFree(h->flags_seen);
        if (h->btiles) (void)hipFree(h->btiles);
        if (h->mvs) (void)hipFree(h->mvs);
        if (h->nbmask) (void)hipFree(h->nbmask);
        if (h->tile_kind) (void)hipFree(h->tile_kind);
        (void)hipFree(h->summary);
        (void)hipFree(h->partial);
        delete h;
        return set_error(LBM_ERR_HIP, "hipMalloc of the neighbour masks failed");
    }
    // the first check always builds: changed = 1, no boundary tiles listed yet
    const uint32_t one = 1u, zero = 0u;
    if (hipMemcpy(h->mask_changed, &one, sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->btiles + h->g.n_tiles, &zero, sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
        (void)lbm_destroy(reinterpret_cast<lbm_handle*>(h));
        return set_error(LBM_ERR_HIP, "initialising the mask state failed");
    }
    *out = reinterpret_cast<lbm_handle*>(h);
    return LBM_OK;
}

int lbm_destroy(lbm_handle* hh) {
    if (!hh) return LBM_OK;
    Handle* h = H(hh);
    if (h->partial) (void)hipFree(h->partial);
    if (h->summary) (void)hipFree(h->summary);
    if (h->tile_kind) (void)hipFree(h->tile_kind);
    if (h->nbmask) (void)hipFree(h->nbmask);
    if (h->mvs) (void)hipFree(h->mvs);
    if (h->btiles) (void)hipFree(h->btiles);
    if (h->flags_seen) (void)hipFree(h->flags_seen);
    if (h->mask_changed) (void)hipFree(h->mask_changed);
    if (h->link_nodes) (void)hipFree(h->link_nodes);
    if (h->link_q) (void)hipFree(h->link_q);
    if (h->force_partial) (void)hipFree(h->force_partial);
    if (h->mlp) (void)hipFree(h->mlp);
    delete h;
    return LBM_OK;
}

int lbm_step_db(lbm_handle* hh, void* f_out, const void* f_in, const void* bc, const uint8_t* flags, double tau,
                void* stream) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (h.d.method != LBM_DOUBLE_BUFFER) return set_error(LBM_ERR_INVALID, "handle is not configured for double buffer");
    if (!f_out || !f_in || !flags || !bc) return set_error(LBM_ERR_INVALID, "NULL buffer");
    if (f_out == f_in) return set_error(LBM_ERR_INVALID, "f_out and f_in must be different buffers");
    if (!(tau > 0.5)) return set_error(LBM_ERR_INVALID, "tau must be > 0.5");
    const int rc = dispatch(h.d, [&](auto l) {
        return decltype(l)::step(h, f_out, f_in, bc, flags, tau, MODE_DB, (hipStream_t)stream);
    });
    if (rc == LBM_OK && h.n_links) h.force_valid = true;
    return rc;
}

int lbm_step_ep(lbm_handle* hh, void* f, const void* bc, const uint8_t* flags, double tau, int is_even,
                void* stream) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (h.d.method != LBM_ESOTERIC_PULL) return set_error(LBM_ERR_INVALID, "handle is not configured for esoteric pull");
    if (!f || !flags || !bc) return set_error(LBM_ERR_INVALID, "NULL buffer");
    if (!(tau > 0.5)) return set_error(LBM_ERR_INVALID, "tau must be > 0.5");
    const int mode = is_even ? MODE_EP_EVEN : MODE_EP_ODD;
    return dispatch(h.d, [&](auto l) {
        return decltype(l)::step(h, f, f, bc, flags, tau, mode, (hipStream_t)stream);
    });
}

static int macro_mode(const Handle& h, int after_odd) {
    if (h.d.method == LBM_DOUBLE_BUFFER) return MODE_DB;
    return after_odd ? MODE_EP_EVEN : MODE_EP_ODD;  // after an odd step the next step is even
}

int lbm_macroscopic(lbm_handle* hh, void* rho, void* u, const void* f, const void* bc, const uint8_t* flags,
                    int after_odd, int u_layout, void* stream) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (!rho || !u || !f || !bc || !flags) return set_error(LBM_ERR_INVALID, "NULL buffer");
    if (u_layout != LBM_U_XYZD && u_layout != LBM_U_DXYZ) return set_error(LBM_ERR_INVALID, "unknown u_layout");
    const int mode = macro_mode(h, after_odd);
    return dispatch(h.d, [&](auto l) {
        return decltype(l)::macro(h, rho, u, f, bc, flags, mode, u_layout, (hipStream_t)stream);
    });
}

int lbm_q_criterion(lbm_handle* hh, void* q, const void* u, const void* f, const void* bc, const uint8_t* flags,
                    double tau, int after_odd, int u_layout, void* stream) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (!q || !u || !f || !bc || !flags) return set_error(LBM_ERR_INVALID, "NULL buffer");
    if (u_layout != LBM_U_XYZD && u_layout != LBM_U_DXYZ) return set_error(LBM_ERR_INVALID, "unknown u_layout");
    if (!(tau > 0.0)) return set_error(LBM_ERR_INVALID, "tau must be positive");
    const int mode = macro_mode(h, after_odd);
    return dispatch(h.d, [&](auto l) {
        return decltype(l)::qcrit(h, q, u, f, bc, flags, tau, mode, u_layout, (hipStream_t)stream);
    });
}

int lbm_reduce_moments_f64(lbm_handle* hh, double* out_dev, const void* f, const void* bc, const uint8_t* flags,
                           int after_odd, void* stream) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (!out_dev || !f || !bc || !flags) return set_error(LBM_ERR_INVALID, "NULL buffer");
    const int mode = macro_mode(h, after_odd);
    return dispatch(h.d, [&](auto l) {
        return decltype(l)::reduce(h, out_dev, f, bc, flags, mode, (hipStream_t)stream);
    });
}

int lbm_init_equilibrium(lbm_handle* hh, void* f, double rho, const double* u, void* stream) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (!f) return set_error(LBM_ERR_INVALID, "NULL buffer");
    return dispatch(h.d, [&](auto l) { return decltype(l)::init(h, f, rho, u, (hipStream_t)stream); });
}

// _error_check, src/lbm/preprocess/boundary_condition.mojo:20-70.
int lbm_set_exterior_wall(lbm_handle* hh, uint8_t* flags, void* bc, const char* side, int flag, const double* u,
                          double rho, void* stream) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    const int D = h.d.D;
    if (!flags || !bc) return set_error(LBM_ERR_INVALID, "NULL buffer");
    const bool flag_ok = flag == LBM_FLAG_FLUID || flag == LBM_FLAG_SOLID ||
                         (flag >= 0 && flag <= LBM_FLAG_FLUID_SOLID_BOUNDARY && ((h.d.included_bcs >> flag) & 1u));
    if (!flag_ok) return set_error(LBM_ERR_INVALID, "Input Boundary Type is not in the config's valid boundary types");
    if (!side || std::strlen(side) != 2 || (side[0] != '-' && side[0] != '+') ||
        (side[1] != 'X' && side[1] != 'Y' && side[1] != 'Z'))
        return set_error(LBM_ERR_INVALID, "Side not valid: expects one of -X, +X, -Y, +Y, -Z, +Z");
    const int axis = side[1] - 'X';
    if (axis >= D) return set_error(LBM_ERR_INVALID, "Side not valid for a 2-D grid");
    const bool u_empty = u == nullptr;
    if (u_empty && std::isnan(rho))
        return set_error(LBM_ERR_INVALID, "Either velocity or density or both have to be specified. Both cant be left as None");
    if (!u_empty)
        for (int i = 0; i < D; ++i)
            if (std::isnan(u[i]))
                return set_error(LBM_ERR_INVALID, "either all components of velocity must be free or specified (cannot be NaN)");
    if (flag == LBM_FLAG_SOLID) {
        if (u_empty || std::isnan(rho)) return set_error(LBM_ERR_INVALID, "For Solid Type you must specify both u and rho");
        if (!h.d.include_moving_boundary)
            for (int i = 0; i < D; ++i)
                if (u[i] != 0.0)
                    return set_error(LBM_ERR_INVALID,
                                     "include_moving_boundary in config was set to False but a non zero velocity was specified");
    }
    if (flag == LBM_FLAG_EQUILIBRIUM && !u_empty && std::isnan(rho))
        return set_error(LBM_ERR_INVALID, "For Equilibrium BC, if u is specified then the density must also be specified");
    // the flags / bc change in place: a static-flags summary built from these buffers is stale
    h.summary_flags = nullptr;
    h.summary_bc = nullptr;
    const int n[3] = {h.d.nx, h.d.ny, h.d.nz};
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    const int count = n[a1] * n[a2];
    const unsigned blocks = (unsigned)((count + 255) / 256);
    const double uu[3] = {u_empty ? NAN : u[0], u_empty ? NAN : u[1], (u_empty || D < 3) ? NAN : u[2]};
    const int plus = side[0] == '+';
    if (h.d.storage == LBM_STORE_F64)
        hipLaunchKernelGGL((wall_kernel<double>), dim3(blocks), dim3(256), 0, (hipStream_t)stream, flags, (double*)bc, h.g,
                           D, axis, plus, (uint8_t)flag, uu[0], uu[1], uu[2], rho);
    else
        hipLaunchKernelGGL((wall_kernel<float>), dim3(blocks), dim3(256), 0, (hipStream_t)stream, flags, (float*)bc, h.g, D,
                           axis, plus, (uint8_t)flag, (float)uu[0], (float)uu[1], (float)uu[2], (float)rho);
    return check_launch("wall_kernel");
}

int lbm_set_links(lbm_handle* hh, const uint32_t* nodes, const float* qdist, uint32_t n_links) {
    return lbm_set_links_clamped(hh, nodes, qdist, n_links, 0.0f);
}

int lbm_set_links_clamped(lbm_handle* hh, const uint32_t* nodes, const float* qdist, uint32_t n_links,
                          float q_clamp) {
    if (int rc = check_handle(hh)) return rc;
    if (!(q_clamp >= 0.0f && q_clamp <= 1.0f)) return set_error(LBM_ERR_INVALID, "q_clamp must lie in [0, 1]");
    Handle& h = *H(hh);
    if (h.d.method != LBM_DOUBLE_BUFFER)
        return set_error(LBM_ERR_UNSUPPORTED, "interpolated bounce-back links need the double-buffer method");
    if (n_links && (!nodes || !qdist)) return set_error(LBM_ERR_INVALID, "NULL link arrays");
    const uint32_t n_cells = (uint32_t)((size_t)h.g.n_tiles * h.g.tile_cells);
    for (uint32_t i = 0; i < n_links; ++i) {
        if (nodes[i] >= n_cells) return set_error(LBM_ERR_INVALID, "link node index outside the grid");
        if (i && nodes[i] <= nodes[i - 1]) return set_error(LBM_ERR_INVALID, "link nodes must be strictly increasing");
    }
    if (hipDeviceSynchronize() != hipSuccess) return set_error(LBM_ERR_HIP, "synchronize before replacing links failed");
    if (h.link_nodes) (void)hipFree(h.link_nodes);
    if (h.link_q) (void)hipFree(h.link_q);
    h.link_nodes = nullptr;
    h.link_q = nullptr;
    h.n_links = 0;
    h.force_valid = false;
    h.node_force = nullptr;  // sized for the old records
    if (!n_links) return LBM_OK;
    // q_dist = max(q_dist, q_clamp) on every link (qd >= 0); negative entries stay plain bounce-back
    std::vector<float> qd(qdist, qdist + (size_t)n_links * (h.d.Q - 1));
    for (float& v : qd)
        if (v >= 0.0f) v = std::max(v, q_clamp);
    const size_t qbytes = qd.size() * sizeof(float);
    if (hipMalloc((void**)&h.link_nodes, (size_t)n_links * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&h.link_q, qbytes) != hipSuccess)
        return set_error(LBM_ERR_HIP, "hipMalloc of the link records failed");
    if (!h.force_partial) {
        const size_t fb = (size_t)h.g.n_tiles * ((h.g.tile_cells + 63) / 64) * h.d.D * sizeof(double);
        if (hipMalloc((void**)&h.force_partial, fb) != hipSuccess)
            return set_error(LBM_ERR_HIP, "hipMalloc of the force partials failed");
    }
    if (hipMemcpy(h.link_nodes, nodes, (size_t)n_links * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h.link_q, qd.data(), qbytes, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(LBM_ERR_HIP, "copy of the link records failed");
    h.n_links = n_links;
    return LBM_OK;
}

int lbm_set_link_node_force(lbm_handle* hh, double* node_force_dev) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (node_force_dev && !h.n_links) return set_error(LBM_ERR_INVALID, "set the links (lbm_set_links) first");
    h.node_force = node_force_dev;
    return LBM_OK;
}

int lbm_set_mlp(lbm_handle* hh, const float* params, int32_t n) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (!params || n != MLP_TOTAL) return set_error(LBM_ERR_INVALID, "the closure takes 1446 parameters");
    for (int32_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i])) return set_error(LBM_ERR_INVALID, "non-finite closure parameter");
    // the float block, then W1^T and W2^T as bf16 MFMA A operands (lane l: row l mod 32,
    // K elements 8 (l / 32) + i; for W2 the K index of step s is hidden unit mfma_row(g, 8 s + i)),
    // then each wave half's b1, b2, w3 values in result-register order
    std::vector<uint8_t> blob(MLP_DEVICE_BYTES, 0);
    std::memcpy(blob.data(), params, MLP_TOTAL * sizeof(float));
    auto bf16 = [](float v) -> uint16_t {  // round to nearest even
        uint32_t b;
        std::memcpy(&b, &v, 4);
        return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
    };
    uint16_t* a1 = (uint16_t*)(blob.data() + MLP_PACK_A1);
    uint16_t* a2 = (uint16_t*)(blob.data() + MLP_PACK_A2);
    for (int l = 0; l < 64; ++l)
        for (int i = 0; i < 8; ++i) {
            const int g = l / 32, m = l % 32, k = 8 * g + i;
            a1[l * 8 + i] = k < MLP_IN ? bf16(params[MLP_W1 + k * MLP_H + m]) : 0;
            for (int s2 = 0; s2 < 2; ++s2) {
                const int kin = mfma_row(g, 8 * s2 + i);
                a2[s2 * 512 + l * 8 + i] = bf16(params[MLP_W2 + kin * MLP_H + m]);
            }
        }
    float* rows = (float*)(blob.data() + MLP_PACK_ROWS);
    for (int g = 0; g < 2; ++g)
        for (int j = 0; j < 16; ++j) {
            rows[g * 48 + j] = params[MLP_B1 + mfma_row(g, j)];
            rows[g * 48 + 16 + j] = params[MLP_B2 + mfma_row(g, j)];
            rows[g * 48 + 32 + j] = params[MLP_W3 + mfma_row(g, j)];
        }
    if (!h.mlp && hipMalloc((void**)&h.mlp, MLP_DEVICE_BYTES) != hipSuccess)
        return set_error(LBM_ERR_HIP, "hipMalloc of the closure parameters failed");
    if (hipDeviceSynchronize() != hipSuccess) return set_error(LBM_ERR_HIP, "synchronize before replacing the closure");
    return hip_rc(hipMemcpy(h.mlp, blob.data(), MLP_DEVICE_BYTES, hipMemcpyHostToDevice), "hipMemcpy");
}

int lbm_link_force(lbm_handle* hh, double* out_dev, void* stream) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (!out_dev) return set_error(LBM_ERR_INVALID, "NULL buffer");
    if (!h.n_links || !h.force_valid) return set_error(LBM_ERR_INVALID, "no step with links has run");
    return dispatch(h.d, [&](auto l) { return decltype(l)::link_force(h, out_dev, (hipStream_t)stream); });
}

int lbm_f16c_convert(const void* in, void* out, size_t n, int to_f16c, void* stream) {
    if (!in || !out) return set_error(LBM_ERR_INVALID, "NULL buffer");
    if (n == 0) return LBM_OK;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(f16c_convert_kernel<0>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, in, out, n, to_f16c);
    return check_launch("f16c_convert_kernel");
}

int lbm_set_option(lbm_handle* hh, int option, int value) {
    if (int rc = check_handle(hh)) return rc;
    Handle& h = *H(hh);
    if (option == LBM_OPT_XCD_REMAP) {
        if (value != 0 && value != 1) return set_error(LBM_ERR_INVALID, "LBM_OPT_XCD_REMAP takes 0 or 1");
        if (value && h.g.n_tiles % 8 != 0)
            return set_error(LBM_ERR_INVALID, "the XCD remap needs a tile count divisible by 8");
        h.g.xcd_chunk = value ? h.g.n_tiles / 8 : 0;
        return LBM_OK;
    }
    if (option == LBM_OPT_STATIC_FLAGS) {
        if (value != 0 && value != 1) return set_error(LBM_ERR_INVALID, "LBM_OPT_STATIC_FLAGS takes 0 or 1");
        h.static_flags = value != 0;
        h.summary_flags = nullptr;  // the next step rebuilds the summary
        h.summary_bc = nullptr;
        return LBM_OK;
    }
    if (option == LBM_OPT_STEP_LANES) {
        const int wzb = step_slab_bits(h.g, h.d.D, value == 0 ? h.g.tile_cells : value);
        if (wzb < 0)
            return set_error(LBM_ERR_INVALID,
                             "LBM_OPT_STEP_LANES takes 0 or a power of two >= 64 that is a whole number of the "
                             "tile's z layers, at most the tile's nodes, with a slab halo of at most 8 nodes per lane");
        h.g.wzb = wzb;
        return LBM_OK;
    }
    if (option == LBM_OPT_STAGED_PULL) {
        if (value < 0 || value > 2) return set_error(LBM_ERR_INVALID, "LBM_OPT_STAGED_PULL takes 0, 1 or 2");
        h.staged_pull = value;
        return LBM_OK;
    }
    if (option == LBM_OPT_ADDR64) {
        if (value != 0 && value != 1) return set_error(LBM_ERR_INVALID, "LBM_OPT_ADDR64 takes 0 or 1");
        h.big = value != 0 || h.f_bytes >= (1ull << 31);
        return LBM_OK;
    }
    return set_error(LBM_ERR_INVALID, "unknown option");
}

int lbm_alloc(size_t bytes, void** ptr) {
    if (!ptr) return set_error(LBM_ERR_INVALID, "ptr is NULL");
    return hip_rc(hipMalloc(ptr, bytes), "hipMalloc");
}
int lbm_free(void* ptr) { return hip_rc(hipFree(ptr), "hipFree"); }
int lbm_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    return hip_rc(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "hipMemcpyAsync");
}
int lbm_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    return hip_rc(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream), "hipMemcpyAsync");
}
int lbm_memset(void* dst, int value, size_t bytes, void* stream) {
    return hip_rc(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream), "hipMemsetAsync");
}
int lbm_synchronize(void* stream) {
    return hip_rc(stream ? hipStreamSynchronize((hipStream_t)stream) : hipDeviceSynchronize(), "synchronize");
}

}  // extern "C"

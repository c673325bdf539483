// lbm_kernels.hpp -- gfx950 stream-collide kernels.
//
// One workgroup per layout tile, one lane per node; lane id == the node's
// column-major index inside the tile, so every population plane of a tile is
// read and written as contiguous 64-lane runs.
//
// Pre-passes, run once while the flags stay the same (or before every step): tile_summary_kernel
// ORs the flags of every tile, node_mask_kernel marks each tile interior or boundary and gives
// every node of a boundary tile a 32-bit neighbour mask. A tile whose own and 26 neighbouring
// tiles hold only FLUID nodes is an interior tile: it pulls every population from x - c_q and
// collides, with no flag traffic. Other tiles load one mask per node beside the populations and
// apply bounce-back, moving walls and equilibrium nodes (no LDS, no barrier in the step).
//
// Populations are addressed through buffer resources: each lane carries a
// 32-bit byte offset and the q-plane offset q * tile_cells rides in the scalar
// soffset operand, so a load or store costs at most one vector add. Buffers of
// 2 GiB or more use a 64-bit pointer variant (BIG).
#pragma once

#include <type_traits>

#include "lbm_physics.hpp"

namespace lbm {

enum Mode { MODE_DB = 0, MODE_EP_EVEN = 1, MODE_EP_ODD = 2 };

template <class R> struct StepParams {
    const void* f_in;          // typed by the storage
    void* f_out;               // == f_in for esoteric pull
    const R* bc;
    const uint8_t* flags;
    const uint32_t* tile_kind; // per tile: 1 for a boundary tile (node_mask_kernel)
    const uint32_t* nbmask;    // per node of a boundary tile: its neighbour mask (node_mask_kernel)
    R tau;
    R inv_tau;                 // 1/tau rounded in R
    R cs;
    int moving;                // include_moving_boundary
    int eqbc;                  // EQUILIBRIUM in the included BCs
    int capture;               // LBM_FLAG_CAPTURE_DENSITY / _VELOCITY bits enabled (the step writes bc)
    uint32_t f_bytes;          // bytes of one f buffer (32-bit addressing only)
    const uint32_t* link_nodes;// interpolated bounce-back records (sorted flag-layout node indices), or NULL
    const float* link_q;       // (Q-1) link distances per record
    uint32_t n_links;
    double* force_partial;     // per wave: D doubles of momentum-exchange force (when link_nodes != NULL)
    double* node_force;        // per link record: D doubles of the node's force this step, or NULL
    const float* mlp;          // learned-closure parameter block (LES_MLP)
    Geo g;
};

template <class R> struct MacroParams {
    const void* f;
    const R* bc;
    const uint8_t* flags;
    R* rho;
    R* u;
    double* partial;           // reduction mode: per-tile (D+1) sums
    const uint32_t* tile_kind; // per tile: 1 for a boundary tile (interior tiles skip the flags), or NULL
    const uint32_t* nbmask;    // per node of a boundary tile: the step's neighbour mask (with tile_kind)
    const uint32_t* link_nodes;  // interpolated bounce-back records (double buffer), or null
    const float* link_q;
    uint32_t n_links;
    int u_layout;
    int moving;
    int eqbc;
    uint32_t f_bytes;
    const R* vel;              // Q-criterion: velocity field in u_layout
    R* qout;                   // Q-criterion output (row-major x, y, z)
    R tau;                     // Q-criterion strain-rate tau
    uint32_t su_offset;        // Q-criterion: byte offset of the LDS velocity tile
    uint32_t stage_offset;     // byte offset of the LDS output staging buffer (fields, Q)
    Geo g;
};

// What macro_kernel writes.
enum MacroKind { MK_FIELDS = 0, MK_REDUCE = 1, MK_QCRIT = 2 };

// ------------------------------------------------------------ tile region masks
// Bit (sx+1) + 3 (sy+1) + 9 (sz+1), s in {-1, 0, 1}, of a tile's mask is set
// when a non-FLUID node lies in region (sx, sy, sz): s_d = 0 is the whole
// extent along axis d, s_d = -1 / +1 the first / last layer. The halo a tile
// reads from its neighbour at offset (dx, dy, dz) is exactly that neighbour's
// region (-dx, -dy, -dz); the tile's own nodes are region (0, 0, 0).
__host__ __device__ constexpr uint32_t region_bit(int sx, int sy, int sz) {
    return 1u << ((sx + 1) + 3 * (sy + 1) + 9 * (sz + 1));
}
// All regions with s_d == s along axis d.
__host__ __device__ constexpr uint32_t axis_regions(int d, int s) {
    uint32_t m = 0;
    for (int sz = -1; sz <= 1; ++sz)
        for (int sy = -1; sy <= 1; ++sy)
            for (int sx = -1; sx <= 1; ++sx)
                if ((d == 0 ? sx : d == 1 ? sy : sz) == s) m |= region_bit(sx, sy, sz);
    return m;
}

// ------------------------------------------------------------ neighbour masks
// Per-node mask of the nodes of boundary tiles, built from the flags by node_mask_kernel (once
// while the flags stay the same): bit q (1 <= q < Q) is set when the pull neighbour x - c_q is
// SOLID (raw flag compare, ops/load_and_store.mojo:216); the upper bits describe the node. A
// mask of 0 is a FLUID node with no SOLID neighbour: nothing to do beyond the plain update.
constexpr uint32_t NB_INVALID = 1u << 0;  // SOLID or EXCLUDE node (is_valid_thread false)
constexpr uint32_t NB_EQ = 1u << 27;      // EQUILIBRIUM node
constexpr uint32_t NB_FSB = 1u << 28;     // FLUID_SOLID_BOUNDARY node (interpolated links)
constexpr uint32_t NB_DECOR = 1u << 29;   // a pull neighbour is SOLID with more flag bits set
constexpr uint32_t NB_MOVING = 1u << 30;  // a SOLID pull neighbour is not a plain resting wall
constexpr uint32_t NB_CAPTURE = 1u << 31; // the node's flag carries CAPTURE_DENSITY / CAPTURE_VELOCITY

// ------------------------------------------------------------ population access
// Cache policy of population stores: bit 1 (nt) streams them past the caches;
// a step never rereads what it writes.
constexpr int kStoreNT = 2;

// Raw bits <-> compute value for each storage.
template <class S> struct Bits;
template <> struct Bits<StoreF32> {
    __device__ __forceinline__ static float get(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
    }
    __device__ __forceinline__ static void put(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, float x) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, v, s, kStoreNT);
    }
};
template <> struct Bits<StoreF16C> {
    __device__ __forceinline__ static float get(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return f16c_decode(__builtin_amdgcn_raw_buffer_load_b16(r, v, s, 0));
    }
    // The code sign-extended (buffer_load_sshort once the value passes a register fence).
    __device__ __forceinline__ static int32_t raw(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return (int16_t)__builtin_amdgcn_raw_buffer_load_b16(r, v, s, 0);
    }
    // The encoder leaves the code in a register's upper half; the buffer-store builtin takes
    // only a lower half (a shift per population), so the d16_hi store is written out. The
    // compiler does not see the store's reads of its scalar operands, which it may have just
    // reloaded with a VALU v_readlane (a spilled SGPR): a VALU write of an SGPR needs five wait
    // states before a vector memory instruction reads it, so the store waits them out itself
    // (the wave stalls; other waves issue meanwhile).
    __device__ __forceinline__ static void put(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, float x) {
        __asm__ volatile("s_nop 4\n\tbuffer_store_short_d16_hi %0, %1, %2, %3 offen nt"
                         :
                         : "v"(f16c_encode_hi(x)), "v"(v), "s"(r), "s"(s));
    }
};
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
template <> struct Bits<StoreF64> {
    __device__ __forceinline__ static double get(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
    }
    __device__ __forceinline__ static void put(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s, double x) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, v, s, kStoreNT);
    }
};

// One population field. ld/st take the lane's element index (node part) and a
// uniform element offset (the q plane).
// An empty asm that takes every register of r as input and output: values loaded before it
// reach it as loaded, so the compiler cannot fold a sign extension into the decode (which
// leaves the extension to the load), and every load is issued before the first decode.
#define LBM_RF(i) "+v"(r[i])
template <int Q> __device__ __forceinline__ void register_fence(int32_t (&r)[Q]);
template <> __device__ __forceinline__ void register_fence<9>(int32_t (&r)[9]) {
    __asm__("" : LBM_RF(0), LBM_RF(1), LBM_RF(2), LBM_RF(3), LBM_RF(4), LBM_RF(5), LBM_RF(6), LBM_RF(7), LBM_RF(8));
}
template <> __device__ __forceinline__ void register_fence<19>(int32_t (&r)[19]) {
    __asm__("" : LBM_RF(0), LBM_RF(1), LBM_RF(2), LBM_RF(3), LBM_RF(4), LBM_RF(5), LBM_RF(6), LBM_RF(7), LBM_RF(8),
            LBM_RF(9), LBM_RF(10), LBM_RF(11), LBM_RF(12), LBM_RF(13), LBM_RF(14), LBM_RF(15), LBM_RF(16),
            LBM_RF(17), LBM_RF(18));
}
template <> __device__ __forceinline__ void register_fence<27>(int32_t (&r)[27]) {
    __asm__("" : LBM_RF(0), LBM_RF(1), LBM_RF(2), LBM_RF(3), LBM_RF(4), LBM_RF(5), LBM_RF(6), LBM_RF(7), LBM_RF(8),
            LBM_RF(9), LBM_RF(10), LBM_RF(11), LBM_RF(12), LBM_RF(13), LBM_RF(14), LBM_RF(15), LBM_RF(16),
            LBM_RF(17), LBM_RF(18), LBM_RF(19), LBM_RF(20), LBM_RF(21), LBM_RF(22), LBM_RF(23), LBM_RF(24),
            LBM_RF(25), LBM_RF(26));
}
#undef LBM_RF

// Element index type of f and bc: 32 bits below 2 GiB per buffer, 64 bits above (BIG), so a
// buffer may hold more than 2^32 values.
template <bool BIG> using Idx = typename std::conditional<BIG, uint64_t, uint32_t>::type;

template <class S, bool BIG> struct Field {
    using T = typename S::T;
    using R = typename S::R;
    using I = Idx<BIG>;
    T* ptr;
    __amdgpu_buffer_rsrc_t rsrc;
    __device__ __forceinline__ Field(const void* p, uint32_t bytes) : ptr((T*)p) {
        if (!BIG) rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, (int)bytes, 0x00020000);
    }
    __device__ __forceinline__ R ld(I lane, uint32_t uni) const {
        if (BIG) return S::ld(ptr + ((uint64_t)lane + uni));
        return Bits<S>::get(rsrc, (uint32_t)lane * (uint32_t)sizeof(T), uni * (uint32_t)sizeof(T));
    }
    // Float16C, 32-bit addressing: the sign-extended code (decoded by f16c_decode_sext).
    __device__ __forceinline__ int32_t ld_code(I lane, uint32_t uni) const {
        return Bits<S>::raw(rsrc, (uint32_t)lane * (uint32_t)sizeof(T), uni * (uint32_t)sizeof(T));
    }
    __device__ __forceinline__ void st(I lane, uint32_t uni, R v) const {
        if (BIG) S::st(ptr + ((uint64_t)lane + uni), v);
        else Bits<S>::put(rsrc, (uint32_t)lane * (uint32_t)sizeof(T), uni * (uint32_t)sizeof(T), v);
    }
};

// ------------------------------------------------------------ tile context
// Per-lane position and the element deltas of a one-node move along each axis
// (periodic, get_adjacent_idx src/lbm/kernels/utils/index.mojo:44-64), plus,
// for boundary tiles, the LDS flag tile with its one-node halo. Halo flags are
// fetched into registers before the population loads are issued and written
// to LDS after, so their latency overlaps the population traffic.
template <class L, class I = uint32_t> struct TileCtx {
    static constexpr int MAXR = 8;  // halo rounds per lane (hcells <= 8 * tile_cells, checked on the host)
    int x, y, z;          // node
    int tid;
    uint32_t loc;         // node index inside its tile
    int wz0, wtz;         // z offset inside the tile of the workgroup's slab and its layer count
    int hx, hxy;          // LDS halo strides
    int center;           // own position in the LDS tile (the workgroup's slab + halo)
    uint32_t tile;        // tile id
    int bx, by, bz;       // tile coordinates (uniform)
    I fbase;              // element index of (x, q = 0) in f
    I fm[3], fp[3];       // element deltas of a -1 / +1 move in f (modulo the index type)
    uint32_t tc;          // tile_cells: element distance between q planes
    int nonfluid;         // tile or halo holds a non-FLUID flag
    uint8_t own;
    uint32_t hv[MAXR];    // fetched halo flags

    // The workgroup covers z layers [part 2^lzb, (part + 1) 2^lzb) of the tile (part 0 and
    // lzb = tzb: the whole tile); lane tid is node tid + part 2^lzb tx ty of that tile.
    // lane_tid: the lane's index in the tile's slab (default: threadIdx.x; a workgroup over several
    // tiles passes its own).
    __device__ __forceinline__ void init_pos(const Geo& g, uint32_t tile_id, uint32_t part, int lzb, int lane_tid = -1) {
        constexpr int D = L::D, Q = L::Q;
        tile = (uint32_t)__builtin_amdgcn_readfirstlane(tile_id);
        tid = lane_tid < 0 ? (int)threadIdx.x : lane_tid;
        const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
        wz0 = (int)part << lzb;
        wtz = 1 << lzb;
        const int wlz = tid >> (g.txb + g.tyb);  // layer inside the slab
        const int lx = tid & (tx - 1), ly = (tid >> g.txb) & (ty - 1), lz = wz0 + wlz;
        loc = (uint32_t)tid + ((uint32_t)wz0 << (g.txb + g.tyb));
        bx = __builtin_amdgcn_readfirstlane((int)(tile % (uint32_t)g.Tx));
        const int byz = (int)(tile / (uint32_t)g.Tx);
        by = __builtin_amdgcn_readfirstlane(byz % g.Ty);
        bz = __builtin_amdgcn_readfirstlane(byz / g.Ty);
        x = (bx << g.txb) + lx;
        y = (by << g.tyb) + ly;
        z = (bz << g.tzb) + lz;
        hx = tx + 2;
        hxy = hx * (ty + 2);
        center = (wlz + (D == 3 ? 1 : 0)) * hxy + (ly + 1) * hx + (lx + 1);
        tc = (uint32_t)g.tile_cells;
        const I TQ = (I)tc * (I)Q;
        fbase = (I)tile * TQ + (I)loc;

        const int Ts[3] = {1, g.Tx, g.Tx * g.Ty};
        const int ls[3] = {1, tx, tx * ty};
        const int lc[3] = {lx, ly, lz};
        const int bc3[3] = {bx, by, bz};
        const int tn[3] = {tx, ty, tz};
        const int Tn[3] = {g.Tx, g.Ty, g.Tz};
        // A move stays inside the tile unless the lane sits on the tile's face; the face case
        // (into the neighbour tile, periodic) depends only on the tile, so it is computed once
        // per workgroup and each lane selects.
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const I m_face = (I)(int64_t)((tn[d] - 1) * ls[d]) +
                             (I)(int64_t)((bc3[d] > 0 ? -1 : Tn[d] - 1) * Ts[d]) * TQ;
            const I p_face = (I)(int64_t)(-(tn[d] - 1) * ls[d]) +
                             (I)(int64_t)((bc3[d] < Tn[d] - 1 ? 1 : -(Tn[d] - 1)) * Ts[d]) * TQ;
            fm[d] = lc[d] > 0 ? (I)(int64_t)(-ls[d]) : m_face;
            fp[d] = lc[d] < tn[d] - 1 ? (I)(int64_t)ls[d] : p_face;
        }
    }

    // Element index in f (q plane excluded) of node x + s * c_q.
    __device__ __forceinline__ I nb(int q, int s) const {
        I a = fbase;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int c = s * L::c(q, d);
            if (c > 0) a += fp[d];
            if (c < 0) a += fm[d];
        }
        return a;
    }

    // Element index in a field with K values per node of node (xn, yn, zn), value k.
    __device__ __forceinline__ I field_index(const Geo& g, int xn, int yn, int zn, int K, int k) const {
        const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
        const uint32_t loc = (uint32_t)((xn & (tx - 1)) + ((yn & (ty - 1)) << g.txb) + ((zn & (tz - 1)) << (g.txb + g.tyb)));
        const uint32_t tl = (uint32_t)((xn >> g.txb) + g.Tx * ((yn >> g.tyb) + g.Ty * (zn >> g.tzb)));
        return ((I)tl * (I)K + (I)k) * (I)tc + (I)loc;
    }

    // Element index of the pull neighbour x - c_q in a field with K values per node.
    __device__ __forceinline__ I pull_index(const Geo& g, int q, int K, int k) const {
        int xn = x - L::c(q, 0), yn = y - L::c(q, 1), zn = z - L::c(q, 2);
        xn += xn < 0 ? g.nx : 0;
        xn -= xn >= g.nx ? g.nx : 0;
        yn += yn < 0 ? g.ny : 0;
        yn -= yn >= g.ny ? g.ny : 0;
        zn += zn < 0 ? g.nz : 0;
        zn -= zn >= g.nz ? g.nz : 0;
        return field_index(g, xn, yn, zn, K, k);
    }

    // Loads the flags of the workgroup's slab and its halo into registers (no LDS traffic yet),
    // with mvs the moving-wall bytes of the same cells in bits 8-15.
    __device__ __forceinline__ void fetch_flags(const Geo& g, const uint8_t* __restrict__ flags,
                                                const uint8_t* __restrict__ mvs = nullptr) {
        constexpr int D = L::D;
        const int hcells = hxy * (D == 3 ? wtz + 2 : 1);
        const int nthr = blockDim.x;
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int h = tid + r * nthr;
            hv[r] = 0;
            if (r * nthr >= hcells) break;
            if (h < hcells) {
                const int iz = (int)__umulhi((uint32_t)h, g.hxy_magic);
                const int rr = h - iz * hxy;
                const int iy = (int)__umulhi((uint32_t)rr, g.hx_magic);
                const int ix = rr - iy * hx;
                int gx = (bx << g.txb) + ix - 1;
                gx += gx < 0 ? g.nx : 0;
                gx -= gx >= g.nx ? g.nx : 0;
                int gy = (by << g.tyb) + iy - 1;
                gy += gy < 0 ? g.ny : 0;
                gy -= gy >= g.ny ? g.ny : 0;
                int gz = 0;
                if (D == 3) {
                    gz = (bz << g.tzb) + wz0 + iz - 1;
                    gz += gz < 0 ? g.nz : 0;
                    gz -= gz >= g.nz ? g.nz : 0;
                }
                const uint32_t fi = (uint32_t)field_index(g, gx, gy, gz, 1, 0);
                hv[r] = flags[fi];
                if (mvs) hv[r] |= (uint32_t)mvs[fi] << 8;
            }
        }
    }

    // Writes the fetched flags (and, with smv, the moving-wall bits) to LDS and synchronises the
    // workgroup.
    __device__ __forceinline__ void commit_flags(const Geo& g, uint8_t* sfl, uint8_t* smv = nullptr) {
        constexpr int D = L::D;
        const int hcells = hxy * (D == 3 ? wtz + 2 : 1);
        const int nthr = blockDim.x;
        uint32_t any = 0;
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int h = tid + r * nthr;
            if (r * nthr >= hcells) break;
            if (h < hcells) {
                sfl[h] = (uint8_t)hv[r];
                if (smv) smv[h] = (uint8_t)(hv[r] >> 8);
            }
            any |= hv[r] & 0xFFu;
        }
        nonfluid = __syncthreads_or((int)any);
        own = sfl[center];
    }

    // Flag of the pull neighbour x - c_q (LDS).
    __device__ __forceinline__ uint8_t pull_flag(const uint8_t* sfl, int q) const {
        return sfl[center - L::c(q, 0) - L::c(q, 1) * hx - L::c(q, 2) * hxy];
    }
};

__device__ __forceinline__ bool node_valid(uint8_t fl) {
    return !(((fl & 0x0F) == LBM_FLAG_SOLID) || (fl & LBM_FLAG_EXCLUDE));  // is_valid_thread
}

// The node's neighbour mask (NB_* layout) from the LDS flag tile of a boundary tile.
template <class L, class I>
__device__ __forceinline__ uint32_t lds_mask(const TileCtx<L, I>& t, const uint8_t* sfl) {
    uint32_t m = node_valid(t.own) ? 0u : NB_INVALID;
    if ((t.own & 0x0F) == LBM_FLAG_EQUILIBRIUM) m |= NB_EQ;
    if ((t.own & 0x0F) == LBM_FLAG_FLUID_SOLID_BOUNDARY) m |= NB_FSB;
    if (t.own & (LBM_FLAG_CAPTURE_DENSITY | LBM_FLAG_CAPTURE_VELOCITY)) m |= NB_CAPTURE;
#pragma unroll
    for (int q = 1; q < L::Q; ++q) {
        const uint8_t nf = t.pull_flag(sfl, q);
        if (nf == LBM_FLAG_SOLID) m |= 1u << q;
        else if ((nf & 0x0F) == LBM_FLAG_SOLID) m |= NB_DECOR;
    }
    return m;
}

// 1 when a non-FLUID node lies in the tile or in the one-node halo it reads from its neighbour
// tiles (periodic): neighbour (dx, dy, dz) contributes its region (-dx, -dy, -dz). Uniform
// across the workgroup.
template <int D>
__device__ __forceinline__ uint32_t boundary_tile(const Geo& g, const uint32_t* __restrict__ summary, int bx, int by,
                                                  int bz) {
    uint32_t any = 0;
    const int zlo = D == 3 ? -1 : 0, zhi = D == 3 ? 1 : 0;
#pragma unroll
    for (int dz = zlo; dz <= zhi; ++dz) {
        int tz = bz + dz;
        tz += tz < 0 ? g.Tz : 0;
        tz -= tz >= g.Tz ? g.Tz : 0;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            int ty = by + dy;
            ty += ty < 0 ? g.Ty : 0;
            ty -= ty >= g.Ty ? g.Ty : 0;
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                int tx = bx + dx;
                tx += tx < 0 ? g.Tx : 0;
                tx -= tx >= g.Tx ? g.Tx : 0;
                const uint32_t t = tx + g.Tx * (ty + g.Ty * tz);
                any |= summary[t] & region_bit(-dx, -dy, -dz);
            }
        }
    }
    return any != 0 ? 1u : 0u;
}

// Stream, part 1 (issued before the flags are known): the pull of every
// population from x - c_q (double buffer; double_buffer_pull_load_f_vec,
// ops/load_and_store.mojo:164-221) or the esoteric slot loads
// (esoteric_pull_load_f_vec, ops/load_and_store.mojo:11-86). Every address is
// in bounds (periodic wrap), whatever the flags say.
template <class L, int MODE, class I, class LD>
__device__ __forceinline__ void for_each_pull(const TileCtx<L, I>& t, LD&& ld) {
    constexpr int Q = L::Q;
    if (MODE == MODE_DB) {
#pragma unroll
        for (int q = 0; q < Q; ++q) ld(q, t.nb(q, -1), (uint32_t)q * t.tc);
    } else {
        ld(0, t.fbase, 0u);
#pragma unroll
        for (int p = 1; p < Q; p += 2) {
            const int n = p + 1;
            if (MODE == MODE_EP_EVEN) {
                ld(p, t.fbase, (uint32_t)p * t.tc);
                ld(n, t.nb(p, +1), (uint32_t)n * t.tc);
            } else {
                ld(p, t.fbase, (uint32_t)n * t.tc);
                ld(n, t.nb(p, +1), (uint32_t)p * t.tc);
            }
        }
    }
}

template <class L, class S, int MODE, bool BIG>
__device__ __forceinline__ void load_populations(const TileCtx<L, Idx<BIG>>& t, const Field<S, BIG>& f,
                                                 typename S::R (&fv)[L::Q]) {
    constexpr int Q = L::Q;
    if constexpr (S::id == LBM_STORE_F16C && !BIG) {
        // all codes loaded sign-extended, then decoded: 3 instructions per population
        int32_t code[Q];
        for_each_pull<L, MODE>(t, [&](int q, Idx<BIG> lane, uint32_t uni) { code[q] = f.ld_code(lane, uni); });
        register_fence<Q>(code);
#pragma unroll
        for (int q = 0; q < Q; ++q) fv[q] = f16c_decode_sext(code[q]);
    } else {
        for_each_pull<L, MODE>(t, [&](int q, Idx<BIG> lane, uint32_t uni) { fv[q] = f.ld(lane, uni); });
    }
}

// Staged pull (Float16C double buffer, 3-D tiles 8 x 8 nodes wide, 256-lane workgroups of four
// z layers). A 64-lane wave that pulls its 8 x 8 layer straight from memory touches the tile's
// own 128-byte line and, for a direction with an x or y component, the line of the neighbour
// tile as well, 2 bytes per lane: on the 256^3 cube the population loads then run at about
// 80 % of the rate of aligned line loads. Here the workgroup first copies each q plane's source
// rows into LDS, 16 bytes (one 8-node row) per lane, straight from memory to LDS, and each lane
// then reads its pull from LDS.
//
// Plane q's box: for every layer and y row of the slab moved by -c_q (the pull sources, from the
// neighbour tile where that leaves the tile), the 16-byte row of each tile column the pulls reach
// (one for c_qx = 0; for c_qx = 1 also the column below, for -1 the one above). For c_qz != 0 the
// box has one more layer after those, the slab's own layer the move leaves out, staged only in
// boundary tiles (interior tiles skip chunks that hold nothing else): there the halfway bounce-back
// reads the node's own population of plane q from the box whenever its row is in it (all but the
// edge row for c_qy != 0) instead of from memory. Rows in (layer, tile column, row) order, planes
// one after the other.
template <class L, int BTX = 1> struct StagedBox {
    static constexpr int SZ = 4;  // slab layers
    static constexpr int ntx(int q) { return BTX + (L::c(q, 0) != 0); }
    static constexpr int nl(int q) { return SZ + (L::c(q, 2) != 0); }           // layers
    static constexpr int rows(int q) { return nl(q) * ntx(q) * 8; }             // 16-byte rows
    static constexpr int pull_rows(int q) { return SZ * ntx(q) * 8; }           // rows of the pulls
    // source layer (relative to the slab) of box layer l: the pulls' layers l - c_qz, then the
    // slab's own layer they leave out
    static constexpr int src_layer(int q, int l) {
        return l < SZ ? l - L::c(q, 2) : (L::c(q, 2) == 1 ? SZ - 1 : 0);
    }
    static constexpr int offset(int q) {  // bytes
        int s = 0;
        for (int p = 0; p < q; ++p) s += rows(p) * 16;
        return s;
    }
    static constexpr int bytes() { return offset(L::Q); }
};

extern __shared__ __attribute__((aligned(16))) uint8_t staged_box[];

// The loads of plane q's box, 64 rows per wave-sized chunk: chunk i of plane q is taken by wave
// (first chunk of q + i) mod (waves of the workgroup); lanes past the box's rows, and in interior
// tiles lanes of rows only the bounce-back reads, are idle. The direct-to-LDS load writes lane j's
// 16 bytes at (chunk address + 16 j), so row order in LDS is lane order. BTX = 2: the workgroup
// covers two x-adjacent tiles (waves 0-3 the first, 4-7 the second), whose boxes share their rows.
template <class L, int BTX, class TC, class FV>
__device__ __forceinline__ void stage_pull(const TC& t, const Geo& g, const void* f_in, uint32_t f_bytes, bool boundary,
                                           FV& fv) {
    using B = StagedBox<L, BTX>;
    constexpr int Q = L::Q, NW = 4 * BTX, SZ = B::SZ;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    const int k = BTX == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 8);  // tile of the wave
    const int bxa = t.bx - k;  // first tile column of the workgroup
    const uint16_t* f = (const uint16_t*)f_in;
    const uint32_t TQ = t.tc * (uint32_t)Q;
    const int tzm = (1 << g.tzb) - 1;
    // tile rows below / above (periodic)
    const int by_m = t.by > 0 ? t.by - 1 : g.Ty - 1, by_p = t.by < g.Ty - 1 ? t.by + 1 : 0;
    int chunk = 0;  // chunks of the planes before q
    if constexpr (BTX == 1) {
        // One tile per workgroup: a row's source tile is the tile itself or a neighbour (periodic)
        // in x, y and z, whose element bases are uniform; each lane picks among them by its row
        // (r = lane % 8), its layer and its tile column (b = lane / 8 gives both), and the plane's
        // offset goes in the scalar offset of a buffer load straight to LDS.
        const int r = lane & 7, b = lane >> 3;
        const uint32_t TxTQ = (uint32_t)g.Tx * TQ, TxyTQ = (uint32_t)g.Ty * TxTQ;
        const int bx_m = t.bx > 0 ? t.bx - 1 : g.Tx - 1, bx_p = t.bx < g.Tx - 1 ? t.bx + 1 : 0;
        const int bz_m = t.bz > 0 ? t.bz - 1 : g.Tz - 1, bz_p = t.bz < g.Tz - 1 ? t.bz + 1 : 0;
        // (read as wave-uniform values, so that a lane's choice selects among the products instead
        // of multiplying a selected factor)
        auto uni = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); };
        const uint32_t Y[3] = {uni((uint32_t)by_m * TxTQ), uni((uint32_t)t.by * TxTQ), uni((uint32_t)by_p * TxTQ)};
        const uint32_t Z[3] = {uni((uint32_t)bz_m * TxyTQ), uni((uint32_t)t.bz * TxyTQ), uni((uint32_t)bz_p * TxyTQ)};
        const uint32_t X[3] = {uni((uint32_t)bx_m * TQ), uni((uint32_t)t.bx * TQ), uni((uint32_t)bx_p * TQ)};
        uint32_t yo[3];  // source row r - c_y (c_y = -1, 0, 1): its tile row and row in the tile
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int sy = r - (d - 1);
            yo[d] = (sy < 0 ? Y[0] : sy > 7 ? Y[2] : Y[1]) + ((uint32_t)(sy & 7) << 3);
        }
        // tile column of a two-column box (c_x = -1: columns bx, bx + 1; c_x = 1: bx - 1, bx)
        const bool second = (b & 1) != 0;
        const uint32_t xo_neg = second ? X[2] : X[1], xo_pos = second ? X[1] : X[0];
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)f_in, 0, (int)f_bytes, 0x00020000);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int cx = L::c(q, 0), cy = L::c(q, 1), cz = L::c(q, 2);
            const int nch = (B::rows(q) + 63) >> 6;
#pragma unroll
            for (int i = 0; i < nch; ++i, ++chunk) {
                if (chunk % NW != wave) continue;                        // wave-uniform
                if (!boundary && i * 64 >= B::pull_rows(q)) continue;    // own-layer rows only
                const int j = i * 64 + lane;
                const int l = B::ntx(q) == 1 ? 8 * i + b : 4 * i + (b >> 1);
                if (j < B::rows(q) && (boundary || l < SZ)) {
                    const int sl = l < SZ ? l - cz : (cz == 1 ? SZ - 1 : 0);  // source layer, -1 .. SZ
                    const int zloc = t.wz0 + sl;                             // in the tile: -1 .. tz
                    const uint32_t zo = (zloc < 0 ? Z[0] : zloc > tzm ? Z[2] : Z[1]) + ((uint32_t)(zloc & tzm) << 6);
                    const uint32_t xo = cx == 0 ? X[1] : (cx == 1 ? xo_pos : xo_neg);
                    const uint32_t e = xo + yo[cy + 1] + zo;
#if defined(__HIP_DEVICE_COMPILE__)  // a gfx950 builtin: absent from the host pass
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        rsrc, (__attribute__((address_space(3))) void*)(staged_box + B::offset(q) + i * 1024), 16,
                        (int)(e << 1), (int)((uint32_t)q * t.tc * 2u), 0, 0);
#endif
                }
            }
        }
    } else
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int cx = L::c(q, 0), cy = L::c(q, 1), cz = L::c(q, 2);
        // first tile column of the box (periodic) and the global z of the slab's first layer
        int bx0 = bxa - (cx == 1 ? 1 : 0);
        bx0 += bx0 < 0 ? g.Tx : 0;
        const int zs = (t.bz << g.tzb) + t.wz0;
        const int nch = (B::rows(q) + 63) >> 6;
#pragma unroll
        for (int i = 0; i < nch; ++i, ++chunk) {
            if (chunk % NW != wave) continue;                        // wave-uniform
            if (!boundary && i * 64 >= B::pull_rows(q)) continue;    // own-layer rows only
            const int j = i * 64 + lane;
            const int r = j & 7, blk = j >> 3;
            const int tx = blk % B::ntx(q), l = blk / B::ntx(q);
            const int sy = r - cy;  // source row, -1 .. 8
            const int sl = l < SZ ? l - cz : (cz == 1 ? SZ - 1 : 0);  // source layer in the slab, -1 .. SZ
            if (j < B::rows(q) && (boundary || l < SZ)) {
                const int yt = sy < 0 ? by_m : (sy > 7 ? by_p : t.by);
                int xt = bx0 + tx;
                xt -= xt >= g.Tx ? g.Tx : 0;
                int zl = zs + sl;
                zl += zl < 0 ? g.nz : 0;
                zl -= zl >= g.nz ? g.nz : 0;
                const uint32_t tile = ((uint32_t)(zl >> g.tzb) * (uint32_t)g.Ty + (uint32_t)yt) * (uint32_t)g.Tx + (uint32_t)xt;
                const uint32_t e = tile * TQ + (uint32_t)q * t.tc + ((uint32_t)(zl & tzm) << 6) + ((uint32_t)(sy & 7) << 3);
#if defined(__HIP_DEVICE_COMPILE__)  // a gfx950 builtin: absent from the host pass
                __builtin_amdgcn_global_load_lds(f + e, staged_box + B::offset(q) + i * 1024, 16, 0, 0);
#endif
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int lx = t.tid & 7, ly = (t.tid >> 3) & 7, lz = t.tid >> 6;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int cx = L::c(q, 0), cy = L::c(q, 1), cz = L::c(q, 2);
        const int xb = 8 * k + lx - cx + (cx == 1 ? 8 : 0);
        const int row = (lz * B::ntx(q) + (xb >> 3)) * 8 + ly;
        const int16_t code = *(const int16_t*)(staged_box + B::offset(q) + ((row * 8 + (xb & 7)) << 1));
        fv[q] = f16c_decode_sext(code);
    }
}

// Halfway bounce-back of a staged boundary tile: the node's own population of plane opp(q) comes
// from the box of that plane when the node's row is in it (its layer always is), else from memory.
template <class L, int BTX, class TC, class F, class FV>
__device__ __forceinline__ void staged_bounce_back(const TC& t, uint32_t m, const F& f, FV& fv) {
    using B = StagedBox<L, BTX>;
    const int k = BTX == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 8);
    const int lx = t.tid & 7, ly = (t.tid >> 3) & 7, lz = t.tid >> 6;
#pragma unroll
    for (int q = 1; q < L::Q; ++q) {
        if ((m >> q) & 1u) {
            const int p = opp(q);
            const int ro = ly + L::c(p, 1);  // box row of the node's own row
            if (ro >= 0 && ro < 8) {
                const int col = k + (L::c(p, 0) == 1 ? 1 : 0);
                const int lo = lz + L::c(p, 2);  // box layer of the node's own layer
                const int row = ((lo >= 0 && lo < B::SZ ? lo : B::SZ) * B::ntx(p) + col) * 8 + ro;
                fv[q] = f16c_decode_sext(*(const int16_t*)(staged_box + B::offset(p) + ((row * 8 + lx) << 1)));
            } else {
                fv[q] = f.ld(t.fbase, (uint32_t)p * t.tc);
            }
        }
    }
}

// Store: double buffer writes every population at x; esoteric_pull_store_f_vec
// (ops/load_and_store.mojo:89-159) writes the parity's slot pattern.
template <class L, class S, int MODE, bool BIG>
__device__ __forceinline__ void store_populations(const TileCtx<L, Idx<BIG>>& t, const Field<S, BIG>& f,
                                                  const typename S::R (&fv)[L::Q]) {
    constexpr int Q = L::Q;
    if (MODE == MODE_DB) {
#pragma unroll
        for (int q = 0; q < Q; ++q) f.st(t.fbase, (uint32_t)q * t.tc, fv[q]);
    } else {
        f.st(t.fbase, 0, fv[0]);
#pragma unroll
        for (int p = 1; p < Q; p += 2) {
            const int n = p + 1;
            if (MODE == MODE_EP_EVEN) {
                f.st(t.nb(p, +1), (uint32_t)n * t.tc, fv[p]);
                f.st(t.fbase, (uint32_t)p * t.tc, fv[n]);
            } else {
                f.st(t.nb(p, +1), (uint32_t)p * t.tc, fv[p]);
                f.st(t.fbase, (uint32_t)n * t.tc, fv[n]);
            }
        }
    }
}

// Stream, part 2 (boundary tiles) is halfway_bounce_back, then bouzidi_links on
// nodes with interpolated links, then apply_boundaries: moving_wall_bc
// (ops/boundary_condition.mojo:17-83) and equilibrium_bc (:87-204).
//
// Halfway bounce-back of the double-buffer pull: the pull from a SOLID
// neighbour (raw flag compare, ops/load_and_store.mojo:216) is replaced by
// f(x, opp q), loaded only for those directions.
template <class L, class S, bool BIG>
__device__ __forceinline__ void halfway_bounce_back(const TileCtx<L, Idx<BIG>>& t, uint32_t m, const Field<S, BIG>& f,
                                                    typename S::R (&fv)[L::Q]) {
#pragma unroll
    for (int q = 1; q < L::Q; ++q)
        if ((m >> q) & 1u) fv[q] = f.ld(t.fbase, (uint32_t)opp(q) * t.tc);
}

// Interpolated bounce-back (linear Bouzidi, Firdaouss & Lallemand 2001; the formulas of
// linkwise_bounceback_kernel, src/lbm/geometry/interpolated_BB.mojo:119-140, and
// nodewise_bounceback_kernel, src/lbm/geometry/nodewise_interpolated_BB.mojo:140-178) on a
// FLUID_SOLID_BOUNDARY node that has a link record: for each direction whose pull neighbour is
// SOLID and whose link distance qd >= 0 (already raised to the handle's q_clamp), the population
// entering along c_q is rebuilt from a = fv[q] = f*(x, opp q) (delivered by halfway bounce-back) and
//   qd <= 1/2: b = fv[opp q],  f_q = (2 qd) a + (1 - 2 qd) b,
//   qd >  1/2: b = f*(x, q),   f_q = (0.5 / qd) a + ((2 qd - 1) / (2 qd)) b.
// fv[opp q] was pulled from x + c_q: f*(x + c_q, opp q) in the fluid, its halfway bounce-back in a
// one-node gap, or the rebuilt value of the opposite link when that link came first. Links run in
// the reference's order (its link bit q_ref rebuilds f_vec[opp q_ref]): opp(1), 1, opp(3), 3, ...
// The link's momentum exchange c_{opp q} (a + f_q) (+ 2 w_q with DDF) is added to force, and the
// mass the interpolation adds, dm = sum (f_q - a), is taken from the rest population so the node
// holds the mass halfway bounce-back gives it (the mass fix of nodewise_interpolated_BB.mojo:170-176,
// which adds dm instead; subtracting it is what conserves mass). Returns the rebuilt directions
// (they take no moving-wall term).
template <class L, class S, bool DDF, bool BIG>
__device__ __forceinline__ uint32_t bouzidi_links(const TileCtx<L, Idx<BIG>>& t, uint32_t m, const Field<S, BIG>& f,
                                                  const uint32_t* __restrict__ link_nodes,
                                                  const float* __restrict__ link_q, uint32_t n_links, uint32_t node,
                                                  typename S::R (&fv)[L::Q], double (&force)[3],
                                                  double* __restrict__ node_force = nullptr) {
    using R = typename S::R;
    constexpr int Q = L::Q, D = L::D;
    uint32_t lo = 0, hi = n_links;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (link_nodes[mid] < node) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= n_links || link_nodes[lo] != node) return 0;
    uint32_t done = 0;
    R dm = 0;  // mass the interpolation adds, taken back from the rest population
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        const int q = opp(k);  // opp(1), 1, opp(3), 3, ...
        const R qd = (R)link_q[(size_t)lo * (Q - 1) + (q - 1)];
        if (!(qd >= (R)0) || !((m >> q) & 1u)) continue;
        const R a = fv[q];
        R fb;
        if (qd <= (R)0.5) {
            const R b = fv[opp(q)];
            fb = (R)2 * qd * a + ((R)1 - (R)2 * qd) * b;
        } else {
            const R b = f.ld(t.fbase, (uint32_t)q * t.tc);
            fb = (R)0.5 / qd * a + ((R)2 * qd - (R)1) / ((R)2 * qd) * b;
        }
        fv[q] = fb;
        dm = dm + (fb - a);
        done |= 1u << q;
        R s = a + fb;
        if (DDF) s = s + (R)2 * weight<L, R>(q);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int c = L::c(q, d);
            if (c != 0) force[d] += (double)(c > 0 ? -s : s);
        }
    }
    if (done) fv[0] = fv[0] - dm;
    if (node_force) {
        // this node's momentum exchange (force starts at zero for the node): the per-node force of
        // calculate_drag_around_object, src/lbm/output/drag.mojo:104-105
#pragma unroll
        for (int d = 0; d < D; ++d) node_force[(size_t)lo * D + d] = force[d];
    }
    return done;
}

// GROUPED: the moving-wall bc loads of six directions are issued together (one wait per group
// instead of one per direction); measured 1.7 % faster on the 256^3 FP32 cube, but the arrays it
// keeps live raise the step's VGPRs (38 -> 96), which costs the Float16C step 6 % (fewer waves)
// and would spill the learned-closure step, so only the FP32 steps without it group the loads.
template <class L, class S, int MODE, bool DDF, bool BIG, bool GROUPED = false>
__device__ __forceinline__ void apply_boundaries(const TileCtx<L, Idx<BIG>>& t, const Geo& g, uint32_t m,
                                                 const uint8_t* __restrict__ flags, const Field<S, BIG>& f,
                                                 const typename S::R* __restrict__ bc, int moving, int eqbc,
                                                 typename S::R (&fv)[L::Q], uint32_t rebuilt = 0) {
    using R = typename S::R;
    constexpr int Q = L::Q, D = L::D;
    if (moving && (m & NB_MOVING)) {
        if constexpr (GROUPED) {
            // SOLID neighbours by their low flag bits: the mask bits, or (rarely: SOLID with more
            // flag bits set somewhere around the node) the neighbours' flags themselves
            uint32_t solid = m & ((2u << (Q - 1)) - 2u);
            if (m & NB_DECOR) {
#pragma unroll
                for (int q = 1; q < Q; ++q)
                    if ((flags[t.pull_index(g, q, 1, 0)] & 0x0F) == LBM_FLAG_SOLID) solid |= 1u << q;
            }
            solid &= ~rebuilt;
            // the bc a direction needs is loaded by the whole wave when any lane needs it (a
            // wave-uniform branch, always in bounds), six directions at a time, and waited for
            // together (a load inside the lane's own branch is waited for in that branch)
            constexpr int G = 6;
#pragma unroll
            for (int q0 = 1; q0 < Q; q0 += G) {
                R uw[G][3], rw[G];
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const int q = q0 + k;
                    if (q < Q && __ballot((solid >> q) & 1u) != 0) {
#pragma unroll
                        for (int d = 0; d < D; ++d)
                            if (L::c(q, d) != 0) uw[k][d] = bc[t.pull_index(g, q, D + 1, d)];
                        rw[k] = bc[t.pull_index(g, q, D + 1, D)];
                    }
                }
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const int q = q0 + k;
                    if (q < Q && ((solid >> q) & 1u))
                        fv[q] = fv[q] + (R)6 * weight<L, R>(q) * rw[k] * dot_c<L, R>(uw[k], q);
                }
            }
        } else {
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                // a SOLID neighbour by its low flag bits: the mask bit, or (rarely: SOLID with more
                // flag bits set somewhere around the node) the neighbour's flag itself
                bool solid = (m >> q) & 1u;
                if (!solid && (m & NB_DECOR)) solid = (flags[t.pull_index(g, q, 1, 0)] & 0x0F) == LBM_FLAG_SOLID;
                if (solid && !((rebuilt >> q) & 1u)) {
                    R uw[3];
#pragma unroll
                    for (int d = 0; d < D; ++d) uw[d] = bc[t.pull_index(g, q, D + 1, d)];
                    const R rw = bc[t.pull_index(g, q, D + 1, D)];
                    fv[q] = fv[q] + (R)6 * weight<L, R>(q) * rw * dot_c<L, R>(uw, q);
                }
            }
        }
    }
    if (eqbc && (m & NB_EQ)) {
        R ub[3];
#pragma unroll
        for (int d = 0; d < D; ++d) ub[d] = bc[t.field_index(g, t.x, t.y, t.z, D + 1, d)];
        const R rb = bc[t.field_index(g, t.x, t.y, t.z, D + 1, D)];
        const int pos[3] = {t.x, t.y, t.z};
        const int nn[3] = {g.nx, g.ny, g.nz};
        R rho = 0, vel[3] = {0, 0, 0};
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            bool oob = false;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int pi = pos[d] - L::c(q, d);
                oob = oob || pi < 0 || pi >= nn[d];
            }
            const R fq = oob ? (DDF ? (R)0 : weight<L, R>(q)) : fv[q];
            rho = rho + fq;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int c = L::c(q, d);
                if (c != 0) vel[d] = vel[d] + (c > 0 ? fq : -fq);
            }
        }
        if (DDF) rho = rho + (R)1;
#pragma unroll
        for (int d = 0; d < D; ++d) vel[d] = vel[d] / rho;
        const bool ufree = isnan(ub[0]);
        R ul[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) ul[d] = ufree ? vel[d] : (d < D ? ub[d] : (R)0);
        const R rl = isnan(rb) ? rho : rb;
        const R uu = dot_uu<L, R>(ul);
#pragma unroll
        for (int q = 0; q < Q; ++q) fv[q] = feq<L, R, DDF>(q, rl, ul, uu);
    }
}

// ------------------------------------------------------------ kernels
template <int Q> struct MaxThreads { static constexpr int value = (Q == 9) ? 1024 : 512; };

// Per-tile region masks of the flags: summary[t] marks the regions of tile t that hold a
// non-FLUID node. With moving walls included it also writes mvs, one byte per node: 1 for a
// SOLID node whose bc velocity is non-zero or whose bc density is not finite (a moving wall).
// Each lane covers 16 consecutive nodes of one tile (tile_cells is a power of two in [32, 1024]);
// a tile is tile_cells / 16 consecutive lanes of one wave, OR-reduced by shuffles. n_chunks =
// cells / 16 lanes in total.
//
// With seen (the flags the masks were last built from, then their moving-wall bytes in mvs), a
// chunk whose flags or moving-wall bytes differ from those sets changed[0] and updates seen and
// mvs; tile_kind_kernel and node_mask_kernel then rebuild, and otherwise return at once. Without
// seen (a build that always rebuilds) the boundary-tile count is reset here.
template <int D, class R>
__global__ void __launch_bounds__(256) tile_summary_kernel(const uint8_t* __restrict__ flags,
                                                           const R* __restrict__ bc, uint8_t* __restrict__ mvs,
                                                           uint32_t* __restrict__ summary,
                                                           uint32_t* __restrict__ btiles, uint8_t* __restrict__ seen,
                                                           uint32_t* __restrict__ changed, Geo g, uint32_t n_chunks) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;  // 16-byte chunk
    if (i == 0 && !seen) btiles[g.n_tiles] = 0;  // boundary-tile count of the tile_kind_kernel that follows
    uint32_t mask = 0;
    if (i < n_chunks) {
        const uint4 v = reinterpret_cast<const uint4*>(flags)[i];
        uint4 mv = make_uint4(0u, 0u, 0u, 0u);
        if ((v.x | v.y | v.z | v.w) != 0u) {
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
            const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
            const uint32_t l0 = (i << 4) & (uint32_t)(g.tile_cells - 1);
            uint32_t solid = 0;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t fb = (w[j >> 2] >> ((j & 3) * 8)) & 0xFFu;
                const int l = (int)l0 + j;
                const int lx = l & (tx - 1), ly = (l >> g.txb) & (ty - 1), lz = l >> (g.txb + g.tyb);
                const uint32_t mx = axis_regions(0, 0) | (lx == 0 ? axis_regions(0, -1) : 0u) |
                                    (lx == tx - 1 ? axis_regions(0, 1) : 0u);
                const uint32_t my = axis_regions(1, 0) | (ly == 0 ? axis_regions(1, -1) : 0u) |
                                    (ly == ty - 1 ? axis_regions(1, 1) : 0u);
                const uint32_t mz = D == 2 ? axis_regions(2, 0)
                                           : (axis_regions(2, 0) | (lz == 0 ? axis_regions(2, -1) : 0u) |
                                              (lz == tz - 1 ? axis_regions(2, 1) : 0u));
                if (fb != 0u) mask |= mx & my & mz;
                solid |= ((fb & 0x0Fu) == LBM_FLAG_SOLID ? 1u : 0u) << j;
            }
            if (mvs && solid) {
                // the bc of all 16 nodes, loaded together (independent loads, one wait), then
                // tested for the SOLID ones
                const uint32_t tile = (i << 4) / (uint32_t)g.tile_cells;
                const uint32_t tc = (uint32_t)g.tile_cells;
                R b[D + 1][16];
#pragma unroll
                for (int d = 0; d <= D; ++d)
#pragma unroll
                    for (int j = 0; j < 16; ++j) b[d][j] = bc[((size_t)tile * (D + 1) + d) * tc + l0 + (uint32_t)j];
                uint32_t mw[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    bool m = !isfinite(b[D][j]);
#pragma unroll
                    for (int d = 0; d < D; ++d) m = m || b[d][j] != (R)0;
                    if (((solid >> j) & 1u) && m) mw[j >> 2] |= 1u << ((j & 3) * 8);
                }
                mv = make_uint4(mw[0], mw[1], mw[2], mw[3]);
            }
        }
        if (seen) {
            const uint4 sv = reinterpret_cast<const uint4*>(seen)[i];
            const uint4 om = mvs ? reinterpret_cast<const uint4*>(mvs)[i] : mv;
            if (((sv.x ^ v.x) | (sv.y ^ v.y) | (sv.z ^ v.z) | (sv.w ^ v.w) | (om.x ^ mv.x) | (om.y ^ mv.y) |
                 (om.z ^ mv.z) | (om.w ^ mv.w)) != 0u) {
                reinterpret_cast<uint4*>(seen)[i] = v;
                if (mvs) reinterpret_cast<uint4*>(mvs)[i] = mv;
                atomicOr(changed, 1u);
            }
        } else if (mvs) {
            reinterpret_cast<uint4*>(mvs)[i] = mv;
        }
    }
    const uint32_t lanes_per_tile = (uint32_t)g.tile_cells >> 4;  // 2 .. 64
    for (uint32_t off = 1; off < lanes_per_tile; off <<= 1) mask |= (uint32_t)__shfl_xor((int)mask, (int)off, 64);
    if ((threadIdx.x & (lanes_per_tile - 1u)) == 0u && i < n_chunks) summary[i / lanes_per_tile] = mask;
}

// The per-tile kind (1: the tile or its one-node halo holds a non-FLUID node) and, for boundary
// tiles, every node's neighbour mask: set_adjacent_flags
// (src/lbm/kernels/utils/set_flags.mojo) done once per flag state instead of in every step. One
// workgroup per tile (tile_cells lanes), the tile's flags and their one-node halo staged in LDS.
// With moving walls included (mvs, from tile_summary_kernel), a node with a SOLID pull neighbour
// whose bc velocity is non-zero or whose bc density is not finite is marked NB_MOVING, and so is a
// node with a decorated SOLID neighbour (its flag is read again in the step). Next to any other
// SOLID node the moving-wall term 6 w rho (c . u) is a signed zero, which the step does not add:
// it can change only the sign of a zero population, and the collision that follows (f - f_eq,
// with f_eq != 0) does not carry that sign into what it stores. The moving-wall bytes are staged
// beside the flags in LDS (sfl, then smv). Workgroups loop over tiles (an interior tile costs
// only its kind).
template <class L>
__global__ void __launch_bounds__(1024) node_mask_kernel(const uint8_t* __restrict__ flags,
                                                         const uint8_t* __restrict__ mvs,
                                                         const uint32_t* __restrict__ btiles,
                                                         const uint32_t* __restrict__ changed,
                                                         uint32_t* __restrict__ nbmask, Geo g, uint32_t halo_cells) {
    extern __shared__ uint8_t sfl[];
    uint8_t* smv = sfl + halo_cells;
    if (changed && !*changed) return;  // the masks of these flags are current
    const uint32_t n_boundary = btiles[g.n_tiles];
    uint32_t i = blockIdx.x;
    if (i >= n_boundary) return;
    TileCtx<L, uint32_t> t;
    t.init_pos(g, btiles[i], 0, g.tzb);
    t.fetch_flags(g, flags, mvs);
    while (true) {
        t.commit_flags(g, sfl, smv);
        uint32_t m = lds_mask(t, sfl);
        if (mvs && !(m & NB_INVALID)) {
            bool mv = (m & NB_DECOR) != 0;
#pragma unroll
            for (int q = 1; q < L::Q; ++q) mv = mv || t.pull_flag(smv, q) != 0;
            if (mv) m |= NB_MOVING;
        }
        // the next tile's flags are fetched while this tile's mask is stored
        const uint32_t next = i + gridDim.x;
        const bool more = next < n_boundary;  // uniform
        TileCtx<L, uint32_t> tn;
        if (more) {
            tn.init_pos(g, btiles[next], 0, g.tzb);
            tn.fetch_flags(g, flags, mvs);
        }
        nbmask[(size_t)t.tile * t.tc + t.loc] = m;
        if (!more) break;
        __syncthreads();  // this tile's LDS reads are done before the next tile's flags overwrite them
        t = tn;
        i = next;
    }
}

// Workgroups of node_mask_kernel (whole tiles): one per CU slot at full occupancy, each looping
// over the boundary tiles. (One-wave z slabs measured slower: their halos hold 2.4 times the cells.)
constexpr uint32_t kMaskGrid = 1024;

// The kind of every tile from the region masks (boundary_tile), one thread per tile (the 27
// summary words of a tile are independent loads of one lane), and the list of boundary tiles
// btiles[0 .. btiles[n_tiles]) for node_mask_kernel (in no particular order; the count starts
// at zero, set by tile_summary_kernel or mask_done_kernel).
template <int D>
__global__ void __launch_bounds__(256) tile_kind_kernel(const uint32_t* __restrict__ summary,
                                                        uint32_t* __restrict__ tile_kind,
                                                        uint32_t* __restrict__ btiles,
                                                        const uint32_t* __restrict__ changed, Geo g) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= (uint32_t)g.n_tiles || (changed && !*changed)) return;
    const int bx = (int)(t % (uint32_t)g.Tx), byz = (int)(t / (uint32_t)g.Tx);
    const uint32_t kind = boundary_tile<D>(g, summary, bx, byz % g.Ty, byz / g.Ty);
    tile_kind[t] = kind;
    if (kind) btiles[atomicAdd(&btiles[g.n_tiles], 1u)] = t;
}

// After node_mask_kernel: clears the change flag and the boundary-tile count for the next check.
template <int UNUSED = 0>
__global__ void __launch_bounds__(64) mask_done_kernel(uint32_t* __restrict__ changed, uint32_t* __restrict__ btiles,
                                                       uint32_t n_tiles) {
    if (threadIdx.x == 0) {
        *changed = 0u;
        btiles[n_tiles] = 0u;
    }
}

// double_buffer_kernel (src/lbm/kernels/double_buffer.mojo:25-96) and
// esoteric_pull_kernel[is_even] (src/lbm/kernels/esoteric_pull.mojo:24-111).
//
// capture_density / capture_velocity (src/lbm/kernels/steps/collide.mojo:54-63): a node whose flag
// carries LBM_FLAG_CAPTURE_DENSITY / _VELOCITY writes the density / velocity its collision starts
// from (after streaming and the boundary conditions) into its own bc entry, value D / values 0..D-1.
// The reference addresses that entry through get_rank4_coord, whose index is broken
// (src/lbm/kernels/utils/index.mojo:120-121); the node's own entry is what it names. Only valid
// (non-SOLID) nodes collide and capture, and no other node reads a valid node's bc in the step, so
// the writes race with nothing.
template <class L, class S, bool DDF, bool BIG>
__device__ __forceinline__ void capture_moments(const TileCtx<L, Idx<BIG>>& t, const StepParams<typename S::R>& p,
                                                const typename S::R (&fv)[L::Q]) {
    using R = typename S::R;
    constexpr int D = L::D;
    const uint8_t fl = p.flags[t.field_index(p.g, t.x, t.y, t.z, 1, 0)];
    R* bcw = const_cast<R*>(p.bc);
    const R rho = density<L, R, DDF>(fv);
    if ((p.capture & LBM_FLAG_CAPTURE_DENSITY) && (fl & LBM_FLAG_CAPTURE_DENSITY))
        bcw[t.field_index(p.g, t.x, t.y, t.z, D + 1, D)] = rho;
    if ((p.capture & LBM_FLAG_CAPTURE_VELOCITY) && (fl & LBM_FLAG_CAPTURE_VELOCITY)) {
        R u[3];
        velocity<L, R>(fv, rho, u);
#pragma unroll
        for (int d = 0; d < D; ++d) bcw[t.field_index(p.g, t.x, t.y, t.z, D + 1, d)] = u[d];
    }
}

// step_tile updates one tile; force_slot numbers its per-wave force records. LINKS
// compiles the interpolated bounce-back path in (double buffer only): kept out of the
// plain kernels, whose register budget it would otherwise set.
template <class L, class S, int MODE, int COLL, int LES, bool DDF, bool BIG, bool LINKS, int STAGE = 0>
__device__ __forceinline__ void step_tile(const StepParams<typename S::R>& p, uint32_t tile, uint32_t force_slot,
                                          uint32_t part, int lzb) {
    using R = typename S::R;
    constexpr int Q = L::Q;
    TileCtx<L, Idx<BIG>> t;
    t.init_pos(p.g, tile, part, lzb, STAGE == 2 ? (int)(threadIdx.x & 255u) : -1);
    const uint32_t kind = p.tile_kind[t.tile];  // uniform
    const bool boundary = (kind & 1u) != 0;
    const Field<S, BIG> fin(p.f_in, p.f_bytes);
    const Field<S, BIG> fout(p.f_out, p.f_bytes);
    uint32_t m = 0;  // issued before the population loads: its latency overlaps theirs
    if (boundary) m = p.nbmask[(size_t)t.tile * t.tc + t.loc];
    R fv[Q];
    if constexpr (STAGE > 0 && S::id == LBM_STORE_F16C && MODE == MODE_DB && !BIG && L::D == 3)
        stage_pull<L, STAGE>(t, p.g, p.f_in, p.f_bytes, boundary, fv);
    else
        load_populations<L, S, MODE, BIG>(t, fin, fv);
    constexpr bool links = LINKS && MODE == MODE_DB;
    double force[3] = {0.0, 0.0, 0.0};
    bool active = true;
    if (m) {
        if (m & NB_INVALID) {
            // stays for the force reduction, or for the learned closure, which needs every lane
            if (!links && LES != LES_MLP) return;
            active = false;
        } else {
            if constexpr (STAGE > 0 && S::id == LBM_STORE_F16C && MODE == MODE_DB && !BIG && L::D == 3)
                staged_bounce_back<L, STAGE>(t, m, fin, fv);
            else if (MODE == MODE_DB)
                halfway_bounce_back<L, S, BIG>(t, m, fin, fv);
            uint32_t rebuilt = 0;
            if (links && (m & NB_FSB))
                rebuilt = bouzidi_links<L, S, DDF, BIG>(t, m, fin, p.link_nodes, p.link_q, p.n_links,
                                                        t.tile * t.tc + t.loc, fv, force, p.node_force);
            apply_boundaries<L, S, MODE, DDF, BIG, S::id == LBM_STORE_F32 && LES != LES_MLP>(
                t, p.g, m, p.flags, fin, p.bc, p.moving, p.eqbc, fv, rebuilt);
            if (p.capture && (m & NB_CAPTURE)) capture_moments<L, S, DDF, BIG>(t, p, fv);
        }
    }
    if (LES == LES_MLP) {
        collide<L, R, COLL, LES, DDF>(fv, p.tau, p.inv_tau, p.cs, p.mlp);  // whole wave (matrix cores)
        if (active) store_populations<L, S, MODE, BIG>(t, fout, fv);
    } else if (active) {
        collide<L, R, COLL, LES, DDF>(fv, p.tau, p.inv_tau, p.cs, p.mlp);
        store_populations<L, S, MODE, BIG>(t, fout, fv);
    }
    if (links) {
        // Momentum-exchange force of this workgroup: wave sums in lane order, one record per wave.
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int waves = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int d = 0; d < L::D; ++d) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) force[d] += __shfl_down(force[d], off, 64);
            if (lane == 0) p.force_partial[((size_t)force_slot * waves + wave) * L::D + d] = force[d];
        }
    }
}

// One workgroup per z slab of 2^wzb layers: 2^(tzb - wzb) workgroups per tile, numbered
// tile-major, so the force records keep the whole-tile order (slot = unit, wave inside it).
template <class L, class S, int MODE, int COLL, int LES, bool DDF, bool BIG, bool LINKS>
__global__ void __launch_bounds__(MaxThreads<L::Q>::value) step_kernel(StepParams<typename S::R> p) {
    const int s = p.g.tzb - p.g.wzb;
    const uint32_t unit = (uint32_t)tile_of_block(p.g, s);
    step_tile<L, S, MODE, COLL, LES, DDF, BIG, LINKS>(p, unit >> s, unit, unit & ((1u << s) - 1u), p.g.wzb);
}

// The staged-pull step (Float16C, double buffer, 8 x 8 x tz tiles in 256-lane z slabs of four
// layers, 32-bit addressing, no interpolated links); dynamic LDS: StagedBox<L, BTX>::bytes().
// BTX = 2: one workgroup per slab of two x-adjacent tiles (Tx even), numbered pair-major, with the
// XCD remap over pairs when the tiles per XCD are even (else hardware order).
template <class L, int COLL, int LES, bool DDF, int BTX>
__global__ void __launch_bounds__(256 * BTX) step_kernel_staged(StepParams<float> p) {
    const int s = p.g.tzb - p.g.wzb;
    uint32_t unit;
    if constexpr (BTX == 1) {
        unit = (uint32_t)tile_of_block(p.g, s);
    } else {
        const uint32_t b = blockIdx.x;
        const uint32_t chunk = (p.g.xcd_chunk & 1) ? 0u : (uint32_t)(p.g.xcd_chunk >> 1) << s;
        unit = chunk ? (b & 7u) * chunk + (b >> 3) : b;
    }
    const uint32_t tile = (unit >> s) * BTX + (BTX == 1 ? 0u : (threadIdx.x >> 8));
    step_tile<L, StoreF16C, MODE_DB, COLL, LES, DDF, false, false, BTX>(p, tile, unit, unit & ((1u << s) - 1u),
                                                                        p.g.wzb);
}

// The step with the learned closure, held to at most 128 VGPRs (two 512-lane workgroups per CU)
// so the population traffic keeps enough waves in flight around the matrix-core work.
template <class L, class S, int MODE, int COLL, bool DDF>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 8)))
    step_kernel_mlp(StepParams<typename S::R> p) {
    const int s = p.g.tzb - p.g.wzb;
    const uint32_t unit = (uint32_t)tile_of_block(p.g, s);
    step_tile<L, S, MODE, COLL, LES_MLP, DDF, false, false>(p, unit >> s, unit, unit & ((1u << s) - 1u), p.g.wzb);
}

// Direction q of a unit move along +axis (sign 1) or -axis (sign -1).
template <class L> constexpr int face_q(int axis, int sign) {
    for (int q = 1; q < L::Q; ++q) {
        bool match = true;
        for (int d = 0; d < L::D; ++d) match = match && L::c(q, d) == (d == axis ? sign : 0);
        if (match) return q;
    }
    return -1;
}

// Where fd_gradient learns that a neighbour is SOLID (raw flag): nowhere (an interior tile, whose
// halo is all FLUID), the LDS flag tile sfl, or the node's neighbour mask (bit q: the pull
// neighbour x - c_q is SOLID; its periodic wrap at the domain edge is never consulted, since the
// gradient takes no neighbour outside the grid).
enum SolidSource { SOLID_NONE = 0, SOLID_LDS = 1, SOLID_MASK = 2 };

// get_velocity_gradient (src/lbm/kernels/utils/finite_difference.mojo:15-136), dx = 1: the
// one-sided differences to the -1 / +1 neighbours along `axis`, a SOLID neighbour (raw flag, from
// the LDS flag tile or the node's mask, by SRC) at half distance, blended by left_weight =
// 1 - left_dx / (left_dx + right_dx), right_weight = 1 - left_weight. Neighbours outside the grid
// are skipped (no periodic wrap), which leaves a single valid side with weight 0. Velocities come
// from the LDS velocity box su (D values per cell, strides lstride, the node at center).
template <class L, class R, int SRC>
__device__ __forceinline__ R fd_gradient(const int (&pos)[3], const int (&n)[3], const R* su, const uint8_t* sfl,
                                         uint32_t mask, int center, const int (&lstride)[3], int comp, int axis) {
    constexpr int D = L::D;
    // A SOLID neighbour sits at half distance (dx 0.5, so the difference is scaled by 1 / dx = 2,
    // exactly); the left weight is 1 - dx_l / (dx_l + dx_r). The quotients take four values,
    // written as constants that round as the division does, so no division is executed.
    constexpr R w_equal = (R)1 - (R)1 / ((R)1 + (R)1);          // dx_l = dx_r (1 / 2 = 0.5 / 1)
    constexpr R w_right_half = (R)1 - (R)1 / ((R)1 + (R)0.5);   // dx_l = 1, dx_r = 0.5
    constexpr R w_left_half = (R)1 - (R)0.5 / ((R)0.5 + (R)1);  // dx_l = 0.5, dx_r = 1
    const R uc = su[center * D + comp];
    R left_grad = 0, right_grad = 0;
    bool left_half = false, right_half = false;
    const bool left_ok = pos[axis] - 1 >= 0;
    if (left_ok) {
        const int c = center - lstride[axis];
        left_half = SRC == SOLID_LDS ? sfl[c] == LBM_FLAG_SOLID
                    : SRC == SOLID_MASK ? ((mask >> face_q<L>(axis, 1)) & 1u) != 0u : false;
        left_grad = (uc - su[c * D + comp]) * (left_half ? (R)2 : (R)1);
    }
    const bool right_ok = pos[axis] + 1 < n[axis];
    if (right_ok) {
        const int c = center + lstride[axis];
        right_half = SRC == SOLID_LDS ? sfl[c] == LBM_FLAG_SOLID
                     : SRC == SOLID_MASK ? ((mask >> face_q<L>(axis, -1)) & 1u) != 0u : false;
        right_grad = (su[c * D + comp] - uc) * (right_half ? (R)2 : (R)1);
    }
    // one side only: dx_l / dx_l = 1 (left) or 0 / dx_r = 0 (right)
    const R left_w = !right_ok ? (R)0
                     : !left_ok ? (R)1
                     : left_half == right_half ? w_equal
                     : right_half ? w_right_half : w_left_half;
    const R right_w = (R)1 - left_w;
    return (left_ok || right_ok) ? right_grad * right_w + left_grad * left_w : (R)0;
}

// Q = 0.25 |omega|^2 - 0.5 |S|^2 of a node (calculate_Q_criterion,
// src/lbm/output/Q_criterion.mojo:90-215): |omega|^2 from finite differences of the velocity box
// su (vorticity_sq: D values per cell, strides ls, the node at center; FLAGS: SOLID neighbours in
// sfl, same cell order, at half distance), S from the node's re-streamed populations fv (q_from);
// q_value does both.
template <class L, class R, int SRC>
__device__ __forceinline__ R vorticity_sq(const int (&pos)[3], const int (&nn)[3], const R* su, const uint8_t* sfl,
                                          uint32_t mask, int center, const int (&ls)[3]) {
    constexpr int D = L::D;
    auto grad = [&](int comp, int axis) {
        return fd_gradient<L, R, SRC>(pos, nn, su, sfl, mask, center, ls, comp, axis);
    };
    R wsq;
    if (D == 2) {
        const R w = grad(1, 0) - grad(0, 1);
        wsq = w * w;
    } else {
        const R wx = grad(2, 1) - grad(1, 2);
        const R wy = grad(0, 2) - grad(2, 0);
        const R wz = grad(1, 0) - grad(0, 1);
        wsq = wx * wx;
        wsq = wsq + wy * wy;
        wsq = wsq + wz * wz;
    }
    return wsq;
}

// Q of a node from |omega|^2 (vorticity_sq) and its re-streamed populations.
template <class L, class R, bool DDF>
__device__ __forceinline__ R q_from(R wsq, const R (&fv)[L::Q], R tau) {
    constexpr int Q = L::Q;
    const R rho = density<L, R, DDF>(fv);
    R u[3];
    velocity<L, R>(fv, rho, u);
    const R uu = dot_uu<L, R>(u);
    R fn[Q], pi[6];
#pragma unroll
    for (int q = 0; q < Q; ++q) fn[q] = fv[q] - feq<L, R, DDF>(q, rho, u, uu);
    pi_neq<L, R>(fn, pi);
    return (R)0.25 * wsq - (R)0.5 * strain_norm_sq<L, R>(pi, rho, tau);
}

template <class L, class R, bool DDF, bool FLAGS>
__device__ __forceinline__ R q_value(const int (&pos)[3], const int (&nn)[3], const R* su, const uint8_t* sfl,
                                     int center, const int (&ls)[3], const R (&fv)[L::Q], R tau) {
    return q_from<L, R, DDF>(vorticity_sq<L, R, FLAGS ? SOLID_LDS : SOLID_NONE>(pos, nn, su, sfl, 0u, center, ls), fv,
                             tau);
}

// Loads the velocities of a tile and its one-node halo (cells outside the grid: 0, never
// read) from the row-major field into su, in the flag tile's cell order with D values per
// cell. Consecutive lanes read consecutive addresses (z, then the component innermost for
// the interleaved layout; z innermost per component plane for the planar one).
template <class R, int D>
__device__ __forceinline__ void load_velocity_tile(const MacroParams<R>& p, int bx, int by, int bz, int hx, int hxy,
                                                   R* su) {
    const int tz = 1 << p.g.tzb;
    const int hy = hxy / hx, hz = D == 3 ? tz + 2 : 1;
    const int hcells = hxy * hz;
    const size_t ncell = (size_t)p.g.nx * p.g.ny * p.g.nz;
    const bool planar = p.u_layout != LBM_U_XYZD;
    for (int e = threadIdx.x; e < hcells * D; e += blockDim.x) {
        int d, h;
        if (planar) {
            d = e / hcells;
            h = e - d * hcells;
        } else {
            h = e / D;
            d = e - h * D;
        }
        const int iz = h % hz, rest = h / hz;
        const int iy = rest % hy, ix = rest / hy;
        const int gx = (bx << p.g.txb) + ix - 1, gy = (by << p.g.tyb) + iy - 1;
        const int gz = D == 3 ? (bz << p.g.tzb) + iz - 1 : 0;
        R v = 0;
        if (gx >= 0 && gx < p.g.nx && gy >= 0 && gy < p.g.ny && gz >= 0 && gz < p.g.nz) {
            const size_t r = ((size_t)gx * p.g.ny + gy) * p.g.nz + gz;
            v = planar ? p.vel[(size_t)d * ncell + r] : p.vel[r * D + d];
        }
        su[((iz * hy + iy) * hx + ix) * D + d] = v;
    }
}

// Row-major outputs of one tile. Lanes hold nodes in the tile's column-major order (x fastest),
// which a row-major (x, y, z) field would receive as scattered words; the values go to LDS in
// row-major order of the tile (K values per node, interleaved or one plane per value) and each
// lane then writes the next element, so a wave writes runs of tz * K (interleaved) or tz (planar)
// consecutive words.
template <class R, int K>
__device__ __forceinline__ void write_rowmajor(const Geo& g, int bx, int by, int bz, int tid, const R* lds, R* out,
                                               bool planar) {
    const int ty = 1 << g.tyb, tz = 1 << g.tzb;
    const int tc = g.tile_cells;
    const size_t ncell = (size_t)g.nx * g.ny * g.nz;
#pragma unroll
    for (int pass = 0; pass < K; ++pass) {
        const int f = tid + pass * tc;
        const int node = planar ? f - (f / tc) * tc : f / K;
        const int k = planar ? f / tc : f - (f / K) * K;
        const int zl = node & (tz - 1), yl = (node >> g.tzb) & (ty - 1), xl = node >> (g.tzb + g.tyb);
        const size_t r = ((size_t)((bx << g.txb) + xl) * g.ny + ((by << g.tyb) + yl)) * g.nz + ((bz << g.tzb) + zl);
        out[planar ? (size_t)k * ncell + r : r * K + k] = lds[f];
    }
}

// Position of a node of the tile in the row-major staging order of write_rowmajor.
__device__ __forceinline__ int rowmajor_slot(const Geo& g, int tid) {
    const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
    const int lx = tid & (tx - 1), ly = (tid >> g.txb) & (ty - 1), lz = tid >> (g.txb + g.tyb);
    return (lx * ty + ly) * tz + lz;
}

// Tile of a macro_kernel workgroup: the step's order (x fastest, XCD-contiguous), so the pulls
// share their halo lines in L2 as the step's do. (Walking tiles z-fastest, or in groups of four
// tiles adjacent in z, so that the workgroups filling one row-major output line run together,
// measured 4-16 % slower at 256^3.)
__device__ __forceinline__ uint32_t macro_tile(const Geo& g) { return (uint32_t)tile_of_block(g); }

// calculate_rho_and_velocity (src/lbm/output/velocity.mojo:25-125), the FP64 moment
// reduction and calculate_Q_criterion (src/lbm/output/Q_criterion.mojo:90-215): re-stream +
// BC (the BCs take no tau), then moments, or Q = 0.25 |omega|^2 - 0.5 |S|^2 with omega from
// finite differences of the velocity field and S from the re-streamed populations.
// Solid nodes (raw flag == SOLID) report bc in the fields and 0 as Q.
// Interior tiles (own nodes and halo all FLUID, from the tile kinds) need no flags: no LDS flag
// tile, no barrier, no boundary work.
template <class L, class S, int MODE, bool DDF, int KIND, bool BIG>
__global__ void __launch_bounds__(MaxThreads<L::Q>::value)
    __attribute__((amdgpu_waves_per_eu(KIND == MK_FIELDS ? 8 : 6, 8))) macro_kernel(MacroParams<typename S::R> p) {
    using R = typename S::R;
    constexpr int Q = L::Q, D = L::D;
    extern __shared__ uint8_t sfl[];
    __shared__ double red[D + 1][16];
    R* stage = (R*)(sfl + p.stage_offset);  // output staging: FIELDS rho plane + u, QCRIT Q
    TileCtx<L, Idx<BIG>> t;
    t.init_pos(p.g, macro_tile(p.g), 0, p.g.tzb);
    const Field<S, BIG> f(p.f, p.f_bytes);
    const bool boundary = p.tile_kind ? (p.tile_kind[t.tile] & 1u) != 0u : true;
    if (boundary) t.fetch_flags(p.g, p.flags);
    R fv[Q];
    load_populations<L, S, MODE, BIG>(t, f, fv);
    if (boundary) {
        t.commit_flags(p.g, sfl);
    } else {
        t.nonfluid = 0;
        t.own = LBM_FLAG_FLUID;
    }
    if (t.nonfluid) {
        const uint32_t m = lds_mask(t, sfl) | NB_MOVING;  // the moving-wall term read for every SOLID neighbour
        uint32_t rebuilt = 0;
        if (MODE == MODE_DB) {
            halfway_bounce_back<L, S, BIG>(t, m, f, fv);
            if (p.link_nodes && (m & NB_FSB)) {
                double unused[3] = {0, 0, 0};
                rebuilt = bouzidi_links<L, S, DDF, BIG>(t, m, f, p.link_nodes, p.link_q, p.n_links,
                                                        t.tile * t.tc + t.loc, fv, unused);
            }
        }
        apply_boundaries<L, S, MODE, DDF, BIG>(t, p.g, m, p.flags, f, p.bc, p.moving, p.eqbc, fv, rebuilt);
    }
    const int slot = rowmajor_slot(p.g, t.tid);
    if (KIND == MK_QCRIT) {
        R qv = 0;
        R* su = (R*)(sfl + p.su_offset);
        load_velocity_tile<R, D>(p, t.bx, t.by, t.bz, t.hx, t.hxy, su);
        __syncthreads();
        if (t.own != LBM_FLAG_SOLID) {
            const int pos[3] = {t.x, t.y, t.z};
            const int nn[3] = {p.g.nx, p.g.ny, p.g.nz};
            const int ls[3] = {1, t.hx, t.hxy};
            qv = boundary ? q_value<L, R, DDF, true>(pos, nn, su, sfl, t.center, ls, fv, p.tau)
                          : q_value<L, R, DDF, false>(pos, nn, su, sfl, t.center, ls, fv, p.tau);
        }
        stage[slot] = qv;
        __syncthreads();
        write_rowmajor<R, 1>(p.g, t.bx, t.by, t.bz, t.tid, stage, p.qout, false);
        return;
    }
    constexpr bool REDUCE = KIND == MK_REDUCE;
    R rho = 0, u[3] = {0, 0, 0};
    bool contributes = false;
    if (REDUCE) {
        if (node_valid(t.own)) {
            rho = density<L, R, DDF>(fv);
            velocity<L, R>(fv, rho, u);
            contributes = true;
        }
    } else {
        if (t.own != LBM_FLAG_SOLID) {
            rho = density<L, R, DDF>(fv);
            velocity<L, R>(fv, rho, u);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) u[d] = p.bc[t.field_index(p.g, t.x, t.y, t.z, D + 1, d)];
            rho = p.bc[t.field_index(p.g, t.x, t.y, t.z, D + 1, D)];
        }
        // rho and u staged together: one barrier
        const bool planar = p.u_layout != LBM_U_XYZD;
        const int tc = p.g.tile_cells;
        stage[slot] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) stage[tc + (planar ? d * tc + slot : slot * D + d)] = u[d];
        __syncthreads();
        write_rowmajor<R, 1>(p.g, t.bx, t.by, t.bz, t.tid, stage, p.rho, false);
        write_rowmajor<R, D>(p.g, t.bx, t.by, t.bz, t.tid, stage + tc, p.u, planar);
        return;
    }
    // Deterministic block reduction in FP64: wave partials by lane order, then waves in order.
    double v[D + 1];
    v[0] = contributes ? (double)rho : 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) v[1 + d] = contributes ? (double)rho * (double)u[d] : 0.0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= D; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
        if (lane == 0) red[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int k = 0; k <= D; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += red[k][w];
            p.partial[(size_t)t.tile * (D + 1) + k] = s;
        }
    }
}

// A 16-byte store that streams past the caches (output fields are not read again by the kernel
// that writes them; the population lines a neighbouring column still needs stay in L2).
typedef float lbm_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt4(float* p, lbm_f4 v) { __builtin_nontemporal_store(v, (lbm_f4*)p); }

// calculate_rho_and_velocity (src/lbm/output/velocity.mojo:25-125) in columns: one 512-lane
// workgroup per column of four z-adjacent 8 x 8 x 8 tiles (Tz a multiple of 4), each lane one node
// of each tile in turn, at the same place in the tile. A wave reads an 8 x 8 layer of a tile per
// population load (whole 256-byte runs), and the column's rho and u wait in LDS until all four
// tiles are done, then leave row-major (z fastest) in whole 128-byte lines: 32 values of rho per
// (x, y), 96 of u, one 16-byte streaming store per lane and value group. A tile alone would leave
// 32-byte pieces whose lines three other tiles, 1,024 tiles away in the walk, complete. Boundary tiles
// (by the tile kinds, uniform per tile) read their nodes' neighbour masks and flags; every wave
// takes the same four tiles, so a boundary tile delays all of the workgroup's waves alike. The
// masks are the step's, built from the same flags and bc by the same call, so the moving-wall term
// is read where the step reads it (NB_MOVING) and skipped next to resting walls, where it is a
// signed zero: it could change only the sign of a zero population, and of a moment that sums to an
// exact zero from such populations. Otherwise the same arithmetic as macro_kernel, node by node.
// The output pointers are 16-byte aligned (checked on the host).
template <class L, class S, int MODE, bool DDF, bool BIG>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6, 8)))
    macro_fields_col(MacroParams<typename S::R> p) {
    using R = typename S::R;
    constexpr int Q = L::Q, D = L::D;
    static_assert(D == 3 && sizeof(R) == 4, "columns of 3-D tiles, 4-byte moments");
    constexpr int ZS = 65;  // LDS stride of a z layer (64 (x, y) + 1: no bank conflicts either way)
    __shared__ R srho[32 * ZS];
    __shared__ R su[3][32 * ZS];
    const uint32_t n_units = (uint32_t)p.g.n_tiles >> 2;  // columns
    if (blockIdx.x >= n_units) return;                     // a launch of more workgroups than columns
    uint32_t c = blockIdx.x;
    if ((n_units & 7u) == 0) c = (c & 7u) * (n_units >> 3) + (c >> 3);
    const uint32_t Tx = (uint32_t)p.g.Tx, Ty = (uint32_t)p.g.Ty;
    const int bx = (int)(c % Tx), by = (int)((c / Tx) % Ty), bz4 = (int)(c / (Tx * Ty));
    const uint32_t t0 = ((uint32_t)(bz4 * 4) * Ty + (uint32_t)by) * Tx + (uint32_t)bx, tstride = Tx * Ty;
    const int tid = threadIdx.x, lxy = tid & 63, lz = tid >> 6;
    const Field<S, BIG> f(p.f, p.f_bytes);
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        TileCtx<L, Idx<BIG>> t;
        t.init_pos(p.g, t0 + (uint32_t)k * tstride, 0, p.g.tzb, tid);
        const bool boundary = (p.tile_kind[t.tile] & 1u) != 0u;  // uniform
        uint32_t m = 0;
        uint8_t own = LBM_FLAG_FLUID;
        if (boundary) {
            const size_t node = (size_t)t.tile * t.tc + t.loc;
            m = p.nbmask[node];
            own = p.flags[node];
        }
        R fv[Q];
        load_populations<L, S, MODE, BIG>(t, f, fv);
        if (m) {
            uint32_t rebuilt = 0;
            if (MODE == MODE_DB) {
                halfway_bounce_back<L, S, BIG>(t, m, f, fv);
                if (p.link_nodes && (m & NB_FSB)) {
                    double unused[3] = {0, 0, 0};
                    rebuilt = bouzidi_links<L, S, DDF, BIG>(t, m, f, p.link_nodes, p.link_q, p.n_links,
                                                            t.tile * t.tc + t.loc, fv, unused);
                }
            }
            apply_boundaries<L, S, MODE, DDF, BIG>(t, p.g, m, p.flags, f, p.bc, p.moving, p.eqbc, fv, rebuilt);
        }
        R rho, u[3] = {0, 0, 0};
        if (own != LBM_FLAG_SOLID) {
            rho = density<L, R, DDF>(fv);
            velocity<L, R>(fv, rho, u);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) u[d] = p.bc[t.field_index(p.g, t.x, t.y, t.z, D + 1, d)];
            rho = p.bc[t.field_index(p.g, t.x, t.y, t.z, D + 1, D)];
        }
        const int a = (k * 8 + lz) * ZS + lxy;
        srho[a] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) su[d][a] = u[d];
    }
    __syncthreads();
    // 64 runs of 32 z per (x, y); lane: run tid / 8, z 4 (tid % 8) .. + 3 of rho (and of each u
    // plane), or three of the 24 16-byte pieces of the interleaved u runs
    const size_t ny = (size_t)p.g.ny, nz = (size_t)p.g.nz, ncell = (size_t)p.g.nx * ny * nz;
    const int x0 = bx * 8, y0 = by * 8;
    const size_t z0 = (size_t)bz4 * 32;
    auto row = [&](int xy) { return ((size_t)(x0 + (xy & 7)) * ny + (size_t)(y0 + (xy >> 3))) * nz + z0; };
    auto gather4 = [&](const R* lds, int xy, int z) {
        return lbm_f4{lds[z * ZS + xy], lds[(z + 1) * ZS + xy], lds[(z + 2) * ZS + xy], lds[(z + 3) * ZS + xy]};
    };
    const int xy = tid >> 3, z4 = (tid & 7) * 4;
    const size_t r = row(xy) + (size_t)z4;
    store_nt4(p.rho + r, gather4(srho, xy, z4));
    if (p.u_layout != LBM_U_XYZD) {
#pragma unroll
        for (int d = 0; d < D; ++d) store_nt4(p.u + (size_t)d * ncell + r, gather4(su[d], xy, z4));
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int g = tid + 512 * i, gxy = g / 24, e = g - 24 * gxy;  // piece e of run gxy
            float w[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int idx = e * 4 + j, z = idx / 3;
                w[j] = su[idx - 3 * z][z * ZS + gxy];
            }
            store_nt4(p.u + row(gxy) * 3 + (size_t)e * 4, lbm_f4{w[0], w[1], w[2], w[3]});
        }
    }
}

// Dynamic LDS of macro_qcrit_col: the velocity box (cells x D values) and the Q staging buffer,
// sizeof(R) bytes each.
constexpr int kQColBoxCells = 10 * 10 * 34, kQColStage = 32 * 65;

// calculate_Q_criterion (src/lbm/output/Q_criterion.mojo:90-215) in the columns of
// macro_fields_col: one 512-lane workgroup per four z-adjacent 8 x 8 x 8 tiles, each lane one node
// of each tile in turn (a wave reads an 8 x 8 layer per population load). The velocity box of the
// column and its one-node halo (10 x 10 x 34 nodes, 1.7 velocities read per node against 2.0 for a
// tile's 10 x 10 x 10 box) is staged in LDS first (interleaved u: straight from memory to LDS,
// without registers); Q then waits in LDS and leaves row-major in whole 128-byte lines, one 16-byte
// streaming store per lane. Boundary tiles take their masks and own
// flags as macro_fields_col does, and the gradient of a boundary tile's node finds its SOLID face
// neighbours (at half distance) in the node's mask; that of an interior tile (halo all FLUID)
// needs none. Same arithmetic as macro_kernel (q_value). Six waves per SIMD (three workgroups per
// CU; 49 KB of LDS) fit in 80 VGPRs because |omega|^2 is done before the population loads are
// issued.
template <class L, class S, int MODE, bool DDF, bool BIG>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6, 8)))
    macro_qcrit_col(MacroParams<typename S::R> p) {
    using R = typename S::R;
    constexpr int Q = L::Q, D = L::D;
    static_assert(D == 3 && sizeof(R) == 4, "columns of 3-D tiles, 4-byte moments");
    constexpr int HX = 10, HY = 10, HZ = 34, HC = HX * HY * HZ;  // velocity box: tile column + halo
    constexpr int ZS = 65;
    static_assert(HC == kQColBoxCells && 32 * ZS == kQColStage, "LDS size the launch gives");
    R* su = (R*)staged_box;        // HC * D: cells (x, y) rows of 34 z, D values each
    R* sq = su + HC * D;           // 32 * ZS
    const uint32_t n_units = (uint32_t)p.g.n_tiles >> 2;  // columns
    if (blockIdx.x >= n_units) return;                     // a launch of more workgroups than columns
    uint32_t c = blockIdx.x;
    if ((n_units & 7u) == 0) c = (c & 7u) * (n_units >> 3) + (c >> 3);
    const uint32_t Tx = (uint32_t)p.g.Tx, Ty = (uint32_t)p.g.Ty;
    const int bx = (int)(c % Tx), by = (int)((c / Tx) % Ty), bz4 = (int)(c / (Tx * Ty));
    const uint32_t t0 = ((uint32_t)(bz4 * 4) * Ty + (uint32_t)by) * Tx + (uint32_t)bx, tstride = Tx * Ty;
    const int tid = threadIdx.x, lxy = tid & 63, lz = tid >> 6;
    uint32_t kinds = 0;  // bit k: tile k of the column is a boundary tile (uniform)
#pragma unroll
    for (int k = 0; k < 4; ++k) kinds |= (p.tile_kind[t0 + (uint32_t)k * tstride] & 1u) << k;
    const int nx = p.g.nx, ny = p.g.ny, nz = p.g.nz;
    const size_t ncell = (size_t)nx * ny * nz;
    const bool planar = p.u_layout != LBM_U_XYZD;
    const int gx0 = bx * 8 - 1, gy0 = by * 8 - 1, gz0 = bz4 * 32 - 1;
    // The box row by row, cells (x, y, z) z fastest: wave w takes rows (x, y) w, w + 8, ... of the
    // 10 x 10; a row is the 34 z of the box, D values each. Interleaved layout: the row's 102
    // consecutive words go straight from memory to LDS (lane j's word at the row's LDS base + 4 j);
    // planar: lane j and j + 64 carry word j and j + 64 of the row through registers. Cells outside
    // the grid are never read by the gradient and are left as they are.
    const int lane = tid & 63, wave = tid >> 6;
    int64_t voff[2];
    int vlds[2];
    bool vok[2];
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
        const int j = lane + 64 * s2;
        const int iz = planar ? j % HZ : j / D, d = planar ? j / HZ : j % D;
        vok[s2] = j < HZ * D && gz0 + iz >= 0 && gz0 + iz < nz;
        voff[s2] = planar ? (int64_t)d * (int64_t)ncell + iz : (int64_t)j;
        vlds[s2] = iz * D + d;
    }
    for (int row = wave; row < HX * HY; row += 8) {
        const int ix = row / HY, iy = row - ix * HY;  // uniform
        const int gx = gx0 + ix, gy = gy0 + iy;
        if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) continue;
        const int64_t rz = ((int64_t)gx * ny + gy) * nz + gz0;  // row-major index of (gx, gy, gz0)
        const int rl = row * (HZ * D);
        if (planar) {
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2)
                if (vok[s2]) su[vlds[s2] + rl] = p.vel[rz + voff[s2]];
        } else {
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2)
                if (vok[s2]) {
#if defined(__HIP_DEVICE_COMPILE__)
                    __builtin_amdgcn_global_load_lds(p.vel + rz * D + voff[s2], (uint8_t*)(su + rl + 64 * s2), 4, 0, 0);
#endif
                }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const Field<S, BIG> f(p.f, p.f_bytes);
    const int nn[3] = {nx, ny, nz};
    const int ls[3] = {HY * HZ, HZ, 1};
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        TileCtx<L, Idx<BIG>> t;
        t.init_pos(p.g, t0 + (uint32_t)k * tstride, 0, p.g.tzb, tid);
        const bool boundary = ((kinds >> k) & 1u) != 0u;  // uniform
        uint32_t m = 0;
        uint8_t own = LBM_FLAG_FLUID;
        if (boundary) {
            const size_t node = (size_t)t.tile * t.tc + t.loc;
            m = p.nbmask[node];
            own = p.flags[node];
        }
        // |omega|^2 from the LDS box first, so that its registers are free again before the
        // population loads are in flight (unused at SOLID nodes); SOLID face neighbours of a
        // boundary tile's node from its mask
        const int pos[3] = {t.x, t.y, t.z};
        const int center = (((tid & 7) + 1) * HY + (tid >> 3 & 7) + 1) * HZ + k * 8 + lz + 1;
        const R wsq = boundary ? vorticity_sq<L, R, SOLID_MASK>(pos, nn, su, nullptr, m, center, ls)
                               : vorticity_sq<L, R, SOLID_NONE>(pos, nn, su, nullptr, 0u, center, ls);
        R fv[Q];
        load_populations<L, S, MODE, BIG>(t, f, fv);
        if (m) {
            uint32_t rebuilt = 0;
            if (MODE == MODE_DB) {
                halfway_bounce_back<L, S, BIG>(t, m, f, fv);
                if (p.link_nodes && (m & NB_FSB)) {
                    double unused[3] = {0, 0, 0};
                    rebuilt = bouzidi_links<L, S, DDF, BIG>(t, m, f, p.link_nodes, p.link_q, p.n_links,
                                                            t.tile * t.tc + t.loc, fv, unused);
                }
            }
            apply_boundaries<L, S, MODE, DDF, BIG>(t, p.g, m, p.flags, f, p.bc, p.moving, p.eqbc, fv, rebuilt);
        }
        sq[(k * 8 + lz) * ZS + lxy] = own != LBM_FLAG_SOLID ? q_from<L, R, DDF>(wsq, fv, p.tau) : (R)0;
    }
    __syncthreads();
    const int xy = tid >> 3, z4 = (tid & 7) * 4;
    const size_t r = ((size_t)(bx * 8 + (xy & 7)) * ny + (size_t)(by * 8 + (xy >> 3))) * nz + (size_t)(bz4 * 32 + z4);
    store_nt4(p.qout + r, lbm_f4{sq[z4 * ZS + xy], sq[(z4 + 1) * ZS + xy], sq[(z4 + 2) * ZS + xy], sq[(z4 + 3) * ZS + xy]});
}

// Sums the per-tile partials in tile order (one workgroup, lane-strided, then lane order).
template <int K>
__global__ void __launch_bounds__(256) reduce_finish(const double* __restrict__ partial, int n_tiles, double* out) {
    __shared__ double s[K][256];
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int i = threadIdx.x; i < n_tiles; i += 256)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += partial[(size_t)i * K + k];
#pragma unroll
    for (int k = 0; k < K; ++k) s[k][threadIdx.x] = acc[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; ++k) {
            double t = 0.0;
            for (int i = 0; i < 256; ++i) t += s[k][i];
            out[k] = t;
        }
    }
}

// Fill every node with f_eq(rho, u); for esoteric pull the odd store pattern of a uniform
// state puts slot q = f_q at every node, so both methods reduce to this fill
// (initialize_f_from_func, src/lbm/preprocess/initial_condition.mojo:86-178).
template <class R> struct QVals { R v[27]; };

template <class S>
__global__ void __launch_bounds__(256) fill_kernel(typename S::T* __restrict__ f, QVals<typename S::R> vals, int Q,
                                                    uint32_t tile_cells, size_t n_elems) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_elems) return;
    const uint32_t q = (uint32_t)((i / tile_cells) % (uint32_t)Q);
    S::st(f + i, vals.v[q]);
}

// Float16C codec over a flat array (lbm_f16c_convert).
template <int UNUSED = 0>
__global__ void __launch_bounds__(256) f16c_convert_kernel(const void* in, void* out, size_t n, int to_f16c) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (to_f16c) ((uint16_t*)out)[i] = (uint16_t)f16c_encode(((const float*)in)[i]);
    else ((float*)out)[i] = f16c_decode(((const uint16_t*)in)[i]);
}

// set_exterior_walls host loop, src/lbm/preprocess/boundary_condition.mojo:160-183.
template <class R>
__global__ void __launch_bounds__(256) wall_kernel(uint8_t* flags, R* bc, Geo g, int D, int axis, int plus,
                                                    uint8_t flag, R u0, R u1, R u2, R rho) {
    const int n[3] = {g.nx, g.ny, g.nz};
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n[a1] * n[a2]) return;
    int idx[3];
    idx[axis] = plus ? n[axis] - 1 : 0;
    idx[a1] = i % n[a1];
    idx[a2] = i / n[a1];
    const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
    const uint32_t loc = (uint32_t)((idx[0] & (tx - 1)) + ((idx[1] & (ty - 1)) << g.txb) + ((idx[2] & (tz - 1)) << (g.txb + g.tyb)));
    const size_t tl = (size_t)((idx[0] >> g.txb) + g.Tx * ((idx[1] >> g.tyb) + g.Ty * (idx[2] >> g.tzb)));
    flags[tl * (size_t)g.tile_cells + loc] = flag;
    const R vals[4] = {u0, u1, u2, rho};
    for (int k = 0; k <= D; ++k) bc[(tl * (size_t)(D + 1) + (size_t)k) * (size_t)g.tile_cells + loc] = k < D ? vals[k] : rho;
}

}  // namespace lbm

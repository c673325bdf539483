// lbm_kernels.hpp -- gfx950 stream-collide kernels.
//
// One workgroup per layout tile, one lane per node; lane id == the node's
// column-major index inside the tile, so every population plane of a tile is
// read and written as contiguous 64-lane runs. The tile's flags plus a one-node
// halo are staged in LDS; a tile whose flags and halo are all FLUID takes a
// branch-free periodic pull path. Collision arithmetic follows the reference's
// operation order (moments summed in q order, u = sum/rho, 1/tau, f_eq as
// written) and the library is built with -ffp-contract=off, so FP32 and FP64
// results are rounded exactly like the CPU restatement.
#pragma once

#include "lbm_common.hpp"

namespace lbm {

enum Mode { MODE_DB = 0, MODE_EP_EVEN = 1, MODE_EP_ODD = 2 };

template <class R> struct StepParams {
    const void* f_in;   // typed by the storage
    void* f_out;        // == f_in for esoteric pull
    const R* bc;
    const uint8_t* flags;
    R tau;
    R cs;
    int moving;         // include_moving_boundary
    int eqbc;           // EQUILIBRIUM in the included BCs
    Geo g;
};

template <class R> struct MacroParams {
    const void* f;
    const R* bc;
    const uint8_t* flags;
    R* rho;
    R* u;
    double* partial;    // reduction mode: per-tile (D+1) sums
    int u_layout;
    int moving;
    int eqbc;
    Geo g;
};

// ------------------------------------------------------------ moment helpers
// Sums skip the terms whose lattice coefficient is zero: the reference adds
// f*0 = +-0 there (Vector.dot / Vector.sum, src/utils/vector.mojo:205-224),
// which leaves every nonzero partial sum unchanged.
template <class L, class R>
__host__ __device__ __forceinline__ R dot_c(const R (&u)[3], int q) {
    R out = 0;
    bool first = true;
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
        const int c = L::c(q, d);
        if (c == 0) continue;
        const R t = c > 0 ? u[d] : -u[d];
        out = first ? t : out + t;
        first = false;
    }
    return out;
}

template <class L, class R>
__host__ __device__ __forceinline__ R dot_uu(const R (&u)[3]) {
    R out = u[0] * u[0];
#pragma unroll
    for (int d = 1; d < L::D; ++d) out = out + u[d] * u[d];
    return out;
}

// f_eq, src/lbm/kernels/utils/equilibrium.mojo:10-38.
template <class L, class R, bool DDF>
__host__ __device__ __forceinline__ R feq(int q, R rho, const R (&u)[3], R uu) {
    const R eu = dot_c<L, R>(u, q);
    const R w = weight<L, R>(q);
    if (DDF) return w * rho * ((R)3 * eu + (R)4.5 * eu * eu - (R)1.5 * uu) + w * (rho - (R)1);
    return w * rho * ((R)1 + (R)3 * eu + (R)4.5 * eu * eu - (R)1.5 * uu);
}

// get_density (moment.mojo:17-41) and get_velocity (moment.mojo:45-76).
template <class L, class R, bool DDF>
__device__ __forceinline__ R density(const R (&f)[L::Q]) {
    R s = f[0];
#pragma unroll
    for (int q = 1; q < L::Q; ++q) s = s + f[q];
    return DDF ? s + (R)1 : s;
}

template <class L, class R>
__device__ __forceinline__ void velocity(const R (&f)[L::Q], R rho, R (&u)[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = 0;
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
        bool first = true;
        R acc = 0;
#pragma unroll
        for (int q = 0; q < L::Q; ++q) {
            const int c = L::c(q, d);
            if (c == 0) continue;
            const R t = c > 0 ? f[q] : -f[q];
            acc = first ? t : acc + t;
            first = false;
        }
        u[d] = acc / rho;
    }
}

__host__ __device__ constexpr int stress_a(int D, int n) { return D == 2 ? (n < 2 ? 0 : 1) : (n < 3 ? 0 : (n < 5 ? 1 : 2)); }
__host__ __device__ constexpr int stress_b(int D, int n) {
    return D == 2 ? (n == 0 ? 0 : 1) : (n == 0 ? 0 : (n == 1 ? 1 : (n == 2 ? 2 : (n == 3 ? 1 : 2))));
}

// get_non_eq_second_order_moment / get_Qiab, moment.mojo:80-150.
template <class L, class R>
__device__ __forceinline__ void pi_neq(const R (&fn)[L::Q], R (&pi)[6]) {
    constexpr int NS = L::D * (L::D + 1) / 2;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        const int a = stress_a(L::D, n), b = stress_b(L::D, n);
        bool first = true;
        R acc = 0;
#pragma unroll
        for (int q = 0; q < L::Q; ++q) {
            const int cc = L::c(q, a) * L::c(q, b);
            if (cc == 0) continue;
            const R t = cc > 0 ? fn[q] : -fn[q];
            acc = first ? t : acc + t;
            first = false;
        }
        pi[n] = acc;
    }
}

// Smagorinsky eddy relaxation time: get_strain_rate_tensor (moment.mojo:154-176),
// its squared norm (:182-215) and get_Smagorinsky_LES_tau (ops/turbulence.mojo:12-38).
// The strain rate uses the base tau (steps/collide.mojo:69).
template <class L, class R>
__device__ __forceinline__ R smagorinsky_tau(const R (&pi)[6], R rho, R tau, R cs) {
    constexpr int NS = L::D * (L::D + 1) / 2;
    const R cs2 = (R)1 / (R)3;
    const R k = (R)(-1) / ((R)2 * rho * cs2 * tau);
    R snorm = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        const R s = pi[n] * k;
        const R ss = s * s;
        snorm = snorm + ((stress_a(L::D, n) != stress_b(L::D, n)) ? ss * (R)2 : ss);
    }
    const R v_eddy = (cs * cs) * sqrt((R)2 * snorm);
    return (R)3 * v_eddy;
}

// KBC shear part. D2Q9: get_shear_D2Q9 (ops/collisions/KBC_.mojo:137-175).
// D3Q19 / D3Q27 have no reference code; the same construction is used there:
// s_q = w_q / (2 cs^4) * sum_ab (c_a c_b - cs^2 delta_ab) dev(Pi)_ab,
// which reproduces the D2Q9 table term by term.
template <class L, class R>
__device__ __forceinline__ void kbc_shear(const R (&pi)[6], R (&s)[L::Q]) {
    if (L::D == 2) {
        const R N = pi[0] - pi[2];
        s[0] = 0;
        s[1] = N * (R)0.25;
        s[2] = N * (R)0.25;
        s[3] = -N * (R)0.25;
        s[4] = -N * (R)0.25;
        s[5] = pi[1] * (R)0.25;
        s[6] = pi[1] * (R)0.25;
        s[7] = -pi[1] * (R)0.25;
        s[8] = -pi[1] * (R)0.25;
        return;
    }
    const R tr3 = (pi[0] + pi[3] + pi[5]) / (R)3;
    const R dev[3] = {pi[0] - tr3, pi[3] - tr3, pi[5] - tr3};
    const R off[3] = {pi[1], pi[2], pi[4]};  // xy, xz, yz
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        R diag = 0, offs = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (L::c(q, d) != 0) diag = diag + dev[d];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const int a = n == 2 ? 1 : 0, b = n == 0 ? 1 : 2;
            const int cc = L::c(q, a) * L::c(q, b);
            if (cc != 0) offs = offs + (cc > 0 ? off[n] : -off[n]);
        }
        s[q] = weight<L, R>(q) * (R)4.5 * (diag + (R)2 * offs);
    }
}

enum Coll { C_SRT = 0, C_KBC = 2 };

// collide, src/lbm/kernels/steps/collide.mojo:23-90 (SRT: ops/collisions/collision.mojo:12-46,
// KBC: ops/collisions/KBC_.mojo:180-238).
template <class L, class R, int COLL, bool LES, bool DDF>
__device__ __forceinline__ void collide(R (&f)[L::Q], R tau, R cs) {
    constexpr int Q = L::Q;
    const R rho = density<L, R, DDF>(f);
    R u[3];
    velocity<L, R>(f, rho, u);
    const R uu = dot_uu<L, R>(u);
    R tau_local = tau;
    if (LES || COLL == C_KBC) {
        R fn[Q];
        R pi[6];
#pragma unroll
        for (int q = 0; q < Q; ++q) fn[q] = f[q] - feq<L, R, DDF>(q, rho, u, uu);
        pi_neq<L, R>(fn, pi);
        if (LES) tau_local = tau_local + smagorinsky_tau<L, R>(pi, rho, tau, cs);
        if (COLL == C_KBC) {
            const R beta = (R)1 / ((R)2 * tau_local);
            const R inv_beta = (R)2 * tau_local;
            R ds[Q];
            kbc_shear<L, R>(pi, ds);
            R sp1 = 0, sp2 = 0;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const R fe = feq<L, R, false>(q, rho, u, uu);
                const R dh = fn[q] - ds[q];
                sp1 = sp1 + ds[q] * dh / fe;
                sp2 = sp2 + dh * dh / fe;
            }
            const R gamma = inv_beta - ((R)2 - inv_beta) * sp1 / (sp2 + (R)1e-32);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const R dh = fn[q] - ds[q];
                f[q] = f[q] - beta * (ds[q] * (R)2 + gamma * dh);
            }
            return;
        }
    }
    const R inv_tau = (R)1 / tau_local;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = f[q] - inv_tau * (f[q] - feq<L, R, DDF>(q, rho, u, uu));
}

// ------------------------------------------------------------ tile context
// Per-lane position, LDS flag tile with halo, and the element deltas of a one
// node move along each axis (periodic, get_adjacent_idx src/lbm/kernels/utils/index.mojo:44-64).
template <class L> struct TileCtx {
    int x, y, z;          // node
    int tid;
    int hx, hxy;          // LDS halo strides
    int center;           // own position in the LDS tile
    uint32_t tile;        // tile id
    int dl_m[3], dl_p[3]; // local-index delta for a -1 / +1 move
    int dt_m[3], dt_p[3]; // tile-index delta for a -1 / +1 move
    int nonfluid;         // tile or halo holds a non-FLUID flag
    uint8_t own;

    __device__ __forceinline__ void init(const Geo& g, const uint8_t* __restrict__ flags, uint8_t* sfl) {
        constexpr int D = L::D;
        tile = tile_of_block(g);
        tid = threadIdx.x;
        const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
        const int lx = tid & (tx - 1), ly = (tid >> g.txb) & (ty - 1), lz = tid >> (g.txb + g.tyb);
        const int bx = (int)(tile % (uint32_t)g.Tx);
        const int byz = (int)(tile / (uint32_t)g.Tx);
        const int by = byz % g.Ty, bz = byz / g.Ty;
        x = (bx << g.txb) + lx;
        y = (by << g.tyb) + ly;
        z = (bz << g.tzb) + lz;

        hx = tx + 2;
        const int hy = ty + 2;
        const int hz = D == 3 ? tz + 2 : 1;
        hxy = hx * hy;
        const int hcells = hxy * hz;
        int any = 0;
        for (int h = tid; h < hcells; h += blockDim.x) {
            const int iz = h / hxy;
            const int r = h - iz * hxy;
            const int iy = r / hx;
            const int ix = r - iy * hx;
            int gx = (bx << g.txb) + ix - 1;
            gx += gx < 0 ? g.nx : 0;
            gx -= gx >= g.nx ? g.nx : 0;
            int gy = (by << g.tyb) + iy - 1;
            gy += gy < 0 ? g.ny : 0;
            gy -= gy >= g.ny ? g.ny : 0;
            int gz = 0;
            if (D == 3) {
                gz = (bz << g.tzb) + iz - 1;
                gz += gz < 0 ? g.nz : 0;
                gz -= gz >= g.nz ? g.nz : 0;
            }
            const uint32_t loc = (uint32_t)((gx & (tx - 1)) + ((gy & (ty - 1)) << g.txb) + ((gz & (tz - 1)) << (g.txb + g.tyb)));
            const uint32_t tl = (uint32_t)((gx >> g.txb) + g.Tx * ((gy >> g.tyb) + g.Ty * (gz >> g.tzb)));
            const uint8_t fl = flags[tl * (uint32_t)g.tile_cells + loc];
            sfl[h] = fl;
            any |= fl;
        }
        nonfluid = __syncthreads_or(any);
        center = (lz + (D == 3 ? 1 : 0)) * hxy + (ly + 1) * hx + (lx + 1);
        own = sfl[center];

        const int Ts[3] = {1, g.Tx, g.Tx * g.Ty};
        const int ls[3] = {1, tx, tx * ty};
        const int lc[3] = {lx, ly, lz};
        const int bc3[3] = {bx, by, bz};
        const int tn[3] = {tx, ty, tz};
        const int Tn[3] = {g.Tx, g.Ty, g.Tz};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            // -1 move
            if (lc[d] > 0) { dl_m[d] = -ls[d]; dt_m[d] = 0; }
            else { dl_m[d] = (tn[d] - 1) * ls[d]; dt_m[d] = (bc3[d] > 0 ? -1 : Tn[d] - 1) * Ts[d]; }
            // +1 move
            if (lc[d] < tn[d] - 1) { dl_p[d] = ls[d]; dt_p[d] = 0; }
            else { dl_p[d] = -(tn[d] - 1) * ls[d]; dt_p[d] = (bc3[d] < Tn[d] - 1 ? 1 : -(Tn[d] - 1)) * Ts[d]; }
        }
    }

    // Flag of the pull neighbour x - c_q (LDS).
    __device__ __forceinline__ uint8_t pull_flag(const uint8_t* sfl, int q) const {
        return sfl[center - L::c(q, 0) - L::c(q, 1) * hx - L::c(q, 2) * hxy];
    }

    // Element index (field with K values per node) of node x + s*c_q, value k.
    __device__ __forceinline__ uint32_t idx(const Geo& g, int K, int q, int s, int k) const {
        int dl = 0, dt = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int c = s * L::c(q, d);
            if (c > 0) { dl += dl_p[d]; dt += dt_p[d]; }
            if (c < 0) { dl += dl_m[d]; dt += dt_m[d]; }
        }
        const uint32_t TK = (uint32_t)g.tile_cells * (uint32_t)K;
        return (tile + (uint32_t)dt) * TK + (uint32_t)(tid + dl) + (uint32_t)k * (uint32_t)g.tile_cells;
    }
    __device__ __forceinline__ uint32_t own_idx(const Geo& g, int K, int k) const {
        return tile * ((uint32_t)g.tile_cells * (uint32_t)K) + (uint32_t)tid + (uint32_t)k * (uint32_t)g.tile_cells;
    }
};

__device__ __forceinline__ bool node_valid(uint8_t fl) {
    return !(((fl & 0x0F) == LBM_FLAG_SOLID) || (fl & LBM_FLAG_EXCLUDE));  // is_valid_thread
}

// Gather: stream (pull / esoteric) + moving wall + equilibrium BC, for one valid node.
// double_buffer_pull_load_f_vec (ops/load_and_store.mojo:164-221), esoteric_pull_load_f_vec (:11-86),
// moving_wall_bc (ops/boundary_condition.mojo:17-83), equilibrium_bc (:87-204).
template <class L, class S, int MODE, bool DDF>
__device__ __forceinline__ void gather(const TileCtx<L>& t, const Geo& g, const uint8_t* sfl,
                                       const typename S::T* __restrict__ f, const typename S::R* __restrict__ bc,
                                       int moving, int eqbc, typename S::R (&fv)[L::Q]) {
    using R = typename S::R;
    constexpr int Q = L::Q, D = L::D;
    const bool plain = !t.nonfluid;
    if (MODE == MODE_DB) {
        if (plain) {
#pragma unroll
            for (int q = 0; q < Q; ++q) fv[q] = S::ld(f + t.idx(g, Q, q, -1, q));
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const uint8_t nf = q ? t.pull_flag(sfl, q) : t.own;
                const uint32_t a = (nf == LBM_FLAG_SOLID) ? t.own_idx(g, Q, opp(q)) : t.idx(g, Q, q, -1, q);
                fv[q] = S::ld(f + a);
            }
        }
    } else {
        fv[0] = S::ld(f + t.own_idx(g, Q, 0));
#pragma unroll
        for (int p = 1; p < Q; p += 2) {
            const int n = p + 1;
            if (MODE == MODE_EP_EVEN) {
                fv[p] = S::ld(f + t.own_idx(g, Q, p));
                fv[n] = S::ld(f + t.idx(g, Q, p, +1, n));
            } else {
                fv[p] = S::ld(f + t.own_idx(g, Q, n));
                fv[n] = S::ld(f + t.idx(g, Q, p, +1, p));
            }
        }
    }
    if (plain) return;
    if (moving) {
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            const uint8_t nf = t.pull_flag(sfl, q);
            if ((nf & 0x0F) == LBM_FLAG_SOLID) {
                R uw[3];
#pragma unroll
                for (int d = 0; d < D; ++d) uw[d] = bc[t.idx(g, D + 1, q, -1, d)];
                const R rw = bc[t.idx(g, D + 1, q, -1, D)];
                fv[q] = fv[q] + (R)6 * weight<L, R>(q) * rw * dot_c<L, R>(uw, q);
            }
        }
    }
    if (eqbc && (t.own & 0x0F) == LBM_FLAG_EQUILIBRIUM) {
        R ub[3];
#pragma unroll
        for (int d = 0; d < D; ++d) ub[d] = bc[t.own_idx(g, D + 1, d)];
        const R rb = bc[t.own_idx(g, D + 1, D)];
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

// double_buffer_kernel (src/lbm/kernels/double_buffer.mojo:25-96) and
// esoteric_pull_kernel[is_even] (src/lbm/kernels/esoteric_pull.mojo:24-111).
template <class L, class S, int MODE, int COLL, bool LES, bool DDF>
__global__ void __launch_bounds__(MaxThreads<L::Q>::value) step_kernel(StepParams<typename S::R> p) {
    using R = typename S::R;
    using T = typename S::T;
    constexpr int Q = L::Q;
    extern __shared__ uint8_t sfl[];
    TileCtx<L> t;
    t.init(p.g, p.flags, sfl);
    if (t.nonfluid && !node_valid(t.own)) return;
    R fv[Q];
    gather<L, S, MODE, DDF>(t, p.g, sfl, (const T*)p.f_in, p.bc, p.moving, p.eqbc, fv);
    collide<L, R, COLL, LES, DDF>(fv, p.tau, p.cs);
    T* fo = (T*)p.f_out;
    if (MODE == MODE_DB) {
#pragma unroll
        for (int q = 0; q < Q; ++q) S::st(fo + t.own_idx(p.g, Q, q), fv[q]);
    } else {
        // esoteric_pull_store_f_vec, ops/load_and_store.mojo:89-159
        S::st(fo + t.own_idx(p.g, Q, 0), fv[0]);
#pragma unroll
        for (int pq = 1; pq < Q; pq += 2) {
            const int n = pq + 1;
            if (MODE == MODE_EP_EVEN) {
                S::st(fo + t.idx(p.g, Q, pq, +1, n), fv[pq]);
                S::st(fo + t.own_idx(p.g, Q, pq), fv[n]);
            } else {
                S::st(fo + t.idx(p.g, Q, pq, +1, pq), fv[pq]);
                S::st(fo + t.own_idx(p.g, Q, n), fv[n]);
            }
        }
    }
}

// calculate_rho_and_velocity (src/lbm/output/velocity.mojo:25-125) and the FP64 moment
// reduction: re-stream + BC (tau = 0), moments; solid nodes (raw flag == SOLID) report bc.
template <class L, class S, int MODE, bool DDF, bool REDUCE>
__global__ void __launch_bounds__(MaxThreads<L::Q>::value) macro_kernel(MacroParams<typename S::R> p) {
    using R = typename S::R;
    using T = typename S::T;
    constexpr int Q = L::Q, D = L::D;
    extern __shared__ uint8_t sfl[];
    __shared__ double red[D + 1][16];
    TileCtx<L> t;
    t.init(p.g, p.flags, sfl);
    R rho = 0, u[3] = {0, 0, 0};
    bool contributes = false;
    if (REDUCE) {
        if (node_valid(t.own)) {
            R fv[Q];
            gather<L, S, MODE, DDF>(t, p.g, sfl, (const T*)p.f, p.bc, p.moving, p.eqbc, fv);
            rho = density<L, R, DDF>(fv);
            velocity<L, R>(fv, rho, u);
            contributes = true;
        }
    } else {
        if (t.own != LBM_FLAG_SOLID) {
            R fv[Q];
            gather<L, S, MODE, DDF>(t, p.g, sfl, (const T*)p.f, p.bc, p.moving, p.eqbc, fv);
            rho = density<L, R, DDF>(fv);
            velocity<L, R>(fv, rho, u);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) u[d] = p.bc[t.own_idx(p.g, D + 1, d)];
            rho = p.bc[t.own_idx(p.g, D + 1, D)];
        }
        const size_t r = ((size_t)t.x * p.g.ny + t.y) * p.g.nz + t.z;
        const size_t ncell = (size_t)p.g.nx * p.g.ny * p.g.nz;
        p.rho[r] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (p.u_layout == LBM_U_XYZD) p.u[r * D + d] = u[d];
            else p.u[(size_t)d * ncell + r] = u[d];
        }
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
    const uint32_t tl = (uint32_t)((idx[0] >> g.txb) + g.Tx * ((idx[1] >> g.tyb) + g.Ty * (idx[2] >> g.tzb)));
    flags[tl * (uint32_t)g.tile_cells + loc] = flag;
    const R vals[4] = {u0, u1, u2, rho};
    for (int k = 0; k <= D; ++k) bc[tl * (uint32_t)g.tile_cells * (uint32_t)(D + 1) + (uint32_t)k * g.tile_cells + loc] = k < D ? vals[k] : rho;
}

}  // namespace lbm

// lbm_physics.hpp -- per-node LBM arithmetic (moments, f_eq, Smagorinsky, KBC,
// SRT) in the reference's operation order. Shared by the kernels and by the
// host-side initialisation (f_eq), and built with -ffp-contract=off so every
// operation rounds as written; f_eq and the SRT update use explicit fused
// multiply-adds (fmadd), which the oracle evaluates the same way.
#pragma once

#include "lbm_common.hpp"

namespace lbm {

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

// a * b + c rounded once (an explicit fused multiply-add; the build otherwise contracts nothing).
__host__ __device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__host__ __device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

// f_eq, src/lbm/kernels/utils/equilibrium.mojo:10-38: w rho (1 + 3 eu + 4.5 eu^2 - 1.5 uu), or with the
// DDF shift w rho (3 eu + 4.5 eu^2 - 1.5 uu) + w (rho - 1), evaluated with three fused multiply-adds:
// A = fma(eu, fma(4.5, eu, 3), -(1.5 uu)), then fma(w rho, A, w rho) or fma(w rho, A, w (rho - 1)).
template <class L, class R, bool DDF>
__host__ __device__ __forceinline__ R feq(int q, R rho, const R (&u)[3], R uu) {
    const R eu = dot_c<L, R>(u, q);
    const R w = weight<L, R>(q);
    const R A = fmadd(eu, fmadd((R)4.5, eu, (R)3), -((R)1.5 * uu));
    const R wr = w * rho;
    return DDF ? fmadd(wr, A, w * (rho - (R)1)) : fmadd(wr, A, wr);
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
__device__ __forceinline__ R strain_norm_sq(const R (&pi)[6], R rho, R tau) {
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
    return snorm;
}

template <class L, class R>
__device__ __forceinline__ R smagorinsky_tau(const R (&pi)[6], R rho, R tau, R cs) {
    const R snorm = strain_norm_sq<L, R>(pi, rho, tau);
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

enum Coll { C_SRT = 0, C_TRT = 1, C_KBC = 2, C_RLBM = 3 };
enum LesMode { LES_OFF = 0, LES_SMAG = 1, LES_MLP = 2 };

// ------------------------------------------------------------ learned closure
// MLP 9 -> 32 -> 32 -> 1 (SiLU, SiLU, softplus) predicting the eddy viscosity nu_t
// from x = (Pi_neq in the stress-pair order xx, xy, xz, yy, yz, zz; rho; |u|;
// Smagorinsky nu_t). Parameter block (floats): W1[k*32+j] (9x32), b1[32],
// W2[k*32+j] (32x32), b2[32], w3[32], b3, then mean[9], inv_std[9] (input
// normalisation), lo[9], hi[9] (training range of the raw features) and
// nu_scale (nu_t = softplus(out) * nu_scale). Outside the training range the
// Smagorinsky value is used. Replaces get_Smagorinsky_LES_tau at
// src/lbm/kernels/steps/collide.mojo:70-73; no reference counterpart.
constexpr int MLP_IN = 9, MLP_H = 32;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_IN * MLP_H, MLP_W2 = MLP_B1 + MLP_H,
              MLP_B2 = MLP_W2 + MLP_H * MLP_H, MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H,
              MLP_NPARAM = MLP_B3 + 1;  // 1409 trainable
constexpr int MLP_MEAN = MLP_NPARAM, MLP_ISTD = MLP_MEAN + MLP_IN, MLP_LO = MLP_ISTD + MLP_IN,
              MLP_HI = MLP_LO + MLP_IN, MLP_NUSCALE = MLP_HI + MLP_IN, MLP_TOTAL = MLP_NUSCALE + 1;  // 1446

__host__ __device__ __forceinline__ float silu(float x) { return x / (1.0f + expf(-x)); }
__host__ __device__ __forceinline__ float softplus(float x) { return x > 20.0f ? x : log1pf(expf(x)); }

// Inference forms used inside collide: fused multiply-adds and the hardware exp2 / log2 /
// reciprocal (about 1 ulp each); the closure's results are compared with its FP64
// restatement to a tolerance, not bit for bit.
__device__ __forceinline__ float silu_fast(float x) {
    return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * x));
}
__device__ __forceinline__ float softplus_fast(float x) {
    return x > 20.0f ? x : 0.6931471805599453f * __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(1.4426950408889634f * x));
}

// Device copy of the parameter block: the 1446 floats, then the two hidden layers' weights
// as bf16 MFMA A operands (v_mfma_f32_32x32x16_bf16, 8 values per lane): W1^T (32 x 16, the 9
// inputs zero-padded) and W2^T as two 32 x 16 K-steps. Built by lbm_set_mlp.
constexpr int MLP_PACK_A1 = 5888;                       // bytes
constexpr int MLP_PACK_A2 = MLP_PACK_A1 + 64 * 16;      // two steps of 64 lanes x 16 bytes
// Then, per wave half g, the 48 per-row values a lane of that half needs, in the order of its
// MFMA result registers j: b1[mfma_row(g, j)], b2[mfma_row(g, j)], w3[mfma_row(g, j)] (16 each).
constexpr int MLP_PACK_ROWS = MLP_PACK_A2 + 2 * 64 * 16;
constexpr int MLP_DEVICE_BYTES = MLP_PACK_ROWS + 2 * 48 * 4;

// Row of the 32 x 32 MFMA result held in register j of a lane in wave half g: the result
// column is the lane index mod 32, rows 8 (j / 4) + 4 g + j % 4.
__host__ __device__ constexpr int mfma_row(int g, int j) { return 8 * (j / 4) + 4 * g + (j % 4); }

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// The closure's nu_t for this lane's node, evaluated by the whole wave: the nodes of lanes
// 0-31 and 32-63 form two 32-column blocks; each hidden layer is W^T (32 x K) times the
// block's inputs (K x 32) on the matrix cores in bf16 with FP32 accumulation, the biases as the
// accumulator inputs. Every lane of the wave must be active. `inside` is false when a raw
// feature lies outside the training range.
typedef const float __attribute__((address_space(4)))* MlpConstPtr;

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// 16 consecutive floats of this lane's wave-half row table (four 16-byte vector loads).
__device__ __forceinline__ f32x16_t mlp_rows16(const f32x4_t* rows, int group) {
    f32x16_t v;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const f32x4_t r = rows[group * 4 + q];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[4 * q + e] = r[e];
    }
    return v;
}

__device__ __forceinline__ float mlp_nu_wave(const float* __restrict__ Pg, const float (&x)[MLP_IN], bool& inside) {
    const MlpConstPtr P = (MlpConstPtr)Pg;
    const int lane = (int)(threadIdx.x & 63u);
    const bool hi = lane >= 32;
    float z[MLP_IN];
    inside = true;
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) {
        inside = inside & (x[k] >= P[MLP_LO + k]) & (x[k] <= P[MLP_HI + k]);
        z[k] = (x[k] - P[MLP_MEAN + k]) * P[MLP_ISTD + k];
    }
    // B operands of layer 1: lanes of half g carry inputs 8 g .. 8 g + 7 of node lane mod 32 of
    // the block. Block 0 (nodes 0-31) takes the low inputs from lanes 0-31 and the high inputs of
    // the same nodes from their partner lanes l ^ 32; block 1 (nodes 32-63) the other way round:
    // one exchange between partner lanes per register.
    bf16x8_t lo8, hi8;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        lo8[i] = (__bf16)z[i];
        hi8[i] = (__bf16)(i == 0 ? z[8] : 0.0f);
    }
    const u32x4_t ulo = __builtin_bit_cast(u32x4_t, lo8), uhi = __builtin_bit_cast(u32x4_t, hi8);
    u32x4_t b0, b1;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const uint32_t t = (uint32_t)__shfl_xor((int)(hi ? ulo[d] : uhi[d]), 32, 64);
        b0[d] = hi ? t : ulo[d];
        b1[d] = hi ? uhi[d] : t;
    }
    const uint8_t* pack = (const uint8_t*)Pg;
    const bf16x8_t a1 = *(const bf16x8_t*)(pack + MLP_PACK_A1 + lane * 16);
    const bf16x8_t a20 = *(const bf16x8_t*)(pack + MLP_PACK_A2 + lane * 16);
    const bf16x8_t a21 = *(const bf16x8_t*)(pack + MLP_PACK_A2 + 1024 + lane * 16);
    const f32x4_t* rows_base = (const f32x4_t*)(pack + MLP_PACK_ROWS + (hi ? 192 : 0));
    // one 32-node block at a time keeps a single block's accumulators live
    float part[2];
#pragma unroll 1
    for (int blk = 0; blk < 2; ++blk) {
        const u32x4_t bin = blk ? b1 : b0;
        // reload the row tables in each block (cache hits) rather than keep 48 values live
        uintptr_t rp = (uintptr_t)rows_base;
        __asm__ volatile("" : "+v"(rp));
        const f32x4_t* rows = (const f32x4_t*)rp;
        const f32x16_t d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, __builtin_bit_cast(bf16x8_t, bin),
                                                                    mlp_rows16(rows, 0), 0, 0, 0);
        // layer 2: the K index of step s, half g, element i is hidden unit mfma_row(g, 8 s + i),
        // which is exactly where layer 1 left it; W2 is packed to match
        bf16x8_t h0, h1;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            h0[i] = (__bf16)silu_fast(d1[i]);
            h1[i] = (__bf16)silu_fast(d1[8 + i]);
        }
        f32x16_t d2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a20, h0, mlp_rows16(rows, 1), 0, 0, 0);
        d2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a21, h1, d2, 0, 0, 0);
        // layer 3: partial dot product over this lane's 16 rows
        const f32x16_t w3 = mlp_rows16(rows, 2);
        float sacc = 0.0f;
#pragma unroll
        for (int j = 0; j < 16; ++j) sacc = __builtin_fmaf(w3[j], silu_fast(d2[j]), sacc);
        part[blk] = sacc;
    }
    const float sa = part[0], sb = part[1];
    // column l of block 0 is on lanes l and l + 32, column l of block 1 on lanes l and l + 32 too
    const float t = __shfl_xor(hi ? sa : sb, 32, 64);
    const float out = P[MLP_B3] + ((hi ? sb : sa) + t);
    return softplus_fast(out) * P[MLP_NUSCALE];
}

// collide, src/lbm/kernels/steps/collide.mojo:23-90 (SRT: ops/collisions/collision.mojo:12-46,
// KBC: ops/collisions/KBC_.mojo:180-238).
template <class L, class R, int COLL, int LES, bool DDF>
// inv_tau_base is 1/tau rounded in R (computed once on the host; equal to the
// per-node (R)1 / tau the reference evaluates when no LES term changes tau).
// LES: LES_OFF, LES_SMAG (the reference's Smagorinsky term) or LES_MLP (the
// learned closure with parameter block mlp, Smagorinsky outside its range).
__device__ __forceinline__ void collide(R (&f)[L::Q], R tau, R inv_tau_base, R cs, const float* mlp = nullptr) {
    constexpr int Q = L::Q;
    const R rho = density<L, R, DDF>(f);
    R u[3];
    velocity<L, R>(f, rho, u);
    const R uu = dot_uu<L, R>(u);
    R tau_local = tau;
    if (LES || COLL == C_KBC || COLL == C_RLBM) {
        R fn[Q];
        R pi[6];
#pragma unroll
        for (int q = 0; q < Q; ++q) fn[q] = f[q] - feq<L, R, DDF>(q, rho, u, uu);
        pi_neq<L, R>(fn, pi);
        if (LES == LES_SMAG) tau_local = tau_local + smagorinsky_tau<L, R>(pi, rho, tau, cs);
        if (LES == LES_MLP) {
            const R tau_s = smagorinsky_tau<L, R>(pi, rho, tau, cs);
            float x[MLP_IN];
#pragma unroll
            for (int n = 0; n < 6; ++n) x[n] = n < L::D * (L::D + 1) / 2 ? (float)pi[n] : 0.0f;
            x[6] = (float)rho;
            x[7] = (float)sqrt(uu);
            x[8] = (float)(tau_s / (R)3);
            bool inside;
            const float nu = mlp_nu_wave(mlp, x, inside);
            tau_local = inside ? tau_local + (R)3 * (R)nu : tau_local + tau_s;
        }
        if (COLL == C_RLBM) {
            // RLBM, ops/collisions/collision.mojo:168-223: f_q = f_eq + (1 - 1/tau) w_q/(2 cs^4) H_q : Pi_neq
            // with H_q,ab = c_a c_b - cs^2 delta_ab (off-diagonal pairs counted twice)
            const R inv_tau = (R)1 / tau_local;
            const R keep = (R)1 - inv_tau;
            const R cs2 = (R)(1.0 / 3.0);
            constexpr int NS = L::D * (L::D + 1) / 2;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                bool first = true;
                R dot = 0;
#pragma unroll
                for (int n = 0; n < NS; ++n) {
                    const int a = stress_a(L::D, n), b = stress_b(L::D, n);
                    const int cc = L::c(q, a) * L::c(q, b);
                    R h;
                    if (a == b) h = (R)cc - cs2;
                    else if (cc == 0) continue;
                    else h = (R)cc * (R)2;
                    const R t = h * pi[n];
                    dot = first ? t : dot + t;
                    first = false;
                }
                const R k = weight<L, R>(q) / (R)(2.0 / 9.0);
                f[q] = feq<L, R, DDF>(q, rho, u, uu) + keep * (k * dot);
            }
            return;
        }
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
    if (COLL == C_TRT) {
        // TRT, ops/collisions/collision.mojo:49-113, with tau_asymm = 1/2 + Lambda / (tau - 1/2),
        // Lambda = 3/16 (steps/collide.mojo:83-85): symmetric and antisymmetric parts of each
        // opposite pair (q, q+1) relax with their own rates; the rest population relaxes with tau.
        const R tau_a = (R)0.5 + (R)0.1875 / (tau_local - (R)0.5);
        const R inv_ts = (R)1 / tau_local;
        const R inv_ta = (R)1 / tau_a;
        f[0] = f[0] - inv_ts * (f[0] - feq<L, R, DDF>(0, rho, u, uu));
#pragma unroll
        for (int q = 1; q < Q; q += 2) {
            const int o = q + 1;
            const R fs = (f[q] + f[o]) * (R)0.5;
            const R fa = (f[q] - f[o]) * (R)0.5;
            const R eq = feq<L, R, DDF>(q, rho, u, uu);
            const R eo = feq<L, R, DDF>(o, rho, u, uu);
            const R es = (eq + eo) * (R)0.5;
            const R ea = (eq - eo) * (R)0.5;
            f[q] = f[q] - (inv_ts * (fs - es) + inv_ta * (fa - ea));
            f[o] = f[o] - (inv_ts * (fs - es) + inv_ta * ((-fa) - (-ea)));
        }
        return;
    }
    // SRT, ops/collisions/collision.mojo:12-46: f - (1/tau)(f - f_eq) as one fused multiply-add
    const R inv_tau = LES != LES_OFF ? (R)1 / tau_local : inv_tau_base;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = fmadd(-inv_tau, f[q] - feq<L, R, DDF>(q, rho, u, uu), f[q]);
}

}  // namespace lbm

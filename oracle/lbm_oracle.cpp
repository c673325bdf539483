/*
 * lbm_oracle.cpp -- CPU restatement of mojo_lbm's stream-collide path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product path never calls it.
 *
 * It restates, per node and in the reference's operation order, the functions
 * of the reference's double-buffer and esoteric-pull kernels. The reference is
 * Mojo GPU code (no CPU stepping path); its per-cell building blocks are written
 * CPU/GPU-generic (src/lbm/kernels/ReadMe.md:5,11,19), which is what this file
 * follows. Arithmetic is templated on the compute type T (float or double);
 * populations are stored as float, double or Float16C codes. Build with
 * -ffp-contract=off so every operation is rounded as written.
 *
 * Pinning (what each function is checked against, in tests/):
 *   lattice tables ........ moment identities (sum w = 1, sum w c = 0, sum w c c = I/3), opposite pairs
 *   f16c_encode/decode .... exhaustive 65,536-code table against the format definition
 *                           (Lehmann et al., PRE 106, 015308 (2022)), round-to-nearest brute force
 *   addr (tiled layout) ... bijection + explicit formula
 *   pull streaming ........ a periodic delta moves one node per step per direction
 *   halfway bounce-back ... a population reflected by a wall returns in one step
 *   moving wall ........... momentum injection closed form; Couette profile
 *   SRT / f_eq ............ f_eq moments reproduce rho and rho u; f = f_eq is a fixed point
 *   DDF shift ............. shifted and unshifted FP64 runs agree to rounding
 *   esoteric pull ......... bitwise equality with double buffer in periodic boxes; slot ownership
 *   Smagorinsky ........... closed form on a synthetic shear f_neq
 *   KBC ................... gamma = 2 reduces to SRT; conservation; D2Q9 shear table
 *   equilibrium BC ........ node reset to f_eq of prescribed rho, u; with free (NaN) u and/or rho the
 *                           local moments rebuilt node by node with out-of-grid pulls at the rest value
 *                           (w_q, 0 with DDF), on both x faces incl. edges and corners
 *   FP64 totals ........... uniform states give N rho0 and N rho0 u0; only non-SOLID, non-EXCLUDE nodes
 *                           count; the lid's momentum injection closed form; equal to fsum of the fields
 *   initialize_f_from_func  moments of f_eq(rho, u(x)); Pi_neq = -2 rho cs^2 (tau - 1) S of the gradient
 *                           term; esoteric store = the double-buffer fill moved by the odd pattern
 *   Q-criterion ........... solid-body rotation (Q = |Omega|^2), uniform strain (Q = -9 eps^2 / (4 tau^2)),
 *                           linear shear against a wall at half distance, one-sided edges
 *   LDC D2Q9 Re=100 ....... Ghia et al. centreline fixture (benchmarks/LDC/*), loose physics check
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/lbm_hip.h"

namespace {

/* ---------------------------------------------------------------- lattices
 * Velocity sets in reference order, opposites adjacent (q odd, q+1):
 * get_D2Q9 src/lbm/lattice.mojo:302-351, get_D3Q19 :221-299, get_D3Q27 :122-218.
 */
struct Lat {
    int D, Q;
    int c[27][3];
    int wnum[27], wden[27];  // weight = wnum / wden, evaluated in T
    int opp[27];
};

const int C27[27][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0},
    {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
    {1, 1, 1}, {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1},
    {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};

const int C9[9][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                      {1, 1, 0},  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}};

Lat make_lat(int D, int Q) {
    Lat L{};
    L.D = D;
    L.Q = Q;
    for (int q = 0; q < Q; ++q) {
        for (int d = 0; d < 3; ++d) L.c[q][d] = (Q == 9) ? C9[q][d] : C27[q][d];
        int nnz = std::abs(L.c[q][0]) + std::abs(L.c[q][1]) + std::abs(L.c[q][2]);
        if (Q == 9) {
            const int num[3] = {4, 1, 1}, den[3] = {9, 9, 36};
            L.wnum[q] = num[nnz]; L.wden[q] = den[nnz];
        } else if (Q == 19) {
            const int num[3] = {1, 1, 1}, den[3] = {3, 18, 36};
            L.wnum[q] = num[nnz]; L.wden[q] = den[nnz];
        } else {
            const int num[4] = {8, 2, 1, 1}, den[4] = {27, 27, 54, 216};
            L.wnum[q] = num[nnz]; L.wden[q] = den[nnz];
        }
    }
    // _get_opposite_indices, src/lbm/lattice.mojo:76-91: search for -c.
    for (int q = 0; q < Q; ++q)
        for (int k = 0; k < Q; ++k)
            if (L.c[k][0] == -L.c[q][0] && L.c[k][1] == -L.c[q][1] && L.c[k][2] == -L.c[q][2]) {
                L.opp[q] = k;
                break;
            }
    return L;
}

/* ---------------------------------------------------------------- Float16C
 * 1-4-11 format of Lehmann et al. (2022) (cited at src/utils/custom_fp.mojo:24-31):
 * code = s | e' << 11 | m, value = m * 2^-25 for e' = 0, (2048 + m) * 2^(e'-26) otherwise.
 * Encoding rounds to the nearest code with ties away from zero and saturates to
 * 0x7FFF for magnitudes that round to 2 or more, infinities and NaN
 * (custom_fp.mojo:71-100 behaviour). Written here from that definition.
 */
float f16c_decode(uint16_t h) {
    int e = (h >> 11) & 0xF;
    int m = h & 0x7FF;
    double mag = (e == 0) ? std::ldexp((double)m, -25) : std::ldexp((double)(2048 + m), e - 26);
    float v = (float)mag;  // exact: at most 12 significant bits
    return (h & 0x8000) ? -v : v;
}

uint16_t f16c_encode(float x) {
    uint32_t bits;
    std::memcpy(&bits, &x, 4);
    uint16_t s = (uint16_t)((bits >> 16) & 0x8000);
    double a = std::fabs((double)x);
    if (!(a < 2.0)) return s | 0x7FFF;  // >= 2, inf, NaN
    // Exponent of the code grid containing a: E = max(floor(log2 a), -14).
    int E = -14;
    if (a >= std::ldexp(1.0, -14)) {
        int ex;
        std::frexp(a, &ex);  // a = f * 2^ex, f in [0.5, 1)
        E = ex - 1;
    }
    double t = std::ldexp(a, 11 - E);      // exact
    double r = std::floor(t + 0.5);        // round half away from zero (t >= 0)
    long code = ((long)(E + 14) << 11) + (long)r;
    if (code > 0x7FFF) code = 0x7FFF;
    return s | (uint16_t)code;
}

/* ---------------------------------------------------------------- layout
 * Blocked product of col-major tile (tx,ty,tz,K) and col-major tiler (Tx,Ty,Tz,1):
 * src/lbm/layoutstruct.mojo:45-64, inverse idx_to_ijk src/lbm/geometry/interpolated_BB.mojo:24-43.
 */
struct Geo {
    int nx, ny, nz, tx, ty, tz, Tx, Ty, Tz;
    size_t tile_cells;
    size_t addr(int x, int y, int z, int k, int K) const {
        size_t lx = x % tx, ly = y % ty, lz = z % tz;
        size_t bx = x / tx, by = y / ty, bz = z / tz;
        return lx + tx * (ly + ty * (lz + tz * (size_t)k)) + tile_cells * K * (bx + Tx * (by + Ty * bz));
    }
};

Geo make_geo(const lbm_desc* d) {
    Geo g;
    g.nx = d->nx; g.ny = d->ny; g.nz = d->nz;
    g.tx = d->tx; g.ty = d->ty; g.tz = d->tz;
    g.Tx = d->nx / d->tx; g.Ty = d->ny / d->ty; g.Tz = d->nz / d->tz;
    g.tile_cells = (size_t)d->tx * d->ty * d->tz;
    return g;
}

/* get_adjacent_idx, src/lbm/kernels/utils/index.mojo:44-64: floor-mod wrap, axes >= D stay 0. */
inline void adjacent(const Geo& g, int D, const int idx[3], const int c[3], int shift, int out[3]) {
    const int n[3] = {g.nx, g.ny, g.nz};
    for (int d = 0; d < 3; ++d) {
        if (d < D) {
            int v = (idx[d] + shift * c[d]) % n[d];
            if (v < 0) v += n[d];
            out[d] = v;
        } else {
            out[d] = 0;
        }
    }
}

/* ---------------------------------------------------------------- storage */
template <class S, class T> struct Store;
template <class T> struct Store<float, T> {
    static T load(const float* p, size_t i) { return (T)p[i]; }
    static void store(float* p, size_t i, T v) { p[i] = (float)v; }
};
template <class T> struct Store<double, T> {
    static T load(const double* p, size_t i) { return (T)p[i]; }
    static void store(double* p, size_t i, T v) { p[i] = (double)v; }
};
/* load_f / store_f with Float16C, src/lbm/kernels/utils/load_and_store.mojo:17-110:
 * the stored value is encoded from its fp32 value. */
template <class T> struct Store<uint16_t, T> {
    static T load(const uint16_t* p, size_t i) { return (T)f16c_decode(p[i]); }
    static void store(uint16_t* p, size_t i, T v) { p[i] = f16c_encode((float)v); }
};

inline bool flag_is_solid(uint8_t f) { return (f & 0x0F) == LBM_FLAG_SOLID; }     // Flags.is[SOLID]
inline bool flag_is_eq(uint8_t f) { return (f & 0x0F) == LBM_FLAG_EQUILIBRIUM; }  // Flags.is[EQUILIBRIUM]
inline bool flag_excluded(uint8_t f) { return (f & LBM_FLAG_EXCLUDE) != 0; }     // Flags.has[EXCLUDE]

/* ---------------------------------------------------------------- per-cell math
 * All in compute type T, in the order the reference writes them.
 */
template <class T> struct Cell {
    const Lat& L;
    T w[27];
    explicit Cell(const Lat& lat) : L(lat) {
        for (int q = 0; q < L.Q; ++q) w[q] = (T)L.wnum[q] / (T)L.wden[q];
    }

    /* Vector.dot, src/utils/vector.mojo:205-217: out = 0; out += a_i * b_i in order. */
    T dot_c(const T* u, int q) const {
        T out = 0;
        for (int d = 0; d < L.D; ++d) out += u[d] * (T)L.c[q][d];
        return out;
    }
    T dot_uu(const T* u) const {
        T out = 0;
        for (int d = 0; d < L.D; ++d) out += u[d] * u[d];
        return out;
    }

    /* f_eq, src/lbm/kernels/utils/equilibrium.mojo:10-38: w rho (1 + 3 eu + 4.5 eu^2 - 1.5 uu), or
     * w rho (3 eu + 4.5 eu^2 - 1.5 uu) + w (rho - 1) with the DDF shift, with the polynomial and the
     * last multiply-add each rounded once (std::fma): A = fma(eu, fma(4.5, eu, 3), -(1.5 uu)),
     * f_eq = fma(w rho, A, w rho) or fma(w rho, A, w (rho - 1)). */
    T feq(int q, T rho, const T* u, T uu, bool ddf) const {
        T eu = dot_c(u, q);
        const T A = std::fma(eu, std::fma((T)4.5, eu, (T)3), -((T)1.5 * uu));
        const T wr = w[q] * rho;
        return ddf ? std::fma(wr, A, w[q] * (rho - (T)1)) : std::fma(wr, A, wr);
    }

    /* get_density, src/lbm/kernels/utils/moment.mojo:17-41 (Vector.sum, vector.mojo:219-224). */
    T density(const T* f, bool ddf) const {
        T s = 0;
        for (int q = 0; q < L.Q; ++q) s += f[q];
        return ddf ? s + (T)1 : s;
    }
    /* get_velocity, src/lbm/kernels/utils/moment.mojo:45-76. */
    void velocity(const T* f, T rho, T* u) const {
        for (int d = 0; d < L.D; ++d) u[d] = 0;
        for (int q = 0; q < L.Q; ++q)
            for (int d = 0; d < L.D; ++d) u[d] += f[q] * (T)L.c[q][d];
        for (int d = 0; d < L.D; ++d) u[d] /= rho;
    }

    int nstress() const { return L.D * (L.D + 1) / 2; }
    /* get_stress_indices, src/lbm/lattice.mojo:354-398. */
    void stress_pair(int n, int& a, int& b) const {
        static const int p2[3][2] = {{0, 0}, {0, 1}, {1, 1}};
        static const int p3[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        if (L.D == 2) { a = p2[n][0]; b = p2[n][1]; }
        else { a = p3[n][0]; b = p3[n][1]; }
    }
    /* get_non_eq_second_order_moment + get_Qiab, src/lbm/kernels/utils/moment.mojo:80-150. */
    void pi_neq(const T* fneq, T* pi) const {
        for (int n = 0; n < nstress(); ++n) {
            int a, b;
            stress_pair(n, a, b);
            T s = 0;
            for (int q = 0; q < L.Q; ++q) s += fneq[q] * (T)L.c[q][a] * (T)L.c[q][b];
            pi[n] = s;
        }
    }

    /* Smagorinsky tau_eddy: get_strain_rate_tensor (moment.mojo:154-176),
     * get_strain_rate_tensor_norm_squared (moment.mojo:182-215),
     * get_Smagorinsky_LES_tau (src/lbm/kernels/ops/turbulence.mojo:12-38).
     * The strain rate uses the base tau (steps/collide.mojo:69). */
    T smagorinsky_tau(const T* pi, T rho, T tau, T Cs) const {
        const T cs2 = (T)1 / (T)3;
        T k = (T)(-1) / ((T)2 * rho * cs2 * tau);
        T snorm = 0;
        for (int n = 0; n < nstress(); ++n) {
            int a, b;
            stress_pair(n, a, b);
            T s = pi[n] * k;
            T ss = s * s;
            snorm += (a != b) ? ss * (T)2 : ss;
        }
        T v_eddy = (Cs * Cs) * std::sqrt((T)2 * snorm);
        return (T)3 * v_eddy;
    }

    /* KBC shear part. D2Q9: get_shear_D2Q9, src/lbm/kernels/ops/collisions/KBC_.mojo:137-175.
     * D3Q19/D3Q27 (no reference code): the same construction, the second-order Hermite
     * projection of the deviatoric non-equilibrium stress,
     *   s_q = w_q / (2 cs^4) * sum_ab (c_a c_b - cs^2 delta_ab) * dev(Pi)_ab,
     * which reproduces the D2Q9 table term by term. */
    void kbc_shear(const T* pi, T* s) const {
        if (L.D == 2) {
            T N = pi[0] - pi[2];
            s[0] = 0;
            s[1] = N * (T)0.25;
            s[2] = N * (T)0.25;
            s[3] = -N * (T)0.25;
            s[4] = -N * (T)0.25;
            s[5] = pi[1] * (T)0.25;
            s[6] = pi[1] * (T)0.25;
            s[7] = -pi[1] * (T)0.25;
            s[8] = -pi[1] * (T)0.25;
            return;
        }
        // dev(Pi) for D = 3: Pi - tr(Pi)/3 I. H:dev = sum_ab c_a c_b dev_ab (the -cs^2 tr(dev)
        // term vanishes): diagonal terms where c_a != 0, off-diagonal terms (xy, xz, yz) with the
        // sign of c_a c_b, each summed in that order; w / (2 cs^4) = 4.5 w.
        T tr3 = (pi[0] + pi[3] + pi[5]) / (T)3;
        const T dev[3] = {pi[0] - tr3, pi[3] - tr3, pi[5] - tr3};
        const T off[3] = {pi[1], pi[2], pi[4]};
        const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
        for (int q = 0; q < L.Q; ++q) {
            T diag = 0, offs = 0;
            for (int dd = 0; dd < 3; ++dd)
                if (L.c[q][dd] != 0) diag = diag + dev[dd];
            for (int n = 0; n < 3; ++n) {
                int cc = L.c[q][pa[n]] * L.c[q][pb[n]];
                if (cc != 0) offs = offs + (cc > 0 ? off[n] : -off[n]);
            }
            s[q] = w[q] * (T)4.5 * (diag + (T)2 * offs);
        }
    }
};

/* ---------------------------------------------------------------- one node */
/* Learned LES closure (no reference counterpart; it replaces get_Smagorinsky_LES_tau at
 * src/lbm/kernels/steps/collide.mojo:70-73): an MLP 9 -> 32 -> 32 -> 1 with SiLU hidden units
 * and a softplus output times nu_scale gives nu_t from the raw features x (Pi_neq in stress-pair
 * order, rho, |u|, Smagorinsky nu_t), normalised by (x - mean) * inv_std. Evaluated here in FP64
 * from the float features, with the two hidden layers' inputs and weights rounded to bf16 as the
 * library's matrix-core evaluation takes them; a feature outside [lo, hi] selects the Smagorinsky
 * value. Parameter
 * block: W1 (9x32, k-major), b1, W2 (32x32, k-major), b2, w3, b3, mean, inv_std, lo, hi, nu_scale. */
const float* g_mlp = nullptr;

/* bf16 rounding (nearest even) of a float: the precision of the hidden layers' inputs and weights
 * when the library evaluates them on the matrix cores. */
double bf16r(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}

bool mlp_nu(const float* P, const float x[9], double* nu) {
    const int W1 = 0, B1 = 288, W2 = 320, B2 = 1344, W3 = 1376, B3 = 1408, MEAN = 1409, ISTD = 1418, LO = 1427,
              HI = 1436, NUS = 1445;
    double z[9];
    for (int k = 0; k < 9; ++k) {
        if (!(x[k] >= P[LO + k] && x[k] <= P[HI + k])) return false;
        z[k] = bf16r((x[k] - P[MEAN + k]) * P[ISTD + k]);
    }
    double h1[32];
    for (int j = 0; j < 32; ++j) {
        double a = P[B1 + j];
        for (int k = 0; k < 9; ++k) a += bf16r(P[W1 + k * 32 + j]) * z[k];
        h1[j] = bf16r((float)(a / (1.0 + std::exp(-a))));
    }
    double out = P[B3];
    for (int j = 0; j < 32; ++j) {
        double a = P[B2 + j];
        for (int k = 0; k < 32; ++k) a += bf16r(P[W2 + k * 32 + j]) * h1[k];
        out += (double)P[W3 + j] * (a / (1.0 + std::exp(-a)));
    }
    const double sp = out > 20.0 ? out : std::log1p(std::exp(out));
    *nu = sp * P[NUS];
    return true;
}

template <class S, class T> struct Stepper {
    const lbm_desc& d;
    Lat L;
    Geo g;
    Cell<T> cell;
    T Cs;
    Stepper(const lbm_desc& desc) : d(desc), L(make_lat(desc.D, desc.Q)), g(make_geo(&desc)), cell(L) {
        Cs = (T)desc.smagorinsky_cs;
    }
    int K() const { return L.Q; }
    int Kbc() const { return L.D + 1; }

    /* set_adjacent_flags, src/lbm/kernels/ops/load_and_store.mojo:270-291 (pull neighbours x - c). */
    void gather_flags(const uint8_t* flags, const int idx[3], uint8_t pf[27]) const {
        pf[0] = flags[g.addr(idx[0], idx[1], idx[2], 0, 1)];
        for (int q = 1; q < L.Q; ++q) {
            int n[3];
            adjacent(g, L.D, idx, L.c[q], -1, n);
            pf[q] = flags[g.addr(n[0], n[1], n[2], 0, 1)];
        }
    }

    /* double_buffer_pull_load_f_vec, src/lbm/kernels/ops/load_and_store.mojo:164-221:
     * raw-flag compare against SOLID selects halfway bounce-back (:216). */
    void load_db(const S* f, const uint8_t pf[27], const int idx[3], T* fv) const {
        for (int q = 0; q < L.Q; ++q) {
            int n[3];
            adjacent(g, L.D, idx, L.c[q], -1, n);
            if (pf[q] == LBM_FLAG_SOLID)
                fv[q] = Store<S, T>::load(f, g.addr(idx[0], idx[1], idx[2], L.opp[q], K()));
            else
                fv[q] = Store<S, T>::load(f, g.addr(n[0], n[1], n[2], q, K()));
        }
    }

    /* esoteric_pull_load_f_vec, src/lbm/kernels/ops/load_and_store.mojo:11-86. */
    void load_ep(const S* f, const int idx[3], bool even, T* fv) const {
        fv[0] = Store<S, T>::load(f, g.addr(idx[0], idx[1], idx[2], 0, K()));
        for (int p = 1; p < L.Q - 1; p += 2) {
            int nq = p + 1;
            int s[3];
            if (even) {
                adjacent(g, L.D, idx, L.c[nq], -1, s);  // pull index of the negative direction
                fv[p] = Store<S, T>::load(f, g.addr(idx[0], idx[1], idx[2], p, K()));
                fv[nq] = Store<S, T>::load(f, g.addr(s[0], s[1], s[2], nq, K()));
            } else {
                adjacent(g, L.D, idx, L.c[p], 1, s);  // push index of the positive direction
                fv[p] = Store<S, T>::load(f, g.addr(idx[0], idx[1], idx[2], nq, K()));
                fv[nq] = Store<S, T>::load(f, g.addr(s[0], s[1], s[2], p, K()));
            }
        }
    }

    /* esoteric_pull_store_f_vec, src/lbm/kernels/ops/load_and_store.mojo:89-159. */
    void store_ep(S* f, const int idx[3], bool even, const T* fv) const {
        Store<S, T>::store(f, g.addr(idx[0], idx[1], idx[2], 0, K()), fv[0]);
        for (int nq = 2; nq < L.Q; nq += 2) {
            int p = nq - 1;
            int s[3];
            if (even) {
                adjacent(g, L.D, idx, L.c[nq], -1, s);
                Store<S, T>::store(f, g.addr(s[0], s[1], s[2], nq, K()), fv[p]);
                Store<S, T>::store(f, g.addr(idx[0], idx[1], idx[2], p, K()), fv[nq]);
            } else {
                adjacent(g, L.D, idx, L.c[p], 1, s);
                Store<S, T>::store(f, g.addr(s[0], s[1], s[2], p, K()), fv[p]);
                Store<S, T>::store(f, g.addr(idx[0], idx[1], idx[2], nq, K()), fv[nq]);
            }
        }
    }

    /* Interpolated bounce-back links (linear Bouzidi, Firdaouss & Lallemand 2001; the formulas of
     * linkwise_bounceback_kernel, src/lbm/geometry/interpolated_BB.mojo:119-140, and
     * nodewise_bounceback_kernel, src/lbm/geometry/nodewise_interpolated_BB.mojo:140-178). Links live on
     * FLUID_SOLID_BOUNDARY nodes: record r holds the node's flag-layout index link_nodes[r] (sorted) and,
     * for q = 1..Q-1, link_q[r*(Q-1) + q-1] = the fraction of the link from x towards the SOLID pull
     * neighbour x - c_q that lies in the fluid (negative: plain halfway bounce-back), raised to q_clamp
     * (nodewise_interpolated_BB.mojo:151). The population entering x along c_q is rebuilt from
     * a = fv[q] = f*(x, opp q) (the one that hit the wall, delivered by halfway bounce-back) and
     *   qd <= 1/2: b = fv[opp q],  f_q = (2 qd) a + (1 - 2 qd) b,
     *   qd >  1/2: b = f*(x, q),   f_q = (0.5 / qd) a + ((2 qd - 1) / (2 qd)) b.
     * fv[opp q] is the population pulled from x + c_q: f*(x + c_q, opp q) in the fluid, its halfway
     * bounce-back f*(x, q) in a one-node gap, or -- when the opposite direction is itself a link --
     * that link's rebuilt value if it came first. Links are taken in the reference's order: its link
     * bit q_ref rebuilds f_vec[opp q_ref], so the rebuilt directions run opp(1), 1, opp(3), 3, ...
     * Each link adds the momentum exchange c_{opp q} (a + f_q) (+ 2 w_q with DDF shifting) to the
     * force on the solid. The mass the interpolation adds, dm = sum (f_q - a), is taken from the rest
     * population (the reference's mass fix, nodewise_interpolated_BB.mojo:170-176, adds dm; subtracting
     * it is what leaves the node the mass halfway bounce-back gives it). Returns a bit mask of the
     * rebuilt directions (no moving-wall term there). */
    const uint32_t* link_nodes = nullptr;
    const float* link_q = nullptr;
    uint32_t n_links = 0;
    float q_clamp = 0.0f;

    uint32_t bouzidi(T* fv, const uint8_t pf[27], const S* f, const uint8_t* flags, const int idx[3],
                     double* force) const {
        if (!n_links || (pf[0] & 0x0F) != LBM_FLAG_FLUID_SOLID_BOUNDARY) return 0;
        const uint32_t node = (uint32_t)g.addr(idx[0], idx[1], idx[2], 0, 1);
        uint32_t lo = 0, hi = n_links;
        while (lo < hi) {
            uint32_t mid = (lo + hi) / 2;
            if (link_nodes[mid] < node) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= n_links || link_nodes[lo] != node) return 0;
        uint32_t done = 0;
        T dm = 0;  // mass the interpolation adds, taken back from the rest population
        for (int k = 1; k < L.Q; ++k) {
            const int q = L.opp[k];  // opp(1), 1, opp(3), 3, ...: the reference's link-bit order
            T qd = (T)link_q[(size_t)lo * (L.Q - 1) + (q - 1)];
            if (!(qd >= (T)0) || pf[q] != LBM_FLAG_SOLID) continue;
            qd = std::max(qd, (T)q_clamp);
            const int oq = L.opp[q];
            const T a = fv[q];
            T fb;
            if (qd <= (T)0.5) {
                const T b = fv[oq];
                fb = (T)2 * qd * a + ((T)1 - (T)2 * qd) * b;
            } else {
                const T b = Store<S, T>::load(f, g.addr(idx[0], idx[1], idx[2], q, K()));
                fb = (T)0.5 / qd * a + ((T)2 * qd - (T)1) / ((T)2 * qd) * b;
            }
            fv[q] = fb;
            dm = dm + (fb - a);
            done |= 1u << q;
            T s = a + fb;
            if (d.ddf_shift) s = s + (T)2 * cell.w[q];
            for (int dd = 0; dd < L.D; ++dd)
                if (L.c[q][dd]) force[dd] += (double)(L.c[q][dd] > 0 ? -s : s);
        }
        if (done) fv[0] = fv[0] - dm;
        return done;
    }

    /* moving_wall_bc, src/lbm/kernels/ops/boundary_condition.mojo:17-83:
     * f_q += 2*3*w_q*rho_w*(c_q . u_w) for q >= 1 whose pull neighbour is SOLID (masked enum). */
    void moving_wall(T* fv, const uint8_t pf[27], const T* bc, const int idx[3], uint32_t skip = 0) const {
        for (int q = 1; q < L.Q; ++q) {
            if ((skip >> q) & 1u) continue;
            if (!flag_is_solid(pf[q])) continue;
            int n[3];
            adjacent(g, L.D, idx, L.c[q], -1, n);
            T u[3];
            for (int dd = 0; dd < L.D; ++dd) u[dd] = bc[g.addr(n[0], n[1], n[2], dd, Kbc())];
            T rho = bc[g.addr(n[0], n[1], n[2], L.D, Kbc())];
            fv[q] += (T)6 * cell.w[q] * rho * cell.dot_c(u, q);
        }
    }

    /* equilibrium_bc + get_density_and_velocity_for_eq_BC,
     * src/lbm/kernels/ops/boundary_condition.mojo:87-204: out-of-bounds pulls count as the rest value. */
    void equilibrium(T* fv, const uint8_t pf[27], const T* bc, const int idx[3]) const {
        if (!flag_is_eq(pf[0])) return;
        bool ddf = d.ddf_shift != 0;
        T ub[3];
        for (int dd = 0; dd < L.D; ++dd) ub[dd] = bc[g.addr(idx[0], idx[1], idx[2], dd, Kbc())];
        T rb = bc[g.addr(idx[0], idx[1], idx[2], L.D, Kbc())];
        const int n[3] = {g.nx, g.ny, g.nz};
        T vel[3] = {0, 0, 0};
        T rho = 0;
        for (int q = 0; q < L.Q; ++q) {
            T rest = ddf ? (T)0 : cell.w[q];
            bool oob = false;
            for (int i = 0; i < 3; ++i) {
                if (i < L.D) {
                    int pi = idx[i] - L.c[q][i];
                    oob = (pi < 0 || pi >= n[i]) || oob;
                }
            }
            T fq = oob ? rest : fv[q];
            rho += fq;
            for (int dd = 0; dd < L.D; ++dd) vel[dd] += fq * (T)L.c[q][dd];
        }
        if (ddf) rho += (T)1;
        for (int dd = 0; dd < L.D; ++dd) vel[dd] /= rho;
        const T* ul = std::isnan(ub[0]) ? vel : ub;
        T rl = std::isnan(rb) ? rho : rb;
        T uu = cell.dot_uu(ul);
        T feq[27];
        for (int q = 0; q < L.Q; ++q) feq[q] = cell.feq(q, rl, ul, uu, ddf);
        for (int q = 0; q < L.Q; ++q) fv[q] = feq[q];
    }

    /* collide, src/lbm/kernels/steps/collide.mojo:23-90 (SRT collision.mojo:12-46, KBC KBC_.mojo:180-238). */
    void collide(T* fv, T tau) const {
        bool ddf = d.ddf_shift != 0;
        T rho = cell.density(fv, ddf);
        T u[3];
        cell.velocity(fv, rho, u);
        T tau_local = tau;
        bool need_neq = d.les != LBM_LES_NONE || d.collision == LBM_COLL_KBC || d.collision == LBM_COLL_RLBM;
        T fneq[27], pi[6];
        if (need_neq) {
            T uu = cell.dot_uu(u);
            for (int q = 0; q < L.Q; ++q) fneq[q] = fv[q] - cell.feq(q, rho, u, uu, ddf);
            cell.pi_neq(fneq, pi);
            if (d.les == LBM_LES_SMAGORINSKY) tau_local += cell.smagorinsky_tau(pi, rho, tau, Cs);
            if (d.les == LBM_LES_MLP) {
                const T tau_s = cell.smagorinsky_tau(pi, rho, tau, Cs);
                float x[9];
                for (int n = 0; n < 6; ++n) x[n] = n < cell.nstress() ? (float)pi[n] : 0.0f;
                x[6] = (float)rho;
                x[7] = (float)std::sqrt(uu);
                x[8] = (float)(tau_s / (T)3);
                double nu = 0;
                tau_local += (g_mlp && mlp_nu(g_mlp, x, &nu)) ? (T)(3.0 * nu) : tau_s;
            }
        }
        if (d.collision == LBM_COLL_RLBM) {
            /* RLBM, src/lbm/kernels/ops/collisions/collision.mojo:168-223: the non-equilibrium part is
             * replaced by its second-order Hermite projection w_q / (2 cs^4) H_q : Pi_neq,
             * H_q,ab = c_a c_b - cs^2 delta_ab, off-diagonal pairs counted twice (get_kbc_Qiab :138-166). */
            T inv_tau = (T)1 / tau_local;
            T keep = (T)1 - inv_tau;
            T uu = cell.dot_uu(u);
            const T cs2 = (T)(1.0 / 3.0);
            for (int q = 0; q < L.Q; ++q) {
                T dot = 0;
                bool first = true;
                for (int n = 0; n < cell.nstress(); ++n) {
                    int a, b;
                    cell.stress_pair(n, a, b);
                    int cc = L.c[q][a] * L.c[q][b];
                    T h;
                    if (a == b) h = (T)cc - cs2;
                    else if (cc == 0) continue;
                    else h = (T)cc * (T)2;
                    T t = h * pi[n];
                    dot = first ? t : dot + t;
                    first = false;
                }
                T k = cell.w[q] / (T)(2.0 / 9.0);
                fv[q] = cell.feq(q, rho, u, uu, ddf) + keep * (k * dot);
            }
            return;
        }
        if (d.collision == LBM_COLL_KBC) {
            T beta = (T)1 / ((T)2 * tau_local);
            T inv_beta = (T)2 * tau_local;
            T uu = cell.dot_uu(u);
            T feq[27], ds[27], dh[27];
            for (int q = 0; q < L.Q; ++q) feq[q] = cell.feq(q, rho, u, uu, false);
            cell.kbc_shear(pi, ds);
            for (int q = 0; q < L.Q; ++q) dh[q] = fneq[q] - ds[q];
            T sp1 = 0, sp2 = 0;
            for (int q = 0; q < L.Q; ++q) sp1 += ds[q] * dh[q] / feq[q];
            for (int q = 0; q < L.Q; ++q) sp2 += dh[q] * dh[q] / feq[q];
            T gamma = inv_beta - ((T)2 - inv_beta) * sp1 / (sp2 + (T)1e-32);
            for (int q = 0; q < L.Q; ++q) fv[q] -= beta * (ds[q] * (T)2 + gamma * dh[q]);
            return;
        }
        if (d.collision == LBM_COLL_TRT) {
            /* TRT, src/lbm/kernels/ops/collisions/collision.mojo:49-113; tau_asymm = 1/2 + (3/16)/(tau - 1/2)
             * (src/lbm/kernels/steps/collide.mojo:83-85). Opposite directions are the pairs (q, q+1), q odd. */
            T uu = cell.dot_uu(u);
            T tau_a = (T)0.5 + (T)0.1875 / (tau_local - (T)0.5);
            T inv_ts = (T)1 / tau_local, inv_ta = (T)1 / tau_a;
            fv[0] = fv[0] - inv_ts * (fv[0] - cell.feq(0, rho, u, uu, ddf));
            for (int q = 1; q < L.Q; q += 2) {
                int o = q + 1;
                T fs = (fv[q] + fv[o]) * (T)0.5, fa = (fv[q] - fv[o]) * (T)0.5;
                T eq = cell.feq(q, rho, u, uu, ddf), eo = cell.feq(o, rho, u, uu, ddf);
                T es = (eq + eo) * (T)0.5, ea = (eq - eo) * (T)0.5;
                fv[q] = fv[q] - (inv_ts * (fs - es) + inv_ta * (fa - ea));
                fv[o] = fv[o] - (inv_ts * (fs - es) + inv_ta * ((-fa) - (-ea)));
            }
            return;
        }
        // SRT
        T uu = cell.dot_uu(u);
        T inv_tau = (T)1 / tau_local;
        for (int q = 0; q < L.Q; ++q) fv[q] = std::fma(-inv_tau, fv[q] - cell.feq(q, rho, u, uu, ddf), fv[q]);
    }

    bool valid(uint8_t flag) const { return !(flag_is_solid(flag) || flag_excluded(flag)); }  // is_valid_thread
    bool eq_enabled() const { return (d.included_bcs >> LBM_FLAG_EQUILIBRIUM) & 1u; }
    uint32_t capture_bits() const {
        return d.included_bcs & (LBM_FLAG_CAPTURE_DENSITY | LBM_FLAG_CAPTURE_VELOCITY);
    }

    /* capture_density / capture_velocity, src/lbm/kernels/steps/collide.mojo:54-63: the density and
     * velocity the collision starts from (get_density / get_velocity after streaming and the
     * boundary conditions) stored into the node's own bc entry (value D / values 0..D-1), for the
     * flag bits the run turned on. The reference addresses the entry through get_rank4_coord, whose
     * index is broken (src/lbm/kernels/utils/index.mojo:120-121); the node's own entry is the one it
     * names. */
    void capture(const T* fv, uint8_t flag, T* bc, const int idx[3]) const {
        const int D = L.D;
        const T rho = cell.density(fv, d.ddf_shift != 0);
        if ((capture_bits() & LBM_FLAG_CAPTURE_DENSITY) && (flag & LBM_FLAG_CAPTURE_DENSITY))
            bc[g.addr(idx[0], idx[1], idx[2], D, D + 1)] = rho;
        if ((capture_bits() & LBM_FLAG_CAPTURE_VELOCITY) && (flag & LBM_FLAG_CAPTURE_VELOCITY)) {
            T u[3] = {0, 0, 0};
            cell.velocity(fv, rho, u);
            for (int dd = 0; dd < D; ++dd) bc[g.addr(idx[0], idx[1], idx[2], dd, D + 1)] = u[dd];
        }
    }

    /* stream + apply_boundary_conditions for one node; returns false for invalid nodes. */
    void gather(const S* f, const T* bc, const uint8_t* flags, const int idx[3], int mode, T* fv,
                uint8_t pf[27], double* force = nullptr) const {
        // mode: 0 double buffer, 1 esoteric even, 2 esoteric odd
        uint32_t rebuilt = 0;
        if (mode == 0) {
            gather_flags(flags, idx, pf);
            load_db(f, pf, idx, fv);
            double fdummy[3] = {0, 0, 0};
            rebuilt = bouzidi(fv, pf, f, flags, idx, force ? force : fdummy);
        } else {
            load_ep(f, idx, mode == 1, fv);
            pf[0] = flags[g.addr(idx[0], idx[1], idx[2], 0, 1)];
            if (d.include_moving_boundary) gather_flags(flags, idx, pf);
        }
        if (d.include_moving_boundary) moving_wall(fv, pf, bc, idx, rebuilt);
        if (eq_enabled()) equilibrium(fv, pf, bc, idx);
    }
};

int g_threads = 1;
float g_q_clamp = 0.0f;  // lower bound of every link fraction (q_clamp of the reference's Bouzidi kernels)

template <class S, class T>
int step_impl(const lbm_desc* d, S* f_out, const S* f_in, const T* bc, const uint8_t* flags, double tau,
              int mode, const uint32_t* link_nodes = nullptr, const float* link_q = nullptr, uint32_t n_links = 0,
              double* force_out = nullptr, double* node_force_out = nullptr) {
    Stepper<S, T> st(*d);
    st.link_nodes = link_nodes;
    st.link_q = link_q;
    st.n_links = n_links;
    st.q_clamp = g_q_clamp;
    const Geo& g = st.g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    const T tau_t = (T)tau;
    std::vector<double> nforce(force_out ? (size_t)ncell * 3 : 0, 0.0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < ncell; ++i) {
        int idx[3] = {(int)(i % g.nx), (int)((i / g.nx) % g.ny), (int)(i / ((long)g.nx * g.ny))};
        uint8_t flag = flags[g.addr(idx[0], idx[1], idx[2], 0, 1)];
        if (!st.valid(flag)) continue;
        T fv[27];
        uint8_t pf[27];
        st.gather(f_in, bc, flags, idx, mode, fv, pf, force_out ? &nforce[(size_t)i * 3] : nullptr);
        if (st.capture_bits()) st.capture(fv, flag, const_cast<T*>(bc), idx);  // bc is the caller's buffer
        st.collide(fv, tau_t);
        if (mode == 0) {
            for (int q = 0; q < st.L.Q; ++q) Store<S, T>::store(f_out, g.addr(idx[0], idx[1], idx[2], q, st.K()), fv[q]);
        } else {
            st.store_ep(f_out, idx, mode == 1, fv);
        }
    }
    if (force_out) {
        for (int k = 0; k < 3; ++k) force_out[k] = 0.0;
        for (long i = 0; i < ncell; ++i)
            for (int k = 0; k < 3; ++k) force_out[k] += nforce[(size_t)i * 3 + k];
    }
    if (force_out && node_force_out) {
        // the force of each link record's node (drag.mojo:104-105), D values per record
        for (long i = 0; i < ncell; ++i) {
            int idx[3] = {(int)(i % g.nx), (int)((i / g.nx) % g.ny), (int)(i / ((long)g.nx * g.ny))};
            const uint32_t node = (uint32_t)g.addr(idx[0], idx[1], idx[2], 0, 1);
            const uint32_t* r = std::lower_bound(link_nodes, link_nodes + n_links, node);
            if (r == link_nodes + n_links || *r != node) continue;
            const uint8_t flag = flags[node];
            if (!st.valid(flag) || (flag & 0x0F) != LBM_FLAG_FLUID_SOLID_BOUNDARY) continue;
            for (int k = 0; k < st.L.D; ++k) node_force_out[(size_t)(r - link_nodes) * st.L.D + k] = nforce[(size_t)i * 3 + k];
        }
    }
    return 0;
}

/* calculate_rho_and_velocity, src/lbm/output/velocity.mojo:25-125 (tau = 0, solid nodes report bc). */
template <class S, class T>
int macro_impl(const lbm_desc* d, T* rho_out, T* u_out, const S* f, const T* bc, const uint8_t* flags, int mode,
               int u_layout, const uint32_t* link_nodes = nullptr, const float* link_q = nullptr,
               uint32_t n_links = 0) {
    Stepper<S, T> st(*d);
    st.link_nodes = link_nodes;
    st.link_q = link_q;
    st.n_links = n_links;
    st.q_clamp = g_q_clamp;
    const Geo& g = st.g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    const int D = st.L.D;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < ncell; ++i) {
        int idx[3] = {(int)(i % g.nx), (int)((i / g.nx) % g.ny), (int)(i / ((long)g.nx * g.ny))};
        uint8_t flag = flags[g.addr(idx[0], idx[1], idx[2], 0, 1)];
        T rho, u[3] = {0, 0, 0};
        if (flag != LBM_FLAG_SOLID) {
            T fv[27];
            uint8_t pf[27];
            st.gather(f, bc, flags, idx, mode, fv, pf);
            rho = st.cell.density(fv, d->ddf_shift != 0);
            st.cell.velocity(fv, rho, u);
        } else {
            for (int dd = 0; dd < D; ++dd) u[dd] = bc[g.addr(idx[0], idx[1], idx[2], dd, D + 1)];
            rho = bc[g.addr(idx[0], idx[1], idx[2], D, D + 1)];
        }
        size_t r = ((size_t)idx[0] * g.ny + idx[1]) * g.nz + idx[2];
        rho_out[r] = rho;
        for (int dd = 0; dd < D; ++dd) {
            if (u_layout == LBM_U_XYZD) u_out[r * D + dd] = u[dd];
            else u_out[(size_t)dd * ncell + r] = u[dd];
        }
    }
    return 0;
}

/* FP64 sums of rho and rho*u over valid nodes, in node order (x fastest). */
template <class S, class T>
int reduce_impl(const lbm_desc* d, double* out, const S* f, const T* bc, const uint8_t* flags, int mode,
                const uint32_t* link_nodes = nullptr, const float* link_q = nullptr, uint32_t n_links = 0) {
    Stepper<S, T> st(*d);
    st.link_nodes = link_nodes;
    st.link_q = link_q;
    st.n_links = n_links;
    st.q_clamp = g_q_clamp;
    const Geo& g = st.g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    const int D = st.L.D;
    std::vector<double> part((size_t)ncell * (D + 1), 0.0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < ncell; ++i) {
        int idx[3] = {(int)(i % g.nx), (int)((i / g.nx) % g.ny), (int)(i / ((long)g.nx * g.ny))};
        uint8_t flag = flags[g.addr(idx[0], idx[1], idx[2], 0, 1)];
        if (!st.valid(flag)) continue;
        T fv[27];
        uint8_t pf[27];
        st.gather(f, bc, flags, idx, mode, fv, pf);
        T rho = st.cell.density(fv, d->ddf_shift != 0);
        T u[3];
        st.cell.velocity(fv, rho, u);
        part[(size_t)i * (D + 1)] = (double)rho;
        for (int dd = 0; dd < D; ++dd) part[(size_t)i * (D + 1) + 1 + dd] = (double)rho * (double)u[dd];
    }
    for (int k = 0; k <= D; ++k) out[k] = 0.0;
    for (long i = 0; i < ncell; ++i)
        for (int k = 0; k <= D; ++k) out[k] += part[(size_t)i * (D + 1) + k];
    return 0;
}

/* get_velocity_gradient, get_adj_finite_difference and forward_difference,
 * src/lbm/kernels/utils/finite_difference.mojo:15-136, with dx = 1. A neighbour outside
 * the grid (no periodic wrap, index.mojo is_index_valid) is skipped; a SOLID neighbour
 * (raw flag compare) sits at half distance. The two one-sided differences are blended by
 * left_weight = 1 - left_dx / (left_dx + right_dx), right_weight = 1 - left_weight, so
 * with a single valid side the valid difference gets weight 0 and the result is 0.
 * u is the row-major velocity field (x, y, z, d) or (d, x, y, z). */
template <class T>
T velocity_gradient(const Geo& g, int D, const T* u, int u_layout, const uint8_t* flags, const int idx[3], int comp,
                    int axis) {
    const int n[3] = {g.nx, g.ny, g.nz};
    const size_t ncell = (size_t)g.nx * g.ny * g.nz;
    auto U = [&](const int p[3]) {
        const size_t r = ((size_t)p[0] * g.ny + p[1]) * g.nz + p[2];
        return u_layout == LBM_U_XYZD ? u[r * D + comp] : u[(size_t)comp * ncell + r];
    };
    const T uc = U(idx);
    T left_grad = 0, left_dx = 0, right_grad = 0, right_dx = 0;
    int a[3] = {idx[0], idx[1], idx[2]};
    a[axis] = idx[axis] - 1;
    const bool left_ok = a[axis] >= 0;
    if (left_ok) {
        const T ua = U(a);
        const T gdx = flags[g.addr(a[0], a[1], a[2], 0, 1)] == LBM_FLAG_SOLID ? (T)0.5 : (T)1;
        left_grad = (uc - ua) * ((T)1 / gdx);
        left_dx = gdx;
    }
    a[axis] = idx[axis] + 1;
    const bool right_ok = a[axis] < n[axis];
    if (right_ok) {
        const T ua = U(a);
        const T gdx = flags[g.addr(a[0], a[1], a[2], 0, 1)] == LBM_FLAG_SOLID ? (T)0.5 : (T)1;
        right_grad = (ua - uc) * ((T)1 / gdx);
        right_dx = gdx;
    }
    const T total = left_dx + right_dx;
    const T left_w = (T)1 - left_dx / total;
    const T right_w = (T)1 - left_w;
    return (left_ok || right_ok) ? right_grad * right_w + left_grad * left_w : (T)0;
}

/* calculate_Q_criterion and calculate_vorticity, src/lbm/output/Q_criterion.mojo:25-215:
 * |omega|^2 from finite differences of the velocity input (Vector.norm_squared, a left-to-right
 * sum from 0), the strain rate S = Pi_neq * (-1 / (2 rho cs^2 tau)) of the re-streamed,
 * boundary-applied populations (get_strain_rate_tensor, moment.mojo:154-176; its squared
 * norm with doubled off-diagonals, :182-215), Q = 0.25 |omega|^2 - 0.5 |S|^2. SOLID nodes
 * (raw flag) get 0; the reference leaves them unwritten. Output row-major (x, y, z). */
template <class S, class T>
int q_impl(const lbm_desc* d, T* q_out, const T* u_in, int u_layout, const S* f, const T* bc, const uint8_t* flags,
           double tau_d, int mode, const uint32_t* link_nodes = nullptr, const float* link_q = nullptr,
           uint32_t n_links = 0) {
    Stepper<S, T> st(*d);
    st.link_nodes = link_nodes;
    st.link_q = link_q;
    st.n_links = n_links;
    st.q_clamp = g_q_clamp;
    const Geo& g = st.g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    const int D = st.L.D;
    const bool ddf = d->ddf_shift != 0;
    const T tau = (T)tau_d;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < ncell; ++i) {
        int idx[3] = {(int)(i % g.nx), (int)((i / g.nx) % g.ny), (int)(i / ((long)g.nx * g.ny))};
        const size_t r = ((size_t)idx[0] * g.ny + idx[1]) * g.nz + idx[2];
        if (flags[g.addr(idx[0], idx[1], idx[2], 0, 1)] == LBM_FLAG_SOLID) {
            q_out[r] = 0;
            continue;
        }
        auto grad = [&](int comp, int axis) { return velocity_gradient<T>(g, D, u_in, u_layout, flags, idx, comp, axis); };
        T w[3];
        int nw;
        if (D == 2) {
            w[0] = grad(1, 0) - grad(0, 1);
            nw = 1;
        } else {
            w[0] = grad(2, 1) - grad(1, 2);
            w[1] = grad(0, 2) - grad(2, 0);
            w[2] = grad(1, 0) - grad(0, 1);
            nw = 3;
        }
        T wsq = 0;
        for (int k = 0; k < nw; ++k) wsq += w[k] * w[k];
        T fv[27];
        uint8_t pf[27];
        st.gather(f, bc, flags, idx, mode, fv, pf);
        const T rho = st.cell.density(fv, ddf);
        T u[3] = {0, 0, 0};
        st.cell.velocity(fv, rho, u);
        const T uu = st.cell.dot_uu(u);
        T fn[27], pi[6];
        for (int q = 0; q < st.L.Q; ++q) fn[q] = fv[q] - st.cell.feq(q, rho, u, uu, ddf);
        st.cell.pi_neq(fn, pi);
        const T cs2 = (T)1 / (T)3;
        const T k = (T)(-1) / ((T)2 * rho * cs2 * tau);
        T snorm = 0;
        for (int n = 0; n < st.cell.nstress(); ++n) {
            int a, b;
            st.cell.stress_pair(n, a, b);
            const T s = pi[n] * k;
            const T ss = s * s;
            snorm += a != b ? ss * (T)2 : ss;
        }
        q_out[r] = (T)0.25 * wsq - (T)0.5 * snorm;
    }
    return 0;
}

/* initialize_f_from_func with a uniform velocity and no gradient term,
 * src/lbm/preprocess/initial_condition.mojo:86-178 (esoteric: odd store pattern, :172-174). */
template <class S, class T>
int init_impl(const lbm_desc* d, S* f, double rho_d, const double* u_d) {
    Stepper<S, T> st(*d);
    const Geo& g = st.g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    T rho = (T)rho_d, u[3] = {0, 0, 0};
    if (u_d)
        for (int dd = 0; dd < st.L.D; ++dd) u[dd] = (T)u_d[dd];
    T uu = st.cell.dot_uu(u);
    T fv[27];
    for (int q = 0; q < st.L.Q; ++q) fv[q] = st.cell.feq(q, rho, u, uu, d->ddf_shift != 0);
    for (long i = 0; i < ncell; ++i) {
        int idx[3] = {(int)(i % g.nx), (int)((i / g.nx) % g.ny), (int)(i / ((long)g.nx * g.ny))};
        if (d->method == LBM_ESOTERIC_PULL) st.store_ep(f, idx, false, fv);
        else
            for (int q = 0; q < st.L.Q; ++q) Store<S, T>::store(f, g.addr(idx[0], idx[1], idx[2], q, st.K()), fv[q]);
    }
    return 0;
}

/* initialize_f_from_func, src/lbm/preprocess/initial_condition.mojo:86-178, for given fields, with
 * fi_neq (:181-234) and calculate_Sab (:236-255) when grad is given:
 *   fi_neq = 0; for a, b: fi_neq += (c_a c_b - cs^2 [a == b]) * 0.5 (g_ab + g_ba);
 *   fi_neq *= ((-w rho) tau / cs^2) * ((tau - 1) / tau); f_q = f_eq + fi_neq,
 * stored in the double-buffer layout or with the esoteric odd store pattern (:172-174).
 * u: row-major (x, y, z, d); grad: row-major (x, y, z, a, b) = d u_a / d x_b, or NULL. */
template <class S, class T>
int init_func_impl(const lbm_desc* d, S* f, double rho_d, const T* u, const T* grad, double tau_d) {
    Stepper<S, T> st(*d);
    const Geo& g = st.g;
    const int D = st.L.D, Q = st.L.Q;
    const T rho = (T)rho_d, tau = (T)tau_d;
    const T cs2 = (T)1 / (T)3;
    const bool ddf = d->ddf_shift != 0;
    for (int x = 0; x < g.nx; ++x)
        for (int y = 0; y < g.ny; ++y)
            for (int z = 0; z < g.nz; ++z) {
                const size_t r = ((size_t)x * g.ny + y) * g.nz + z;
                T uv[3] = {0, 0, 0};
                for (int dd = 0; dd < D; ++dd) uv[dd] = u[r * D + dd];
                const T uu = st.cell.dot_uu(uv);
                T fv[27];
                for (int q = 0; q < Q; ++q) {
                    T fi = st.cell.feq(q, rho, uv, uu, ddf);
                    if (grad) {
                        const T* gr = grad + r * D * D;
                        T neq = 0;
                        for (int a = 0; a < D; ++a)
                            for (int b = 0; b < D; ++b) {
                                const T sab = (T)0.5 * (gr[a * D + b] + gr[b * D + a]);
                                T qiab = (T)st.L.c[q][a] * (T)st.L.c[q][b];
                                if (a == b) qiab -= cs2;
                                neq += qiab * sab;
                            }
                        neq *= (((-st.cell.w[q]) * rho) * tau * (T)3) * ((tau - (T)1) / tau);
                        fi += neq;
                    }
                    fv[q] = fi;
                }
                const int idx[3] = {x, y, z};
                if (d->method == LBM_ESOTERIC_PULL) st.store_ep(f, idx, false, fv);
                else
                    for (int q = 0; q < Q; ++q) Store<S, T>::store(f, g.addr(x, y, z, q, st.K()), fv[q]);
            }
    return 0;
}

int check(const lbm_desc* d) {
    if (!d) return -1;
    bool lat_ok = (d->D == 2 && d->Q == 9) || (d->D == 3 && (d->Q == 19 || d->Q == 27));
    if (!lat_ok) return -1;
    if (d->tx <= 0 || d->ty <= 0 || d->tz <= 0) return -1;
    if (d->nx % d->tx || d->ny % d->ty || d->nz % d->tz) return -1;
    if (d->D == 2 && (d->nz != 1 || d->tz != 1)) return -1;
    return 0;
}

}  // namespace

extern "C" {

/* Parameter block of the learned closure (1446 floats, kept by the caller) for les == 2. */
void orc_set_mlp(const float* params) { g_mlp = params; }

/* The closure as the oracle's collide evaluates it: nu[s] for raw features X[s] (n x 9), -1 outside. */
void orc_mlp_closure(const float* params, const float* X, long n, double* nu) {
    for (long s = 0; s < n; ++s) {
        double v = 0;
        nu[s] = mlp_nu(params, X + 9 * s, &v) ? v : -1.0;
    }
}

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
void orc_set_q_clamp(float q) { g_q_clamp = q; }
int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

uint16_t orc_f16c_encode(float x) { return f16c_encode(x); }
float orc_f16c_decode(uint16_t h) { return f16c_decode(h); }
void orc_f16c_encode_n(const float* x, uint16_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = f16c_encode(x[i]);
}
void orc_f16c_decode_n(const uint16_t* h, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = f16c_decode(h[i]);
}

/* Lattice tables for tests: c (Q x 3 ints), weights (double, num/den), opposite indices. */
int orc_lattice(int D, int Q, int* c_out, double* w_out, int* opp_out) {
    lbm_desc d{};
    d.D = D; d.Q = Q;
    bool ok = (D == 2 && Q == 9) || (D == 3 && (Q == 19 || Q == 27));
    if (!ok) return -1;
    Lat L = make_lat(D, Q);
    for (int q = 0; q < Q; ++q) {
        for (int k = 0; k < 3; ++k) c_out[3 * q + k] = L.c[q][k];
        w_out[q] = (double)L.wnum[q] / (double)L.wden[q];
        opp_out[q] = L.opp[q];
    }
    return 0;
}

size_t orc_addr(const lbm_desc* d, int x, int y, int z, int k, int K) { return make_geo(d).addr(x, y, z, k, K); }

#define ORC_DISPATCH(CALL_F32, CALL_F16C, CALL_F64)                   \
    if (check(d)) return -1;                                          \
    switch (d->storage) {                                             \
        case LBM_STORE_F32: return CALL_F32;                          \
        case LBM_STORE_F16C: return CALL_F16C;                        \
        case LBM_STORE_F64: return CALL_F64;                          \
        default: return -1;                                           \
    }

int orc_step_db(const lbm_desc* d, void* f_out, const void* f_in, const void* bc, const uint8_t* flags, double tau) {
    ORC_DISPATCH((step_impl<float, float>(d, (float*)f_out, (const float*)f_in, (const float*)bc, flags, tau, 0)),
                 (step_impl<uint16_t, float>(d, (uint16_t*)f_out, (const uint16_t*)f_in, (const float*)bc, flags, tau, 0)),
                 (step_impl<double, double>(d, (double*)f_out, (const double*)f_in, (const double*)bc, flags, tau, 0)))
}

/* Double-buffer step with interpolated bounce-back links; force_out[0..2] = total momentum-exchange force,
 * node_force_out (NULL or n_links * D doubles) = the force of each link record's node. */
int orc_step_db_links(const lbm_desc* d, void* f_out, const void* f_in, const void* bc, const uint8_t* flags,
                      double tau, const uint32_t* link_nodes, const float* link_q, uint32_t n_links,
                      double* force_out, double* node_force_out) {
    ORC_DISPATCH((step_impl<float, float>(d, (float*)f_out, (const float*)f_in, (const float*)bc, flags, tau, 0,
                                          link_nodes, link_q, n_links, force_out, node_force_out)),
                 (step_impl<uint16_t, float>(d, (uint16_t*)f_out, (const uint16_t*)f_in, (const float*)bc, flags, tau,
                                             0, link_nodes, link_q, n_links, force_out, node_force_out)),
                 (step_impl<double, double>(d, (double*)f_out, (const double*)f_in, (const double*)bc, flags, tau, 0,
                                            link_nodes, link_q, n_links, force_out, node_force_out)))
}

int orc_step_ep(const lbm_desc* d, void* f, const void* bc, const uint8_t* flags, double tau, int is_even) {
    int mode = is_even ? 1 : 2;
    ORC_DISPATCH((step_impl<float, float>(d, (float*)f, (const float*)f, (const float*)bc, flags, tau, mode)),
                 (step_impl<uint16_t, float>(d, (uint16_t*)f, (const uint16_t*)f, (const float*)bc, flags, tau, mode)),
                 (step_impl<double, double>(d, (double*)f, (const double*)f, (const double*)bc, flags, tau, mode)))
}

int orc_macroscopic(const lbm_desc* d, void* rho, void* u, const void* f, const void* bc, const uint8_t* flags,
                    int after_odd, int u_layout) {
    int mode = d->method == LBM_ESOTERIC_PULL ? (after_odd ? 1 : 2) : 0;
    ORC_DISPATCH((macro_impl<float, float>(d, (float*)rho, (float*)u, (const float*)f, (const float*)bc, flags, mode, u_layout)),
                 (macro_impl<uint16_t, float>(d, (float*)rho, (float*)u, (const uint16_t*)f, (const float*)bc, flags, mode, u_layout)),
                 (macro_impl<double, double>(d, (double*)rho, (double*)u, (const double*)f, (const double*)bc, flags, mode, u_layout)))
}

int orc_reduce_moments(const lbm_desc* d, double* out, const void* f, const void* bc, const uint8_t* flags,
                       int after_odd) {
    int mode = d->method == LBM_ESOTERIC_PULL ? (after_odd ? 1 : 2) : 0;
    ORC_DISPATCH((reduce_impl<float, float>(d, out, (const float*)f, (const float*)bc, flags, mode)),
                 (reduce_impl<uint16_t, float>(d, out, (const uint16_t*)f, (const float*)bc, flags, mode)),
                 (reduce_impl<double, double>(d, out, (const double*)f, (const double*)bc, flags, mode)))
}

/* Double-buffer macroscopic fields / FP64 totals on a grid with interpolated bounce-back links. */
int orc_macroscopic_links(const lbm_desc* d, void* rho, void* u, const void* f, const void* bc, const uint8_t* flags,
                          int u_layout, const uint32_t* nodes, const float* q, uint32_t n) {
    if (d->method != LBM_DOUBLE_BUFFER) return -1;
    ORC_DISPATCH((macro_impl<float, float>(d, (float*)rho, (float*)u, (const float*)f, (const float*)bc, flags, 0, u_layout, nodes, q, n)),
                 (macro_impl<uint16_t, float>(d, (float*)rho, (float*)u, (const uint16_t*)f, (const float*)bc, flags, 0, u_layout, nodes, q, n)),
                 (macro_impl<double, double>(d, (double*)rho, (double*)u, (const double*)f, (const double*)bc, flags, 0, u_layout, nodes, q, n)))
}

int orc_reduce_moments_links(const lbm_desc* d, double* out, const void* f, const void* bc, const uint8_t* flags,
                             const uint32_t* nodes, const float* q, uint32_t n) {
    if (d->method != LBM_DOUBLE_BUFFER) return -1;
    ORC_DISPATCH((reduce_impl<float, float>(d, out, (const float*)f, (const float*)bc, flags, 0, nodes, q, n)),
                 (reduce_impl<uint16_t, float>(d, out, (const uint16_t*)f, (const float*)bc, flags, 0, nodes, q, n)),
                 (reduce_impl<double, double>(d, out, (const double*)f, (const double*)bc, flags, 0, nodes, q, n)))
}

/* Q-criterion of the current state (after_odd as for orc_macroscopic); u is the velocity field
 * of orc_macroscopic in u_layout, q the row-major output, both in the grid float type. */
int orc_q_criterion(const lbm_desc* d, void* q, const void* u, int u_layout, const void* f, const void* bc,
                    const uint8_t* flags, double tau, int after_odd) {
    int mode = d->method == LBM_ESOTERIC_PULL ? (after_odd ? 1 : 2) : 0;
    ORC_DISPATCH((q_impl<float, float>(d, (float*)q, (const float*)u, u_layout, (const float*)f, (const float*)bc, flags, tau, mode)),
                 (q_impl<uint16_t, float>(d, (float*)q, (const float*)u, u_layout, (const uint16_t*)f, (const float*)bc, flags, tau, mode)),
                 (q_impl<double, double>(d, (double*)q, (const double*)u, u_layout, (const double*)f, (const double*)bc, flags, tau, mode)))
}

int orc_q_criterion_links(const lbm_desc* d, void* q, const void* u, int u_layout, const void* f, const void* bc,
                          const uint8_t* flags, double tau, const uint32_t* nodes, const float* qd, uint32_t n) {
    ORC_DISPATCH((q_impl<float, float>(d, (float*)q, (const float*)u, u_layout, (const float*)f, (const float*)bc, flags, tau, 0, nodes, qd, n)),
                 (q_impl<uint16_t, float>(d, (float*)q, (const float*)u, u_layout, (const uint16_t*)f, (const float*)bc, flags, tau, 0, nodes, qd, n)),
                 (q_impl<double, double>(d, (double*)q, (const double*)u, u_layout, (const double*)f, (const double*)bc, flags, tau, 0, nodes, qd, n)))
}

/* One finite-difference velocity gradient d u_comp / d x_axis at node (x, y, z), FP64. */
double orc_velocity_gradient(const lbm_desc* d, const double* u, int u_layout, const uint8_t* flags, int x, int y,
                             int z, int comp, int axis) {
    const int idx[3] = {x, y, z};
    return velocity_gradient<double>(make_geo(d), d->D, u, u_layout, flags, idx, comp, axis);
}

/* f from a row-major velocity field (and gradient field, or NULL) in the grid float type. */
int orc_init_f_from_fields(const lbm_desc* d, void* f, double rho, const void* u, const void* grad, double tau) {
    ORC_DISPATCH((init_func_impl<float, float>(d, (float*)f, rho, (const float*)u, (const float*)grad, tau)),
                 (init_func_impl<uint16_t, float>(d, (uint16_t*)f, rho, (const float*)u, (const float*)grad, tau)),
                 (init_func_impl<double, double>(d, (double*)f, rho, (const double*)u, (const double*)grad, tau)))
}

int orc_init_equilibrium(const lbm_desc* d, void* f, double rho, const double* u) {
    ORC_DISPATCH((init_impl<float, float>(d, (float*)f, rho, u)),
                 (init_impl<uint16_t, float>(d, (uint16_t*)f, rho, u)),
                 (init_impl<double, double>(d, (double*)f, rho, u)))
}

/* set_exterior_walls host loop, src/lbm/preprocess/boundary_condition.mojo:78-183 (no validation here;
 * the validation restatement lives in the tests). bc in the grid float type. */
int orc_set_exterior_wall(const lbm_desc* d, uint8_t* flags, void* bc, int axis, int plus, int flag,
                          const double* u, double rho) {
    if (check(d)) return -1;
    Geo g = make_geo(d);
    int D = d->D;
    int n[3] = {g.nx, g.ny, g.nz};
    int fixed = plus ? n[axis] - 1 : 0;
    int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    for (int i = 0; i < n[a1]; ++i)
        for (int j = 0; j < n[a2]; ++j) {
            int idx[3];
            idx[axis] = fixed; idx[a1] = i; idx[a2] = j;
            flags[g.addr(idx[0], idx[1], idx[2], 0, 1)] = (uint8_t)flag;
            for (int k = 0; k <= D; ++k) {
                double v = (k < D) ? (u ? u[k] : NAN) : rho;
                size_t a = g.addr(idx[0], idx[1], idx[2], k, D + 1);
                if (d->storage == LBM_STORE_F64) ((double*)bc)[a] = v;
                else ((float*)bc)[a] = (float)v;
            }
        }
    return 0;
}

}  // extern "C"

// lbm_les_data.hip -- training-data kernels of the learned LES closure (include/lbm_hip.h,
// "Learned closure"): box filter of fine fields onto a coarse grid, a-priori features and
// eddy-viscosity targets on the coarse grid, and per-node equilibrium initialisation from
// fields. No reference counterpart for the first two; the initialisations are
// initialize_f_from_func (src/lbm/preprocess/initial_condition.mojo:86-234) for given fields.
#include <cmath>
#include <string>

#include "lbm_launch.hpp"

namespace lbm {

constexpr int LES_CH = 10;  // rho, u (3), u_i u_j (xx, xy, xz, yy, yz, zz)

// out[((I cy + J) cz + K) * 10 + c] = mean over fine nodes factor*I + offset + a (a < factor, per
// axis) of (rho, u, u u); rho row-major (x, y, z), u row-major (x, y, z, d).
template <class R>
__global__ void __launch_bounds__(256) box_filter_kernel(const R* __restrict__ rho, const R* __restrict__ u, int nx,
                                                         int ny, int nz, int factor, int offset, int cx, int cy,
                                                         int cz, float* __restrict__ out) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= (int64_t)cx * cy * cz) return;
    const int K = (int)(c % cz), J = (int)((c / cz) % cy), I = (int)(c / ((int64_t)cz * cy));
    double acc[LES_CH] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int cnt = 0;
    for (int a = 0; a < factor; ++a)
        for (int b = 0; b < factor; ++b)
            for (int e = 0; e < factor; ++e) {
                const int x = factor * I + offset + a, y = factor * J + offset + b, z = factor * K + offset + e;
                if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
                const size_t r = ((size_t)x * ny + y) * nz + z;
                const double ux = u[r * 3], uy = u[r * 3 + 1], uz = u[r * 3 + 2];
                acc[0] += rho[r];
                acc[1] += ux;
                acc[2] += uy;
                acc[3] += uz;
                acc[4] += ux * ux;
                acc[5] += ux * uy;
                acc[6] += ux * uz;
                acc[7] += uy * uy;
                acc[8] += uy * uz;
                acc[9] += uz * uz;
                ++cnt;
            }
    const double inv = cnt ? 1.0 / cnt : 0.0;
#pragma unroll
    for (int k = 0; k < LES_CH; ++k) out[c * LES_CH + k] = (float)(acc[k] * inv);
}

// Strain rate of the filtered velocity at coarse cell (I,J,K) by central differences:
// s = (xx, xy, xz, yy, yz, zz).
__device__ __forceinline__ void coarse_strain(const float* __restrict__ F, int cy, int cz, int I, int J, int K,
                                              double (&s)[6]) {
    auto U = [&](int i, int j, int k, int d) -> double {
        return F[(((size_t)i * cy + j) * cz + k) * LES_CH + 1 + d];
    };
    double g[3][3];  // g[d][a] = d u_d / d x_a
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g[d][0] = 0.5 * (U(I + 1, J, K, d) - U(I - 1, J, K, d));
        g[d][1] = 0.5 * (U(I, J + 1, K, d) - U(I, J - 1, K, d));
        g[d][2] = 0.5 * (U(I, J, K + 1, d) - U(I, J, K - 1, d));
    }
    s[0] = g[0][0];
    s[1] = 0.5 * (g[0][1] + g[1][0]);
    s[2] = 0.5 * (g[0][2] + g[2][0]);
    s[3] = g[1][1];
    s[4] = 0.5 * (g[1][2] + g[2][1]);
    s[5] = g[2][2];
}

// For each listed coarse cell (2 <= I,J,K <= n-3): the eddy viscosity that best maps the
// filtered strain onto the subgrid stress over the 3x3x3 neighbourhood (least squares of
// tau_d = -2 nu S, clipped at 0),
//   nu = max(0, -<tau_d : S> / (2 <S : S>)),  tau_ij = <u_i u_j> - <u_i><u_j>,
// and the features collide sees when the coarse LBM runs with that closure:
// Pi_neq = -2 rho cs^2 (tau0 + 3 nu) S, rho, |u|, and the Smagorinsky nu_t of that Pi_neq
// (get_Smagorinsky_LES_tau with the base tau, src/lbm/kernels/ops/turbulence.mojo:12-38).
__global__ void __launch_bounds__(256) les_samples_kernel(const float* __restrict__ F, int cx, int cy, int cz,
                                                          double tau0, double cs, const uint32_t* __restrict__ cells,
                                                          int64_t n, float* __restrict__ X, float* __restrict__ Y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cells[i];
    const int K = (int)(c % (uint32_t)cz), J = (int)((c / (uint32_t)cz) % (uint32_t)cy),
              I = (int)(c / ((uint32_t)cz * (uint32_t)cy));
    if (I < 2 || J < 2 || K < 2 || I > cx - 3 || J > cy - 3 || K > cz - 3) {
        for (int k = 0; k < 9; ++k) X[i * 9 + k] = 0.0f;
        Y[i] = -1.0f;  // marks a cell outside the sampling region
        return;
    }
    double num = 0.0, den = 0.0;
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
            for (int e = -1; e <= 1; ++e) {
                double s[6];
                coarse_strain(F, cy, cz, I + a, J + b, K + e, s);
                const float* p = F + (((size_t)(I + a) * cy + (J + b)) * cz + (K + e)) * LES_CH;
                const double ux = p[1], uy = p[2], uz = p[3];
                double t[6] = {p[4] - ux * ux, p[5] - ux * uy, p[6] - ux * uz,
                               p[7] - uy * uy, p[8] - uy * uz, p[9] - uz * uz};
                const double tr3 = (t[0] + t[3] + t[5]) / 3.0;
                t[0] -= tr3;
                t[3] -= tr3;
                t[5] -= tr3;
                num += t[0] * s[0] + t[3] * s[3] + t[5] * s[5] + 2.0 * (t[1] * s[1] + t[2] * s[2] + t[4] * s[4]);
                den += s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + 2.0 * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4]);
            }
    const double nu = den > 1e-30 ? fmax(0.0, -num / (2.0 * den)) : 0.0;
    double s[6];
    coarse_strain(F, cy, cz, I, J, K, s);
    const float* p = F + (((size_t)I * cy + J) * cz + K) * LES_CH;
    const double rho = p[0], ux = p[1], uy = p[2], uz = p[3];
    const double cs2 = 1.0 / 3.0;
    const double kpi = -2.0 * rho * cs2 * (tau0 + 3.0 * nu);
    double pi[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pi[k] = kpi * s[k];
    // Smagorinsky nu_t of this Pi_neq with the base tau
    const double kb = -1.0 / (2.0 * rho * cs2 * tau0);
    double sn = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double sb = pi[k] * kb;
        sn += (k == 1 || k == 2 || k == 4) ? 2.0 * sb * sb : sb * sb;
    }
    const double nu_s = cs * cs * sqrt(2.0 * sn);
#pragma unroll
    for (int k = 0; k < 6; ++k) X[i * 9 + k] = (float)pi[k];
    X[i * 9 + 6] = (float)rho;
    X[i * 9 + 7] = (float)sqrt(ux * ux + uy * uy + uz * uz);
    X[i * 9 + 8] = (float)nu_s;
    Y[i] = (float)nu;
}

// f = f_eq(rho(x), u(x)) at every node (double-buffer layout); rho row-major (x, y, z),
// u row-major (x, y, z, d), both in the grid float type.
template <class L, class S, bool DDF>
__global__ void __launch_bounds__(256) init_fields_kernel(typename S::T* __restrict__ f,
                                                          const typename S::R* __restrict__ rho,
                                                          const typename S::R* __restrict__ u, Geo g) {
    using R = typename S::R;
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= (int64_t)g.nx * g.ny * g.nz) return;
    const int z = (int)(r % g.nz), y = (int)((r / g.nz) % g.ny), x = (int)(r / ((int64_t)g.nz * g.ny));
    R uu3[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < L::D; ++d) uu3[d] = u[r * L::D + d];
    const R rr = rho[r];
    const R uu = dot_uu<L, R>(uu3);
    const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
    const size_t loc = (size_t)((x & (tx - 1)) + ((y & (ty - 1)) << g.txb) + ((z & (tz - 1)) << (g.txb + g.tyb)));
    const size_t tl = (size_t)(x >> g.txb) + (size_t)g.Tx * ((size_t)(y >> g.tyb) + (size_t)g.Ty * (size_t)(z >> g.tzb));
#pragma unroll
    for (int q = 0; q < L::Q; ++q) S::st(f + (tl * L::Q + q) * (size_t)g.tile_cells + loc, feq<L, R, DDF>(q, rr, uu3, uu));
}

// Element index in f of node (x, y, z), population q.
__device__ __forceinline__ size_t f_elem(const Geo& g, int x, int y, int z, int Q, int q) {
    const int tx = 1 << g.txb, ty = 1 << g.tyb, tz = 1 << g.tzb;
    const size_t loc = (size_t)((x & (tx - 1)) + ((y & (ty - 1)) << g.txb) + ((z & (tz - 1)) << (g.txb + g.tyb)));
    const size_t tl = (size_t)(x >> g.txb) + (size_t)g.Tx * ((size_t)(y >> g.tyb) + (size_t)g.Ty * (size_t)(z >> g.tzb));
    return (tl * Q + q) * (size_t)g.tile_cells + loc;
}

// initialize_f_from_func (src/lbm/preprocess/initial_condition.mojo:86-178) over given fields:
// f_q = f_eq(rho, u) plus, when grad is given, fi_neq (:181-234):
//   fi_neq = sum_ab (c_a c_b - cs^2 delta_ab) 0.5 (g_ab + g_ba), then
//   fi_neq *= ((-w_q rho) tau / cs^2) ((tau - 1) / tau),
// with g_ab = d u_a / d x_b. Stored in the double-buffer layout or, for esoteric pull, with the
// odd store pattern (:172-174): f_p (p odd) to slot p of x + c_p, the rest to x. u row-major
// (x, y, z, d) and grad row-major (x, y, z, a, b) cover the x-slab [x0, x0 + nxc).
template <class L, class S, bool DDF, bool EP, bool NEQ>
__global__ void __launch_bounds__(256) init_func_kernel(typename S::T* __restrict__ f, typename S::R rho,
                                                        const typename S::R* __restrict__ u,
                                                        const typename S::R* __restrict__ grad, typename S::R tau,
                                                        int x0, int nxc, Geo g) {
    using R = typename S::R;
    constexpr int D = L::D, Q = L::Q;
    const int64_t nyz = (int64_t)g.ny * g.nz;
    const int64_t rl = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (rl >= (int64_t)nxc * nyz) return;
    const int x = x0 + (int)(rl / nyz), y = (int)((rl / g.nz) % g.ny), z = (int)(rl % g.nz);
    R uv[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d) uv[d] = u[rl * D + d];
    const R uu = dot_uu<L, R>(uv);
    R fv[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) fv[q] = feq<L, R, DDF>(q, rho, uv, uu);
    if (NEQ) {
        R gr[3][3];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) gr[a][b] = grad[(rl * D + a) * D + b];
        const R cs2 = (R)1 / (R)3;
        const R tail = (tau - (R)1) / tau;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            R acc = 0;
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    const R sab = (R)0.5 * (gr[a][b] + gr[b][a]);
                    R qiab = (R)L::c(q, a) * (R)L::c(q, b);
                    if (a == b) qiab = qiab - cs2;
                    acc = acc + qiab * sab;
                }
            const R pre = (((-weight<L, R>(q)) * rho) * tau) * (R)3;
            acc = acc * (pre * tail);
            fv[q] = fv[q] + acc;
        }
    }
    if (!EP) {
#pragma unroll
        for (int q = 0; q < Q; ++q) S::st(f + f_elem(g, x, y, z, Q, q), fv[q]);
        return;
    }
    S::st(f + f_elem(g, x, y, z, Q, 0), fv[0]);
#pragma unroll
    for (int p = 1; p < Q; p += 2) {
        int xn = x + L::c(p, 0), yn = y + L::c(p, 1), zn = z + L::c(p, 2);
        xn += xn < 0 ? g.nx : 0;
        xn -= xn >= g.nx ? g.nx : 0;
        yn += yn < 0 ? g.ny : 0;
        yn -= yn >= g.ny ? g.ny : 0;
        zn += zn < 0 ? g.nz : 0;
        zn -= zn >= g.nz ? g.nz : 0;
        S::st(f + f_elem(g, xn, yn, zn, Q, p), fv[p]);
        S::st(f + f_elem(g, x, y, z, Q, p + 1), fv[p + 1]);
    }
}

template <class L, class S, bool DDF, bool EP>
static void init_func_launch(const Handle& h, void* f, double rho, const void* u, const void* grad, double tau,
                             int x0, int nxc, unsigned blocks, hipStream_t st) {
    using R = typename S::R;
    if (grad)
        hipLaunchKernelGGL((init_func_kernel<L, S, DDF, EP, true>), dim3(blocks), dim3(256), 0, st,
                           (typename S::T*)f, (R)rho, (const R*)u, (const R*)grad, (R)tau, x0, nxc, h.g);
    else
        hipLaunchKernelGGL((init_func_kernel<L, S, DDF, EP, false>), dim3(blocks), dim3(256), 0, st,
                           (typename S::T*)f, (R)rho, (const R*)u, (const R*)nullptr, (R)tau, x0, nxc, h.g);
}

template <class L, class S>
int init_func(const Handle& h, void* f, double rho, const void* u, const void* grad, double tau, int x0, int nxc,
              hipStream_t st) {
    const int64_t n = (int64_t)nxc * h.g.ny * h.g.nz;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    const bool ddf = h.d.ddf_shift != 0, ep = h.d.method == LBM_ESOTERIC_PULL;
    if (ddf && ep) init_func_launch<L, S, true, true>(h, f, rho, u, grad, tau, x0, nxc, blocks, st);
    else if (ddf) init_func_launch<L, S, true, false>(h, f, rho, u, grad, tau, x0, nxc, blocks, st);
    else if (ep) init_func_launch<L, S, false, true>(h, f, rho, u, grad, tau, x0, nxc, blocks, st);
    else init_func_launch<L, S, false, false>(h, f, rho, u, grad, tau, x0, nxc, blocks, st);
    return check_launch("init_func_kernel");
}

template <class L, class S>
int init_fields(const Handle& h, void* f, const void* rho, const void* u, hipStream_t st) {
    using R = typename S::R;
    const int64_t n = (int64_t)h.g.nx * h.g.ny * h.g.nz;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (h.d.ddf_shift)
        hipLaunchKernelGGL((init_fields_kernel<L, S, true>), dim3(blocks), dim3(256), 0, st, (typename S::T*)f,
                           (const R*)rho, (const R*)u, h.g);
    else
        hipLaunchKernelGGL((init_fields_kernel<L, S, false>), dim3(blocks), dim3(256), 0, st, (typename S::T*)f,
                           (const R*)rho, (const R*)u, h.g);
    return check_launch("init_fields_kernel");
}

}  // namespace lbm

using namespace lbm;

extern "C" {

int lbm_box_filter(const void* rho, const void* u, int storage, int32_t nx, int32_t ny, int32_t nz, int32_t factor,
                   int32_t offset, int32_t cx, int32_t cy, int32_t cz, float* out, void* stream) {
    if (!rho || !u || !out || factor < 1 || nx <= 0 || ny <= 0 || nz <= 0 || cx <= 0 || cy <= 0 || cz <= 0)
        return set_error(LBM_ERR_INVALID, "bad box filter arguments");
    const int64_t n = (int64_t)cx * cy * cz;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (storage == LBM_STORE_F64)
        hipLaunchKernelGGL((box_filter_kernel<double>), dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (const double*)rho, (const double*)u, nx, ny, nz, factor, offset, cx, cy, cz, out);
    else
        hipLaunchKernelGGL((box_filter_kernel<float>), dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (const float*)rho, (const float*)u, nx, ny, nz, factor, offset, cx, cy, cz, out);
    return check_launch("box_filter_kernel");
}

int lbm_init_fields(lbm_handle* hh, void* f, const void* rho, const void* u, void* stream) {
    if (!hh || !f || !rho || !u) return set_error(LBM_ERR_INVALID, "NULL argument");
    const Handle& h = *reinterpret_cast<Handle*>(hh);
    if (h.d.method != LBM_DOUBLE_BUFFER)
        return set_error(LBM_ERR_UNSUPPORTED, "field initialisation is provided for the double-buffer layout");
    const hipStream_t st = (hipStream_t)stream;
#define LBM_CASE(QQ, SS, SID) \
    if (h.d.Q == QQ && h.d.storage == SID) return init_fields<Lattice<QQ>, SS>(h, f, rho, u, st);
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

int lbm_init_f_from_fields(lbm_handle* hh, void* f, double rho, const void* u, const void* grad, double tau,
                           int32_t x0, int32_t nx_slab, void* stream) {
    if (!hh || !f || !u) return set_error(LBM_ERR_INVALID, "NULL argument");
    const Handle& h = *reinterpret_cast<Handle*>(hh);
    if (x0 < 0 || nx_slab <= 0 || (int64_t)x0 + nx_slab > h.g.nx)
        return set_error(LBM_ERR_INVALID, "the x-slab [x0, x0 + nx_slab) must lie inside the grid");
    if (grad && !(tau != 0.0)) return set_error(LBM_ERR_INVALID, "the non-equilibrium term needs tau != 0");
    const hipStream_t st = (hipStream_t)stream;
#define LBM_CASE(QQ, SS, SID) \
    if (h.d.Q == QQ && h.d.storage == SID) return init_func<Lattice<QQ>, SS>(h, f, rho, u, grad, tau, x0, nx_slab, st);
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

int lbm_les_samples(const float* coarse, int32_t cx, int32_t cy, int32_t cz, double tau0, double cs,
                    const uint32_t* cells, int64_t n, float* X, float* y, void* stream) {
    if (!coarse || !cells || !X || !y || n < 0 || cx < 5 || cy < 5 || cz < 5)
        return set_error(LBM_ERR_INVALID, "bad sample arguments");
    if (!n) return LBM_OK;
    hipLaunchKernelGGL(les_samples_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, coarse,
                       cx, cy, cz, tau0, cs, cells, n, X, y);
    return check_launch("les_samples_kernel");
}

}  // extern "C"

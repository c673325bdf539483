// Exercises every oracle entry point on small grids, for the AddressSanitizer /
// UndefinedBehaviorSanitizer build in tests/test_oracle_sanitize.py. Exits non-zero on a
// failed call or a non-finite result.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/lbm_hip.h"

extern "C" {
int orc_step_db(const lbm_desc*, void*, const void*, const void*, const uint8_t*, double);
int orc_step_db_links(const lbm_desc*, void*, const void*, const void*, const uint8_t*, double, const uint32_t*,
                      const float*, uint32_t, double*, double*);
int orc_step_ep(const lbm_desc*, void*, const void*, const uint8_t*, double, int);
int orc_macroscopic(const lbm_desc*, void*, void*, const void*, const void*, const uint8_t*, int, int);
int orc_reduce_moments(const lbm_desc*, double*, const void*, const void*, const uint8_t*, int);
int orc_q_criterion(const lbm_desc*, void*, const void*, int, const void*, const void*, const uint8_t*, double, int);
int orc_init_f_from_fields(const lbm_desc*, void*, double, const void*, const void*, double);
int orc_init_equilibrium(const lbm_desc*, void*, double, const double*);
int orc_set_exterior_wall(const lbm_desc*, uint8_t*, void*, int, int, int, const double*, double);
size_t orc_addr(const lbm_desc*, int, int, int, int, int);
}

static int fail(const char* what) {
    std::fprintf(stderr, "failed: %s\n", what);
    return 1;
}

template <class R>
static int run(lbm_desc d) {
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    const size_t es = d.storage == LBM_STORE_F16C ? 2 : (d.storage == LBM_STORE_F64 ? 8 : 4);
    std::vector<uint8_t> f(n * d.Q * es), f2(n * d.Q * es), flags(n, 0);
    std::vector<R> bc(n * (d.D + 1), 0), rho(n), u(n * d.D), q(n), grad(n * d.D * d.D);
    const double lid[3] = {0.05, 0.0, 0.0}, zero[3] = {0.0, 0.0, 0.0};
    if (orc_init_equilibrium(&d, f.data(), 1.0, nullptr)) return fail("init");
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < d.D; ++a) {
            u[i * d.D + a] = (R)(0.01 * std::sin(0.1 * (double)(i + a)));
            for (int b = 0; b < d.D; ++b) grad[(i * d.D + a) * d.D + b] = (R)(1e-3 * std::cos((double)(i + a + b)));
        }
    if (orc_init_f_from_fields(&d, f.data(), 1.0, u.data(), grad.data(), 0.7)) return fail("init_f_from_fields");
    f2 = f;
    if (orc_set_exterior_wall(&d, flags.data(), bc.data(), 1, 1, LBM_FLAG_SOLID, lid, 1.0)) return fail("wall");
    for (int axis = 0; axis < d.D; ++axis)
        for (int plus = 0; plus < 2; ++plus)
            if (!(axis == 1 && plus == 1) && orc_set_exterior_wall(&d, flags.data(), bc.data(), axis, plus, LBM_FLAG_SOLID, zero, 1.0))
                return fail("walls");
    for (int s = 0; s < 3; ++s) {
        if (d.method == LBM_DOUBLE_BUFFER) {
            if (orc_step_db(&d, f2.data(), f.data(), bc.data(), flags.data(), 0.6)) return fail("step_db");
            f.swap(f2);
        } else if (orc_step_ep(&d, f.data(), bc.data(), flags.data(), 0.6, s % 2 == 0)) {
            return fail("step_ep");
        }
    }
    if (d.method == LBM_DOUBLE_BUFFER) {
        // one link on every fluid node next to the -X wall, half of them on each branch
        std::vector<uint32_t> nodes;
        std::vector<float> qd;
        for (int y = 1; y < d.ny - 1; ++y)
            for (int z = (d.D == 3 ? 1 : 0); z < (d.D == 3 ? d.nz - 1 : 1); ++z) nodes.push_back((uint32_t)orc_addr(&d, 1, y, z, 0, 1));
        std::sort(nodes.begin(), nodes.end());
        for (size_t i = 0; i < nodes.size(); ++i)
            for (int k = 1; k < d.Q; ++k) qd.push_back(i % 2 ? 0.3f : 0.8f);
        for (uint32_t nd : nodes) flags[nd] = LBM_FLAG_FLUID_SOLID_BOUNDARY;
        double force[3] = {0, 0, 0};
        std::vector<double> node_force(nodes.size() * (size_t)d.D);
        if (orc_step_db_links(&d, f2.data(), f.data(), bc.data(), flags.data(), 0.6, nodes.data(), qd.data(),
                              (uint32_t)nodes.size(), force, node_force.data()))
            return fail("step_db_links");
        f.swap(f2);
    }
    if (orc_macroscopic(&d, rho.data(), u.data(), f.data(), bc.data(), flags.data(), 1, LBM_U_XYZD)) return fail("macro");
    double tot[4];
    if (orc_reduce_moments(&d, tot, f.data(), bc.data(), flags.data(), 1)) return fail("reduce");
    if (orc_q_criterion(&d, q.data(), u.data(), LBM_U_XYZD, f.data(), bc.data(), flags.data(), 0.6, 1)) return fail("q");
    for (int k = 0; k <= d.D; ++k)
        if (!std::isfinite(tot[k])) return fail("finite totals");
    return 0;
}

int main() {
    int bad = 0;
    const int lats[3][2] = {{2, 9}, {3, 19}, {3, 27}};
    for (auto& l : lats)
        for (int storage = 0; storage < 3; ++storage)
            for (int method = 0; method < 2; ++method) {
                lbm_desc d{};
                d.D = l[0];
                d.Q = l[1];
                d.nx = 16; d.ny = 16; d.nz = l[0] == 3 ? 8 : 1;
                d.tx = 8; d.ty = 8; d.tz = l[0] == 3 ? 4 : 1;
                d.method = method;
                d.collision = LBM_COLL_SRT;
                d.storage = storage;
                d.ddf_shift = storage == LBM_STORE_F16C;
                d.les = l[0] == 3 ? LBM_LES_SMAGORINSKY : LBM_LES_NONE;
                d.include_moving_boundary = 1;
                d.included_bcs = 3;
                d.smagorinsky_cs = 0.1;
                bad += storage == LBM_STORE_F64 ? run<double>(d) : run<float>(d);
            }
    std::printf("%s\n", bad ? "sanitized oracle run FAILED" : "sanitized oracle run ok");
    return bad ? 1 : 0;
}

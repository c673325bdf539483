/*
 * A plain C host of liblbm_hip.so: no Python, no torch, only include/lbm_hip.h and the
 * library's own device-memory helpers, the way a Mojo host without a GPU runtime would
 * drive it. D3Q19 lid-driven cube, both methods: create a handle, allocate the buffers,
 * fill f_eq at rest, write the walls (the lid last but one, as the reference scripts do),
 * step, and print the FP64 mass / momentum totals and one velocity sample as JSON.
 * Exit status: 0 on success, 1 on any library error.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/lbm_hip.h"

#define CHECK(call)                                                                    \
    do {                                                                               \
        int rc_ = (call);                                                              \
        if (rc_ != LBM_OK) {                                                           \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, lbm_last_error());    \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

static int run(int method, int n, int steps) {
    lbm_desc d = {0};
    d.D = 3; d.Q = 19; d.nx = n; d.ny = n; d.nz = n; d.tx = 8; d.ty = 8; d.tz = 8;
    d.method = method; d.collision = LBM_COLL_SRT; d.storage = LBM_STORE_F32;
    d.include_moving_boundary = 1; d.included_bcs = 3; d.smagorinsky_cs = 0.1;
    lbm_handle* h = NULL;
    CHECK(lbm_create(&d, &h));
    const size_t fb = lbm_f_bytes(&d), bb = lbm_bc_bytes(&d), gb = lbm_flags_bytes(&d);
    void *f = NULL, *f2 = NULL, *bc = NULL, *flags = NULL, *rho = NULL, *u = NULL, *tot = NULL;
    CHECK(lbm_alloc(fb, &f));
    CHECK(lbm_alloc(fb, &f2));
    CHECK(lbm_alloc(bb, &bc));
    CHECK(lbm_alloc(gb, &flags));
    CHECK(lbm_alloc((size_t)n * n * n * 4, &rho));
    CHECK(lbm_alloc((size_t)n * n * n * 3 * 4, &u));
    CHECK(lbm_alloc(4 * sizeof(double), &tot));
    CHECK(lbm_memset(bc, 0, bb, NULL));
    CHECK(lbm_memset(flags, 0, gb, NULL));
    CHECK(lbm_init_equilibrium(h, f, 1.0, NULL, NULL));
    CHECK(lbm_init_equilibrium(h, f2, 1.0, NULL, NULL));
    const double lid[3] = {0.1, 0.0, 0.0}, zero[3] = {0.0, 0.0, 0.0};
    CHECK(lbm_set_exterior_wall(h, (uint8_t*)flags, bc, "+Y", LBM_FLAG_SOLID, lid, 1.0, NULL));
    const char* sides[5] = {"-Y", "+X", "-X", "-Z", "+Z"};
    for (int i = 0; i < 5; ++i) CHECK(lbm_set_exterior_wall(h, (uint8_t*)flags, bc, sides[i], LBM_FLAG_SOLID, zero, 1.0, NULL));
    const double tau = 0.5 + 3.0 * 0.1 * n / 100.0;
    double m0[4];
    CHECK(lbm_reduce_moments_f64(h, (double*)tot, f, bc, (const uint8_t*)flags, 1, NULL));
    CHECK(lbm_memcpy_d2h(m0, tot, sizeof(m0), NULL));
    for (int s = 0; s < steps; ++s) {
        if (method == LBM_DOUBLE_BUFFER) {
            CHECK(lbm_step_db(h, f2, f, bc, (const uint8_t*)flags, tau, NULL));
            void* t = f; f = f2; f2 = t;
        } else {
            CHECK(lbm_step_ep(h, f, bc, (const uint8_t*)flags, tau, s % 2 == 0, NULL));
        }
    }
    const int after_odd = steps > 0 && (steps - 1) % 2 == 1;  /* the last step was an odd one */
    double m1[4];
    CHECK(lbm_reduce_moments_f64(h, (double*)tot, f, bc, (const uint8_t*)flags, after_odd, NULL));
    CHECK(lbm_memcpy_d2h(m1, tot, sizeof(m1), NULL));
    CHECK(lbm_macroscopic(h, rho, u, f, bc, (const uint8_t*)flags, after_odd, LBM_U_XYZD, NULL));
    float* uh = (float*)malloc((size_t)n * n * n * 3 * 4);
    CHECK(lbm_memcpy_d2h(uh, u, (size_t)n * n * n * 3 * 4, NULL));
    CHECK(lbm_synchronize(NULL));
    const size_t probe = (((size_t)(n / 2) * n + (n - 2)) * n + n / 2) * 3;  /* below the lid, mid-plane */
    printf("{\"method\": %d, \"n\": %d, \"steps\": %d, \"mass0\": %.17g, \"mass\": %.17g, \"momentum\": [%.6g, %.6g, %.6g], "
           "\"u_below_lid\": %.6g}\n", method, n, steps, m0[0], m1[0], m1[1], m1[2], m1[3], (double)uh[probe]);
    free(uh);
    CHECK(lbm_free(f)); CHECK(lbm_free(f2)); CHECK(lbm_free(bc)); CHECK(lbm_free(flags));
    CHECK(lbm_free(rho)); CHECK(lbm_free(u)); CHECK(lbm_free(tot));
    CHECK(lbm_destroy(h));
    return 0;
}

int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 32;
    const int steps = argc > 2 ? atoi(argv[2]) : 200;
    int count = 0;
    CHECK(lbm_device_count(&count));
    if (count < 1) {
        fprintf(stderr, "no HIP device visible\n");
        return 1;
    }
    if (run(LBM_DOUBLE_BUFFER, n, steps) || run(LBM_ESOTERIC_PULL, n, steps)) return 1;
    return 0;
}

/*
 * lbm_hip.h -- C ABI of the MI355X (gfx950) lattice Boltzmann stream-collide library.
 *
 * The library replaces the GPU side of mojo_lbm's solver step: the fused
 * pull / esoteric-pull stream -> bounce-back / moving-wall / equilibrium BC ->
 * SRT / KBC (+ Smagorinsky LES) collide -> store kernel, the rho/u output
 * kernel and the host setup loops that fill f, flags and bc.
 *
 * Conventions
 *  - Every buffer is caller-owned device memory, laid out exactly like the
 *    reference's TileTensors: rank-4 (x,y,z,k) fields use the blocked product of
 *    a column-major tile (tx,ty,tz,K) and a column-major tiler (Tx,Ty,Tz,1)
 *    (src/lbm/layoutstruct.mojo:45-64):
 *        addr(x,y,z,k) = lx + tx*(ly + ty*(lz + tz*k)) + tx*ty*tz*K*(bx + Tx*(by + Ty*bz))
 *    with K = Q for f, K = D+1 for bc (u_0..u_{D-1}, rho) and K = 1 for flags.
 *  - f is float (LBM_STORE_F32), uint16 Float16C codes (LBM_STORE_F16C) or
 *    double (LBM_STORE_F64). bc, rho, u and tau are in the grid float type:
 *    float for F32/F16C, double for F64.
 *  - Calls are asynchronous on the given hipStream_t (passed as void*; NULL is
 *    the default stream). Nothing is retained past a call except state owned by
 *    the handle.
 *  - Every function returns LBM_OK (0) or a negative status; lbm_last_error()
 *    returns the calling thread's last message. No exception crosses the ABI.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Streaming methods: LBM_method, src/lbm/constants.mojo:20-31. */
enum { LBM_DOUBLE_BUFFER = 0, LBM_ESOTERIC_PULL = 1 };

/* Collision operators: Collisions, src/lbm/constants.mojo:41-55. */
enum { LBM_COLL_SRT = 0, LBM_COLL_TRT = 1, LBM_COLL_KBC = 2, LBM_COLL_RLBM = 3 };

/* Population storage: f32 (default), Float16C codes (src/utils/custom_fp.mojo:16-100),
 * or f64 (the FP64 reference runs). */
enum { LBM_STORE_F32 = 0, LBM_STORE_F16C = 1, LBM_STORE_F64 = 2 };

/* Turbulence closure applied in collide (src/lbm/kernels/steps/collide.mojo:66-73).
 * LBM_LES_MLP is the learned closure of lbm_set_mlp (3-D lattices; no reference counterpart). */
enum { LBM_LES_NONE = 0, LBM_LES_SMAGORINSKY = 1, LBM_LES_MLP = 2 };

/* Node flags: enum in the low 4 bits, bit masks above (src/lbm/constants.mojo:58-117). */
enum {
    LBM_FLAG_FLUID = 0,
    LBM_FLAG_SOLID = 1,
    LBM_FLAG_EQUILIBRIUM = 2,
    LBM_FLAG_FLUID_SOLID_BOUNDARY = 3,
    LBM_FLAG_CAPTURE_VELOCITY = 0x20,
    LBM_FLAG_CAPTURE_DENSITY = 0x40,
    LBM_FLAG_EXCLUDE = 0x80
};

/* Layout of the velocity output of lbm_macroscopic (both row-major). */
enum { LBM_U_XYZD = 0, LBM_U_DXYZ = 1 };

/* Status codes. */
enum {
    LBM_OK = 0,
    LBM_ERR_INVALID = -1,     /* a validation error (bad desc, bad argument) */
    LBM_ERR_UNSUPPORTED = -2, /* a valid configuration this library does not provide */
    LBM_ERR_HIP = -3,         /* a HIP runtime error */
    LBM_ERR_NO_DEVICE = -4    /* no usable gfx950 device */
};

/*
 * Run description: the union of LBM_Grid's compile-time shape/tile parameters
 * (src/lbm/grid.mojo:55-106) and LBM_Config's toggles (src/lbm/config.mojo:50-190).
 */
typedef struct lbm_desc {
    int32_t D, Q;              /* (2,9) D2Q9, (3,19) D3Q19, (3,27) D3Q27 (src/lbm/lattice.mojo) */
    int32_t nx, ny, nz;        /* lattice nodes; nz = 1 for D = 2 */
    int32_t tx, ty, tz;        /* layout tile = launch block (grid.mojo:99, set_default_block_shape) */
    int32_t method;            /* LBM_DOUBLE_BUFFER | LBM_ESOTERIC_PULL */
    int32_t collision;         /* LBM_COLL_* */
    int32_t storage;           /* LBM_STORE_* (Float16C: config.mojo use_float16c) */
    int32_t ddf_shift;         /* store f - w_q (config.mojo DDF_shift) */
    int32_t les;               /* LBM_LES_* (config.mojo LES) */
    int32_t include_moving_boundary; /* config.mojo include_moving_boundary (reference default: 1) */
    uint32_t included_bcs;     /* bit (1u << flag enum) per BC type in the run; FLUID and SOLID are implied
                                  (config.mojo INCLUDED_BCs); LBM_FLAG_CAPTURE_DENSITY and
                                  LBM_FLAG_CAPTURE_VELOCITY (the flag values themselves, bits 6 and 5)
                                  turn on config.mojo capture_density / capture_velocity: a step then
                                  writes into bc at nodes whose flag carries the bit */
    double smagorinsky_cs;     /* Cs; the reference hard-codes 0.1 (steps/collide.mojo:71) */
} lbm_desc;

typedef struct lbm_handle lbm_handle;

/* Library version (major*10000 + minor*100 + patch). */
int lbm_version(void);

/* Last error message of the calling thread ("" when none). */
const char* lbm_last_error(void);

/* Number of visible HIP devices (0 when there is no GPU). */
int lbm_device_count(int* count);

/*
 * Validates a description the way the reference's comptime asserts and host
 * checks do (grid.mojo:419-440 divisibility, esoteric_pull.mojo:76-78 lattice
 * pairing, config.mojo:128-150 BC set, KBC lattice restriction) and builds a
 * handle bound to the current device. Replaces Solver.__init__
 * (src/lbm/solver.mojo:63-76).
 */
int lbm_create(const lbm_desc* desc, lbm_handle** out);
int lbm_destroy(lbm_handle* h);

/* Byte sizes of the f, bc and flags buffers for a description. */
size_t lbm_f_bytes(const lbm_desc* desc);
size_t lbm_bc_bytes(const lbm_desc* desc);
size_t lbm_flags_bytes(const lbm_desc* desc);

/*
 * One double-buffer step f_in -> f_out: replaces Solver.double_buffer_step
 * (src/lbm/solver.mojo:81-111) and double_buffer_kernel
 * (src/lbm/kernels/double_buffer.mojo:25-96). f_out and f_in must not alias.
 * With capture turned on in included_bcs, a valid node whose flag carries
 * LBM_FLAG_CAPTURE_DENSITY / LBM_FLAG_CAPTURE_VELOCITY writes the density /
 * velocity its collision starts from into its own bc entry (value D / values
 * 0..D-1; steps/collide.mojo:54-63), so bc must then be writable device memory.
 */
int lbm_step_db(lbm_handle* h, void* f_out, const void* f_in, const void* bc,
                const uint8_t* flags, double tau, void* stream);

/*
 * One in-place esoteric-pull step: replaces Solver.esoteric_pull_step[is_even]
 * (src/lbm/solver.mojo:115-154) and esoteric_pull_kernel
 * (src/lbm/kernels/esoteric_pull.mojo:24-111). Solver.even_step/odd_step pass
 * is_even = 1 / 0 (solver.mojo:159-187). Captures into bc as lbm_step_db does.
 */
int lbm_step_ep(lbm_handle* h, void* f, const void* bc, const uint8_t* flags,
                double tau, int is_even, void* stream);

/*
 * rho (nx*ny*nz, row-major (x,y,z)) and u (row-major (x,y,z,d) or (d,x,y,z))
 * after re-streaming and re-applying the BCs: replaces calculate_rho_and_velocity
 * (src/lbm/output/velocity.mojo:25-125). after_odd selects the esoteric load
 * pattern (ignored for double buffer). Results do not depend on alignment; 16-byte aligned rho
 * and u (with 8x8x8 tiles, a multiple of 4 tiles in z, FP32 or Float16C) take the faster
 * whole-line kernel.
 */
int lbm_macroscopic(lbm_handle* h, void* rho, void* u, const void* f, const void* bc,
                    const uint8_t* flags, int after_odd, int u_layout, void* stream);

/*
 * Q-criterion Q = 0.25 |omega|^2 - 0.5 |S|^2 per node (nx*ny*nz, row-major (x,y,z), grid float
 * type): replaces calculate_Q_criterion (src/lbm/output/Q_criterion.mojo:90-215) and
 * OutputRequest.get_Q_criterion / Q_criterion_frame (src/lbm/outputrequest.mojo:136-223).
 * omega comes from finite differences of u, the velocity field lbm_macroscopic wrote for the
 * same state (u_layout as there), with the reference's rule (finite_difference.mojo:15-136):
 * a SOLID neighbour sits at half distance, a neighbour outside the grid is skipped and a
 * derivative with one valid side is 0. S = Pi_neq * (-1 / (2 rho cs^2 tau)) from the
 * re-streamed, boundary-applied populations (after_odd as for lbm_macroscopic). SOLID nodes
 * (raw flag) get 0; the reference leaves them unwritten. A 16-byte aligned q takes the faster
 * whole-line kernel under the same conditions as lbm_macroscopic.
 */
int lbm_q_criterion(lbm_handle* h, void* q, const void* u, const void* f, const void* bc,
                    const uint8_t* flags, double tau, int after_odd, int u_layout, void* stream);

/*
 * FP64 totals over valid nodes (not SOLID, not EXCLUDE) of the same moments
 * lbm_macroscopic reports: out_dev[0] = sum rho, out_dev[1+d] = sum rho*u_d.
 * out_dev is device memory of D+1 doubles. Deterministic (fixed reduction
 * order). No reference counterpart: the conservation check used to state mass
 * drift.
 */
int lbm_reduce_moments_f64(lbm_handle* h, double* out_dev, const void* f, const void* bc,
                           const uint8_t* flags, int after_odd, void* stream);

/*
 * Fill f with the equilibrium of (rho, u) at every node, stored with the
 * method's pattern (double buffer: plain; esoteric pull: the odd store
 * pattern): replaces initialize_fluid_at_rest / initialize_f_from_func without
 * the gradient term (src/lbm/preprocess/initial_condition.mojo:47-178).
 * u points to D host doubles (NULL = at rest).
 */
int lbm_init_equilibrium(lbm_handle* h, void* f, double rho, const double* u, void* stream);

/*
 * Write one exterior face: flags = flag and bc = (u, rho) at every node of the
 * face, in the order the reference's host loop does (later calls overwrite
 * shared edges): replaces set_exterior_walls
 * (src/lbm/preprocess/boundary_condition.mojo:78-183), including _error_check
 * (:20-70). side is one of "-X", "+X", "-Y", "+Y", "-Z", "+Z". u points to D
 * host doubles or is NULL (free velocity, stored as NaN); rho NaN = free.
 */
int lbm_set_exterior_wall(lbm_handle* h, uint8_t* flags, void* bc, const char* side,
                          int flag, const double* u, double rho, void* stream);

/*
 * Interpolated (linear Bouzidi) bounce-back links for lbm_step_db: replaces the
 * link setup of RigidStationaryObject / get_rigid_sphere and the Bouzidi update of
 * nodewise_bounceback_kernel (src/lbm/geometry/rigidSphere.mojo:14-170,
 * src/lbm/geometry/nodewise_interpolated_BB.mojo:46-186), which the reference
 * leaves without a launcher. Links live on FLUID_SOLID_BOUNDARY nodes; record r
 * is the node's flag-layout index nodes[r] (strictly increasing) and, for
 * q = 1..Q-1, qdist[r*(Q-1) + q-1] = the fraction of the link from the node
 * towards its SOLID pull neighbour x - c_q that lies in the fluid (negative: plain
 * halfway bounce-back). Host arrays, copied into the handle (n_links = 0 clears).
 * lbm_macroscopic and lbm_reduce_moments_f64 apply the same links when re-streaming.
 */
int lbm_set_links(lbm_handle* h, const uint32_t* nodes, const float* qdist, uint32_t n_links);
/* lbm_set_links with every link fraction qd >= 0 raised to q_clamp (0 <= q_clamp <= 1): the
 * q_clamp argument of linkwise_bounceback_kernel / nodewise_bounceback_kernel
 * (src/lbm/geometry/interpolated_BB.mojo:60,117, nodewise_interpolated_BB.mojo:69,151).
 * lbm_set_links is this call with q_clamp = 0. */
int lbm_set_links_clamped(lbm_handle* h, const uint32_t* nodes, const float* qdist, uint32_t n_links,
                          float q_clamp);

/*
 * Momentum-exchange force on the solid from the interpolated links of the last
 * lbm_step_db: out_dev[0..D-1] doubles in device memory, summed in a fixed order.
 * Replaces calculate_drag_around_object / RigidStationaryObject.sum_force
 * (src/lbm/output/drag.mojo:16-105, src/lbm/geometry/rigidstationary.mojo:155-169).
 */
int lbm_link_force(lbm_handle* h, double* out_dev, void* stream);

/*
 * Per-node momentum-exchange force: while set, every lbm_step_db writes, for each link record r of
 * the last lbm_set_links, node_force_dev[r*D + d] = the force of that node's links this step (the
 * terms lbm_link_force sums), into caller-owned device memory of n_links * D doubles. Replaces the
 * force_tensor[tid, d] output of calculate_drag_around_object / nodewise_bounceback_kernel
 * (src/lbm/output/drag.mojo:104-105, src/lbm/geometry/nodewise_interpolated_BB.mojo:166-167); with
 * the DDF shift the stored populations are shifted back (+ 2 w_q per link) so the force is the
 * physical one. A record whose node is not a FLUID_SOLID_BOUNDARY node is not written. NULL stops
 * the output; lbm_set_links clears it (the record count changes).
 */
int lbm_set_link_node_force(lbm_handle* h, double* node_force_dev);

/*
 * Float16C codec on the device for n values: to_f16c = 1 converts floats to
 * uint16 codes, 0 converts codes to floats. The same codec the F16C storage
 * uses (Float16C.to_fp16c / to_fp32, src/utils/custom_fp.mojo:36-100), for
 * converting fields in and out of the population format.
 */
int lbm_f16c_convert(const void* in, void* out, size_t n, int to_f16c, void* stream);

/*
 * Parameters of the learned LES closure used by desc.les == LBM_LES_MLP: replaces the
 * Smagorinsky eddy viscosity of get_Smagorinsky_LES_tau at
 * src/lbm/kernels/steps/collide.mojo:70-73 with an MLP 9 -> 32 -> 32 -> 1 (SiLU, SiLU,
 * softplus) of (Pi_neq xx, xy, xz, yy, yz, zz; rho; |u|; Smagorinsky nu_t), and falls back
 * to Smagorinsky where a feature lies outside the training range. params: n = 1446 host
 * floats: W1 (9x32, k-major), b1[32], W2 (32x32, k-major), b2[32], w3[32], b3, mean[9],
 * inv_std[9], lo[9], hi[9], nu_scale. Copied into the handle. No reference counterpart.
 */
int lbm_set_mlp(lbm_handle* h, const float* params, int32_t n);

/*
 * Training kernels of the closure (device arrays; no reference counterpart). X: n x 9 raw
 * features, row-major; y: n target eddy viscosities; params: the 1446-float block above
 * (only the first 1409, the weights, are trained; normalisation and nu_scale are read).
 *   lbm_mlp_forward:   nu[s] = softplus(MLP(x_s)) * nu_scale (no range fallback).
 *   lbm_mlp_loss_grad: loss = mean_s (softplus(MLP(x_s)) - y_s / nu_scale)^2 into *loss_dev
 *                      (double, may be NULL) and d loss / d weights into grad_dev (1409
 *                      floats); deterministic; needs lbm_mlp_workspace_bytes(n) of scratch.
 *   lbm_mlp_adam:      one Adam step (bias-corrected, step t >= 1) on the 1409 weights with
 *                      moment buffers m, v (1409 floats each, zero before t = 1).
 */
int lbm_mlp_forward(const float* params_dev, const float* X_dev, int64_t n, float* nu_dev, void* stream);
/* The closure of handle h (lbm_set_mlp) exactly as collide evaluates it (bf16 matrix cores): nu[s]
 * for raw features X[s], or -1 where a feature lies outside the training range. */
int lbm_mlp_closure(lbm_handle* h, const float* X_dev, int64_t n, float* nu_dev, void* stream);
size_t lbm_mlp_workspace_bytes(int64_t n);
int lbm_mlp_loss_grad(const float* params_dev, const float* X_dev, const float* y_dev, int64_t n,
                      float* grad_dev, double* loss_dev, void* workspace_dev, void* stream);
int lbm_mlp_adam(float* params_dev, float* m_dev, float* v_dev, const float* grad_dev, int32_t t,
                 float lr, float beta1, float beta2, float eps, void* stream);
/*
 * lbm_mlp_loss_grad without a workspace argument and without any allocation: no hipMalloc, hipFree
 * or allocating library call, for callers that cannot allocate (code under HIP graph capture,
 * training loops in harnesses that forbid allocation). No counterpart in the reference. Its scratch
 * is a static device array of 5,779,456 bytes per GPU (the gradient partials of the largest launch,
 * 1024 workgroups, so every n > 0 fits), kept until the process ends; its address on each device is
 * looked up on the first call made with that device current. Same kernels as lbm_mlp_loss_grad, same
 * results bit for bit. Every call on one GPU shares that scratch: calls that overlap on one GPU (two
 * streams, two host threads) race and must be serialised by the caller.
 */
int lbm_mlp_loss_grad_noalloc(const float* params_dev, const float* X_dev, const float* y_dev, int64_t n,
                              float* grad_dev, double* loss_dev, void* stream);

/*
 * Training data of the closure (device arrays; no reference counterpart).
 *   lbm_box_filter:  coarse cell (I,J,K) of a cx x cy x cz grid averages the fine nodes
 *                    factor*I + offset + a (0 <= a < factor per axis, nodes outside the fine
 *                    grid skipped) of rho (row-major (x,y,z)) and u (row-major (x,y,z,d)) in the
 *                    grid float type of `storage`, writing 10 floats per cell, row-major
 *                    (I,J,K,c): rho, u_x, u_y, u_z, then <u_i u_j> for xx, xy, xz, yy, yz, zz.
 *   lbm_les_samples: for the listed coarse cells (row-major ids; cells within 2 of the
 *                    border get y = -1), the eddy viscosity that maps the filtered strain
 *                    onto the subgrid stress by least squares over the 3x3x3 neighbourhood
 *                    (clipped at 0) into y, and into X (n x 9) the features collide computes
 *                    when the coarse LBM runs with that eddy viscosity: Pi_neq =
 *                    -2 rho cs^2 (tau0 + 3 nu) S in stress-pair order, rho, |u| and the
 *                    Smagorinsky nu_t of that Pi_neq with Smagorinsky constant cs.
 *   lbm_init_fields: f = f_eq(rho(x), u(x)) per node (double buffer; rho, u laid out as
 *                    lbm_macroscopic writes them with LBM_U_XYZD; initialize_f_from_func
 *                    without the gradient term, src/lbm/preprocess/initial_condition.mojo:86-178).
 */
int lbm_box_filter(const void* rho, const void* u, int storage, int32_t nx, int32_t ny, int32_t nz,
                   int32_t factor, int32_t offset, int32_t cx, int32_t cy, int32_t cz, float* out,
                   void* stream);
int lbm_les_samples(const float* coarse, int32_t cx, int32_t cy, int32_t cz, double tau0, double cs,
                    const uint32_t* cells, int64_t n, float* X, float* y, void* stream);
int lbm_init_fields(lbm_handle* h, void* f, const void* rho, const void* u, void* stream);

/*
 * f from a velocity field on the device: replaces initialize_f_from_func
 * (src/lbm/preprocess/initial_condition.mojo:86-178) with its non-equilibrium term fi_neq
 * (:181-255). Per node f_q = f_eq(rho, u) + fi_neq, fi_neq = -w_q rho (tau - 1) / cs^2 *
 * sum_ab (c_a c_b - cs^2 delta_ab) S_ab with S_ab = (g_ab + g_ba) / 2, in the reference's
 * operation order; stored in the double-buffer layout or, for esoteric pull, the odd store
 * pattern the reference uses (:172-174). u (row-major (x,y,z,d)) and grad (row-major
 * (x,y,z,a,b) = d u_a / d x_b, or NULL for no fi_neq) are lattice-unit values in the grid
 * float type covering the x-slab [x0, x0 + nx_slab), so a large grid is filled slab by slab.
 */
int lbm_init_f_from_fields(lbm_handle* h, void* f, double rho, const void* u, const void* grad, double tau,
                           int32_t x0, int32_t nx_slab, void* stream);

/*
 * Launch tuning of one handle (no reference counterpart: the reference fixes
 * the launch shape at compile time, src/lbm/grid.mojo:99-100).
 *   LBM_OPT_XCD_REMAP: 1 = give each of the 8 XCDs a contiguous run of tiles
 *                      (default when the tile count is a multiple of 8),
 *                      0 = hardware round-robin order.
 *   LBM_OPT_STATIC_FLAGS: 1 = the caller promises that flags and bc do not change
 *                      while the same flags / bc buffers are passed: the per-tile
 *                      flag summary is rebuilt only when either pointer changes,
 *                      this option is set again or lbm_set_exterior_wall runs on the
 *                      handle, not before every step. A caller that edits flags or bc
 *                      in place by other means (a copy into the same buffers, a new
 *                      allocation that reuses a freed address) sets the option again
 *                      before the next step. 0 (default) = rebuilt before every step.
 *   LBM_OPT_ADDR64:    1 = use the 64-bit addressing kernels (which buffers of 2 GiB and
 *                      more always use) for any size, so they can be checked on small grids;
 *                      0 (default) = only for buffers of 2 GiB and more.
 *   LBM_OPT_STEP_LANES: lanes of one lbm_step_db / lbm_step_ep workgroup: a power of
 *                      two >= 64 that is a whole number of the tile's z layers, so
 *                      2^k workgroups share a tile, each with a z slab of it (the
 *                      (x,y,z,q) layout is unchanged); 0 or tile_cells = one
 *                      workgroup per tile. Default: 256 for Float16C storage or
 *                      the learned closure on 3-D tiles of more than 256 nodes,
 *                      else one workgroup per tile. Results are identical for
 *                      every value.
 *   LBM_OPT_STAGED_PULL: 1 (default) = Float16C double-buffer steps on tiles 8 x 8
 *                      nodes wide in 256-lane workgroups (32-bit addressing, no
 *                      interpolated links) copy each direction's source rows into
 *                      LDS first and pull from there; 2 = the same with one 512-lane
 *                      workgroup per pair of x-adjacent tiles (an even tile count
 *                      in x; else as 1); 0 = pull straight from memory. Results are
 *                      identical for every value.
 * Returns LBM_ERR_INVALID for an unknown option or value.
 */
enum { LBM_OPT_XCD_REMAP = 1, LBM_OPT_STATIC_FLAGS = 2, LBM_OPT_ADDR64 = 3, LBM_OPT_STEP_LANES = 4,
       LBM_OPT_STAGED_PULL = 5 };
int lbm_set_option(lbm_handle* h, int option, int value);

/* Device memory helpers for hosts without their own GPU runtime (the Mojo host
 * of ContextTileTensor, src/utils/contextTileTensor.mojo:113-158, 280-293). */
int lbm_alloc(size_t bytes, void** ptr);
int lbm_free(void* ptr);
int lbm_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes, void* stream);
int lbm_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes, void* stream);
int lbm_memset(void* dst_dev, int value, size_t bytes, void* stream);
int lbm_synchronize(void* stream);

/*
 * Batched cases. Many independent double-buffer grids of different sizes and
 * relaxation times -- e.g. a sweep of 2-D cylinder channels -- advance together,
 * one kernel launch per step for all of them, runs of steps replayed from a
 * captured hipGraph. Drives what the reference does one case per script with
 * Solver.double_buffer_step (src/lbm/solver.mojo:89-110) and reads the drag of
 * calculate_drag_around_object (src/lbm/output/drag.mojo:16-105) per case; the
 * reference has no batched launch. The batch owns its buffers. Each case is
 * padded to whole tiles with SOLID nodes (a case relying on periodic wrap needs
 * sizes that are tile multiples). Case arrays are host memory in row-major
 * (x, y, z) node order (x slowest, as lbm_macroscopic's rho): flags one byte
 * per node, bc D+1 doubles per node (u..., rho), link records as lbm_set_links
 * but with row-major node ids. proto gives D, Q, tile, storage, collision, LES,
 * DDF, BC set (its nx/ny/nz are ignored; method must be double buffer).
 */
typedef struct lbm_batch lbm_batch;
int lbm_batch_create(const lbm_desc* proto, int32_t n_cases, const int32_t* sizes /* n_cases*3 */,
                     const double* taus, lbm_batch** out);
int lbm_batch_destroy(lbm_batch* b);
/* NULL flags: all FLUID; NULL bc: u = 0, rho = 1. Takes effect at the next run / init. */
int lbm_batch_set_case(lbm_batch* b, int32_t c, const uint8_t* flags, const double* bc,
                       const uint32_t* link_nodes, const float* qdist, uint32_t n_links);
/* Re-lays out every case for another tile shape; the state must be initialised again. */
int lbm_batch_set_tile(lbm_batch* b, int32_t tx, int32_t ty, int32_t tz);
/*
 *   LBM_BATCH_OPT_GRID_BLOCKS: 0 = one workgroup per tile (default), n > 0 = a
 *                              persistent grid of n workgroups (rounded up to a
 *                              multiple of 8) looping over the tiles.
 *   LBM_BATCH_OPT_GRAPH_STEPS: steps per captured hipGraph (even; 0 = no graph;
 *                              default 100).
 */
enum { LBM_BATCH_OPT_GRID_BLOCKS = 2, LBM_BATCH_OPT_GRAPH_STEPS = 3 };
int lbm_batch_set_option(lbm_batch* b, int option, int value);
/* Both buffers of every case to f_eq(rho = 1, u = 0); resets the step parity. */
int lbm_batch_init_equilibrium(lbm_batch* b, void* stream);
int lbm_batch_run(lbm_batch* b, int64_t steps, void* stream);
/* Momentum-exchange force of the last step per case: out_dev[c*D + d] (device doubles). */
int lbm_batch_force(lbm_batch* b, double* out_dev, void* stream);
/* Padded single-grid description of case c and its staged tiled flags / bc (grid float
 * type) / link records, to rerun the case alone with lbm_create + lbm_set_links. */
int lbm_batch_case_desc(lbm_batch* b, int32_t c, lbm_desc* out);
int lbm_batch_export_case(lbm_batch* b, int32_t c, uint8_t* flags, void* bc, uint32_t* link_nodes,
                          float* qdist, uint32_t* n_links);
/* Current populations of case c (its tiled layout) to / from device memory; set_f writes both buffers. */
int lbm_batch_get_f(lbm_batch* b, int32_t c, void* dst_dev, void* stream);
int lbm_batch_set_f(lbm_batch* b, int32_t c, const void* src_dev, void* stream);
/*
 * Times `steps` steps for each candidate tile shape, then persistent grid sizes
 * for the best one, keeps the fastest (best = {tx, ty, tz, grid_blocks}) and
 * re-initialises the state to rest.
 */
int lbm_batch_autotune(lbm_batch* b, int32_t steps, void* stream, int32_t* best /* 4 */);
/* out = {cases, real nodes, padded nodes, tiles, tx, ty, tz, grid_blocks, graph_steps}. */
int lbm_batch_info(lbm_batch* b, int64_t* out /* 9 */);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */

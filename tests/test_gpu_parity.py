"""GPU parity of the HIP stream-collide path against the CPU oracle.

Every case runs the same seeded input through the HIP library (C ABI) and the
oracle and compares populations after a few steps. The contract is the stated
tolerance (max |df| <= 1e-6 per step in FP32); because both sides round every
operation in the same order (the library is built with -ffp-contract=off), the
populations are also expected to be bit-identical, which is asserted separately.
"""
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U

pytestmark = pytest.mark.gpu

SIZES = {"D2Q9": (32, 32, 1), "D3Q19": (16, 16, 16), "D3Q27": (16, 16, 16)}
STORES = {"f32": 0, "f16c": 1, "f64": 2}
VARIANTS = {
    "srt": dict(collision=0, les=0, ddf=0),
    "srt_ddf": dict(collision=0, les=0, ddf=1),
    "srt_les": dict(collision=0, les=1, ddf=0),
    "kbc": dict(collision=2, les=0, ddf=0),
    "kbc_les_ddf": dict(collision=2, les=1, ddf=1),
    "trt": dict(collision=1, les=0, ddf=0),
    "trt_les_ddf": dict(collision=1, les=1, ddf=1),
    "rlbm": dict(collision=3, les=0, ddf=0),
    "rlbm_les_ddf": dict(collision=3, les=1, ddf=1),
}
TAU = 0.55  # low viscosity so LES and KBC terms are active


def _case(lat, store, variant, method):
    kw = dict(VARIANTS[variant])
    if store == "f16c":
        kw["ddf"] = 1  # Float16C is used with DDF shifting
    desc = U.make_desc(lat, SIZES[lat], method=method, storage=STORES[store], **kw)
    f, bc, flags = U.perturbed_state(desc, seed=zlib.crc32(f"{lat}{store}{variant}{method}".encode()) % 1000)
    return desc, f, bc, flags


def _tol(store):
    return 1e-12 if store == "f64" else 1e-6


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("store", list(STORES))
@pytest.mark.parametrize("lat", list(SIZES))
def test_double_buffer_matches_oracle(gpu, lat, store, variant):
    desc, f, bc, flags = _case(lat, store, variant, method=0)
    print(f"case {lat} {store} {variant} db", flush=True)
    run = U.DeviceRun(desc, f, bc, flags)
    steps = 4
    for _ in range(steps):
        run.step_db(TAU)
    got = run.f_host()
    ref = U.oracle_run_db(desc, f, bc, flags, TAU, steps)
    assert U.max_abs_diff(desc, got, ref) <= _tol(store)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), "not bit-identical to the oracle"


@pytest.mark.parametrize("variant", ["srt", "srt_les", "kbc", "trt", "rlbm"])
@pytest.mark.parametrize("store", list(STORES))
@pytest.mark.parametrize("lat", list(SIZES))
def test_esoteric_pull_matches_oracle(gpu, lat, store, variant):
    desc, f, bc, flags = _case(lat, store, variant, method=1)
    print(f"case {lat} {store} {variant} ep", flush=True)
    run = U.DeviceRun(desc, f, bc, flags)
    steps = 5
    even = True
    for _ in range(steps):
        run.step_ep(TAU, even)
        even = not even
    got = run.f_host()
    ref = U.oracle_run_ep(desc, f, bc, flags, TAU, steps)
    assert U.max_abs_diff(desc, got, ref) <= _tol(store)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), "not bit-identical to the oracle"


@pytest.mark.parametrize("lat", list(SIZES))
def test_equilibrium_bc_channel(gpu, lat):
    """Equilibrium inlet (u given) and outlet (rho given, u free) with walls."""
    desc = U.make_desc(lat, SIZES[lat], eq=True)
    f, bc, flags = O.alloc(desc)
    O.init_equilibrium(desc, f)
    D = desc.D
    walls = ["+Y", "-Y"] + (["+Z", "-Z"] if D == 3 else [])
    for s in walls:
        O.set_exterior_wall(desc, flags, bc, s, 1, [0.0] * D, 1.0)
    O.set_exterior_wall(desc, flags, bc, "-X", 2, [0.05] + [0.0] * (D - 1), 1.0)
    O.set_exterior_wall(desc, flags, bc, "+X", 2, None, 1.0)
    print(f"case {lat} eq-channel", flush=True)
    run = U.DeviceRun(desc, f, bc, flags)
    for _ in range(6):
        run.step_db(0.8)
    got = run.f_host()
    ref = U.oracle_run_db(desc, f, bc, flags, 0.8, 6)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("lat,n,tile", [("D2Q9", (64, 48, 1), (16, 16, 1)), ("D2Q9", (64, 64, 1), (32, 8, 1)),
                                        ("D3Q19", (32, 32, 24), (8, 8, 8)), ("D3Q19", (32, 16, 16), (16, 4, 8)),
                                        ("D3Q27", (24, 16, 16), (8, 8, 4))])
def test_scattered_obstacles(gpu, lat, n, tile, method):
    """Solid and excluded nodes scattered over tile corners, edges and faces of a periodic box with
    a moving lid: exercises the per-tile region summary that decides which tiles take the boundary
    path (a tile next to a flagged node in its halo must not take the interior path)."""
    desc = U.make_desc(lat, n, tile=tile, method=method, les=1)
    f, bc, flags = U.perturbed_state(desc, seed=sum(n) + method, amp=1e-3, walls=False)
    rng = np.random.default_rng(sum(n))
    idx = U.index_map(desc, 1)[..., 0]
    picks = rng.integers(0, [desc.nx, desc.ny, desc.nz], size=(40, 3))
    tx, ty, tz = tile
    # snap half of the picks onto tile borders (first / last layer along each axis)
    for p in picks[:20]:
        for d, t in enumerate((tx, ty, tz)):
            if rng.random() < 0.7 and t > 1:
                p[d] = (p[d] // t) * t + (0 if rng.random() < 0.5 else t - 1)
    for k, (x, y, z) in enumerate(picks):
        flags[idx[x, y, z]] = 0x80 if k % 7 == 0 else 1  # mostly SOLID, some EXCLUDE
        bc[U.index_map(desc, desc.D + 1)[x, y, z, :]] = [0.02] * desc.D + [1.0]
    O.set_exterior_wall(desc, flags, bc, "+Y", 1, [0.05] + [0.0] * (desc.D - 1), 1.0)
    print(f"case {lat} {n} {tile} obstacles method={method}", flush=True)
    run = U.DeviceRun(desc, f, bc, flags)
    steps = 4
    if method == 0:
        for _ in range(steps):
            run.step_db(TAU)
        ref = U.oracle_run_db(desc, f, bc, flags, TAU, steps)
    else:
        for k in range(steps):
            run.step_ep(TAU, k % 2 == 0)
        ref = U.oracle_run_ep(desc, f, bc, flags, TAU, steps)
    got = run.f_host()
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("lat", list(SIZES))
def test_macroscopic_and_reduction(gpu, lat, method):
    desc, f, bc, flags = _case(lat, "f32", "srt", method)
    run = U.DeviceRun(desc, f, bc, flags)
    print(f"case {lat} macro method={method}", flush=True)
    if method == 0:
        run.step_db(TAU)
        ref_f = U.oracle_run_db(desc, f, bc, flags, TAU, 1)
        after_odd = False
    else:
        run.step_ep(TAU, True)
        run.step_ep(TAU, False)
        ref_f = U.oracle_run_ep(desc, f, bc, flags, TAU, 2)
        after_odd = True
    for layout in (0, 1):
        rho, u = run.macroscopic(after_odd, layout)
        rr, ur = O.macroscopic(desc, ref_f, bc, flags, after_odd, layout)
        assert np.array_equal(rho, rr)
        assert np.array_equal(u, ur)
    tot = run.reduce(after_odd)
    tref = O.reduce_moments(desc, ref_f, bc, flags, after_odd)
    np.testing.assert_allclose(tot, tref, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("lat,store,method", [("D2Q9", "f32", 0), ("D3Q19", "f32", 0), ("D3Q19", "f32", 1),
                                              ("D3Q27", "f64", 0), ("D3Q19", "f16c", 0), ("D3Q27", "f16c", 1)])
def test_moving_wall_nodes(gpu, lat, store, method):
    """The moving-wall term is applied per node: where a SOLID pull neighbour has a non-zero bc
    velocity (here also tiny ones and a wall whose only moving part is a single node) the term is
    added as the reference adds it, and next to resting walls it is left out (a signed zero).
    Bit-identical to the oracle, which adds the term for every SOLID neighbour."""
    desc = U.make_desc(lat, SIZES[lat], method=method, storage=STORES[store], ddf=1 if store == "f16c" else 0)
    print(f"case moving-wall nodes {lat} {store} method {method}", flush=True)
    f, bc, flags = U.perturbed_state(desc, seed=11, walls=False)
    D = desc.D
    O.set_exterior_wall(desc, flags, bc, "+Y", 1, [0.1] + [0.0] * (D - 1), 1.0)        # lid
    O.set_exterior_wall(desc, flags, bc, "-Y", 1, [0.0] * D, 1.0)                      # resting
    O.set_exterior_wall(desc, flags, bc, "+X", 1, [0.0, 1e-7] + [0.0] * (D - 2), 1.0)  # tiny velocity
    O.set_exterior_wall(desc, flags, bc, "-X", 1, [0.0] * D, 0.5)                      # resting, other density
    if D == 3:
        O.set_exterior_wall(desc, flags, bc, "-Z", 1, [0.0] * D, 1.0)
        # one moving node in an otherwise resting +Z wall
        O.set_exterior_wall(desc, flags, bc, "+Z", 1, [0.0] * D, 1.0)
        node = U.index_map(desc, D + 1)[desc.nx // 2, desc.ny // 2, desc.nz - 1]
        bc[node[2]] = -3e-3
    run = U.DeviceRun(desc, f, bc, flags)
    steps = 3
    if method == 0:
        for _ in range(steps):
            run.step_db(TAU)
        ref = U.oracle_run_db(desc, f, bc, flags, TAU, steps)
    else:
        for k in range(steps):
            run.step_ep(TAU, k % 2 == 0)
        ref = U.oracle_run_ep(desc, f, bc, flags, TAU, steps)
    got = run.f_host()
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), "populations differ next to the walls"


CAPTURE_CASES = [("D2Q9", "f32", 0, 0), ("D2Q9", "f16c", 1, 1), ("D2Q9", "f64", 0, 1), ("D3Q19", "f32", 0, 0),
                 ("D3Q19", "f16c", 1, 0), ("D3Q19", "f16c", 1, 1), ("D3Q19", "f64", 1, 0), ("D3Q27", "f32", 1, 1),
                 ("D3Q27", "f16c", 1, 0)]


@pytest.mark.parametrize("lat,store,ddf,method", CAPTURE_CASES)
def test_capture_matches_oracle(gpu, lat, store, ddf, method):
    """capture_density / capture_velocity: fluid nodes carrying CAPTURE_DENSITY, CAPTURE_VELOCITY or
    both (next to walls and the moving lid too) write their pre-collision moments into bc during the
    step; populations and the whole bc buffer bit-identical to the oracle after every step (double
    buffer and esoteric pull; the Float16C double buffer on 8x8x8 tiles takes the staged pull)."""
    print(f"case capture {lat} {store} ddf={ddf} method={method}", flush=True)
    desc = U.make_desc(lat, SIZES[lat], method=method, storage=STORES[store], ddf=ddf)
    desc.included_bcs |= 0x60
    f, bc, flags = U.perturbed_state(desc, seed=41 + method)
    rng = np.random.default_rng(7)
    fluid = flags == 0
    marks = rng.integers(0, 4, flags.size)
    flags[fluid] |= np.array([0, 0x40, 0x20, 0x60], np.uint8)[marks[fluid]]
    run = U.DeviceRun(desc, f, bc, flags)
    a, b = f.copy(), f.copy()
    bco = bc.copy()
    for k in range(3):
        if method == 0:
            run.step_db(TAU)
            O.step_db(desc, b, a, bco, flags, TAU)
            a, b = b, a
        else:
            run.step_ep(TAU, k % 2 == 0)
            O.step_ep(desc, a, bco, flags, TAU, k % 2 == 0)
        assert np.array_equal(run.f_host().view(np.uint8), a.view(np.uint8)), f"populations, step {k + 1}"
        got = run.bc.cpu().numpy()
        assert np.array_equal(got.view(np.uint8), bco.view(np.uint8)), f"bc, step {k + 1}"
    assert np.count_nonzero(bco != bc) > 0


COLUMN_CASES = [("D3Q19", "f32", 0, (16, 16, 32), False, True), ("D3Q19", "f32", 1, (24, 16, 32), False, True),
                ("D3Q19", "f16c", 0, (16, 24, 32), True, True), ("D3Q19", "f64", 1, (16, 16, 32), False, True),
                ("D3Q27", "f32", 0, (16, 16, 32), True, True), ("D3Q27", "f16c", 1, (8, 16, 64), False, True),
                # periodic boxes, with a sphere and its links in the double-buffer cases
                ("D3Q19", "f32", 0, (32, 32, 32), False, False), ("D3Q19", "f32", 1, (32, 16, 64), False, False),
                ("D3Q27", "f16c", 0, (32, 32, 32), False, False), ("D3Q19", "f64", 0, (16, 16, 64), False, False)]


@pytest.mark.parametrize("lat,store,method,grid,eq,walls", COLUMN_CASES)
def test_macroscopic_columns(gpu, lat, store, method, grid, eq, walls):
    """The fields output in columns of four z-adjacent tiles (8x8x8 tiles, a multiple of 4 tiles in
    z; FP64 and outputs off a 16-byte boundary take the tile kernel): rho and u bit-identical to the
    oracle in both u layouts, with walls, the moving lid, SOLID nodes reporting bc, an equilibrium
    outlet face and interpolated links on a sphere (double buffer), and in periodic boxes."""
    print(f"case macro columns {lat} {store} method={method} grid={grid} eq={eq} walls={walls}", flush=True)
    desc = U.make_desc(lat, grid, method=method, storage=STORES[store], ddf=1 if store == "f16c" else 0, eq=eq)
    f, bc, flags = U.perturbed_state(desc, seed=17 + method, walls=walls)
    if eq:
        O.set_exterior_wall(desc, flags, bc, "+X", 2, None, 1.0)
    links = None
    if method == 0:
        from mojo_lbm_amd import geometry as G

        gr = U.grid_of(desc)
        center = (grid[0] / 2 + 0.3, grid[1] / 2 - 0.4, grid[2] / 2 + 0.2)
        inside = G.add_sphere(gr, flags, center, 3.1)
        links = G.sphere_links(gr, flags, center, 3.1, inside)
    run = U.DeviceRun(desc, f, bc, flags)
    if links:
        run.set_links(*links)
    if method == 0:
        if links:
            run.step_db(TAU)
            a, b = f.copy(), f.copy()
            O.step_db_links(desc, b, a, bc, flags, TAU, *links)
            ref_f = b
        else:
            run.step_db(TAU)
            ref_f = U.oracle_run_db(desc, f, bc, flags, TAU, 1)
        after_odd = False
    else:
        run.step_ep(TAU, True)
        run.step_ep(TAU, False)
        ref_f = U.oracle_run_ep(desc, f, bc, flags, TAU, 2)
        after_odd = True
    for layout in (0, 1):
        kw = dict(links=links) if links else {}
        rr, ur = O.macroscopic(desc, ref_f, bc, flags, after_odd, layout, **kw)
        for misalign in (False, True):
            rho, u = run.macroscopic(after_odd, layout, misalign=misalign)
            assert np.array_equal(rho.view(np.uint8), rr.view(np.uint8)), f"rho, layout {layout}, misalign {misalign}"
            assert np.array_equal(u.view(np.uint8), ur.view(np.uint8)), f"u, layout {layout}, misalign {misalign}"

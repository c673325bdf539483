"""Launch options of a handle: the XCD tile order, static flags, 64-bit addressing, the
workgroup slab height and the staged (LDS) Float16C pull change nothing in the results (bit
for bit), and static flags still pick up a new flags buffer."""
import numpy as np
import pytest

from mojo_lbm_amd import _lib
from oracle import oracle as O
from tests import _util as U

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("option,value", [(_lib.OPT_STATIC_FLAGS, 1), (_lib.OPT_XCD_REMAP, 0)])
def test_option_keeps_results(gpu, option, value):
    desc = U.make_desc("D3Q19", (32, 32, 32))
    f, bc, flags = U.perturbed_state(desc, seed=4)
    a = U.DeviceRun(desc, f, bc, flags)
    b = U.DeviceRun(desc, f, bc, flags)
    _lib.check(b.L.lbm_set_option(b.h.ptr, option, value))
    for _ in range(5):
        a.step_db(0.6)
        b.step_db(0.6)
    assert np.array_equal(a.f_host(), b.f_host())


def test_static_flags_follow_a_new_flags_buffer(gpu):
    import torch

    desc = U.make_desc("D3Q19", (16, 16, 16))
    f, bc, flags = U.perturbed_state(desc, seed=8)
    run = U.DeviceRun(desc, f, bc, flags)
    _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_STATIC_FLAGS, 1))
    run.step_db(0.7)
    # an obstacle appears in a second flags buffer: the summary must be rebuilt for it
    flags2 = flags.copy()
    amap = U.index_map(desc, 1)[..., 0]
    flags2[amap[5:9, 5:9, 5:9].ravel()] = 1
    run.flags = torch.from_numpy(flags2).cuda()
    run.step_db(0.7)
    ref = f.copy()
    tmp = f.copy()
    O.step_db(desc, tmp, ref, bc, flags, 0.7)
    O.step_db(desc, ref, tmp, bc, flags2, 0.7)
    assert np.array_equal(run.f_host(), ref)


def test_static_flags_follow_in_place_edits(gpu):
    """Flags / bc edited in place in the same buffers: lbm_set_exterior_wall on the handle drops the
    static summary by itself; after any other in-place edit, setting the option again does."""
    import ctypes as C

    import torch

    desc = U.make_desc("D3Q19", (16, 16, 16))
    f, bc, flags = U.perturbed_state(desc, seed=12, walls=False)
    run = U.DeviceRun(desc, f, bc, flags)
    _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_STATIC_FLAGS, 1))
    run.step_db(0.7)  # summary built: every tile interior
    u = (C.c_double * 3)(0.05, 0.0, 0.0)
    _lib.check(run.L.lbm_set_exterior_wall(run.h.ptr, run.flags.data_ptr(), run.bc.data_ptr(), b"+Y", 1, u,
                                           1.0, run.stream))
    run.step_db(0.7)
    flags2, bc2 = flags.copy(), bc.copy()
    O.set_exterior_wall(desc, flags2, bc2, "+Y", 1, [0.05, 0.0, 0.0], 1.0)
    # the same two buffers as the device run, so SOLID nodes (never written) hold the same values
    a, b = f.copy(), f.copy()
    O.step_db(desc, b, a, bc, flags, 0.7)
    O.step_db(desc, a, b, bc2, flags2, 0.7)
    assert np.array_equal(run.f_host(), a)
    # an obstacle copied into the same flags buffer, then the option set again
    flags3 = flags2.copy()
    amap = U.index_map(desc, 1)[..., 0]
    flags3[amap[5:9, 4:8, 6:10].ravel()] = 1
    run.flags.copy_(torch.from_numpy(flags3))
    _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_STATIC_FLAGS, 1))
    run.step_db(0.7)
    O.step_db(desc, b, a, bc2, flags3, 0.7)
    assert np.array_equal(run.f_host(), b)


def test_masks_follow_in_place_edits_without_static_flags(gpu):
    """Without static flags the neighbour masks are rebuilt only when the flags or a wall's moving
    state change: flags edited in place (an obstacle), then only the bc of a resting wall (it starts
    moving), then nothing; every step bit-identical to the oracle, esoteric pull and double buffer."""
    import torch

    for method in (0, 1):
        desc = U.make_desc("D3Q19", (16, 16, 16), method=method)
        print(f"case in-place edits, no static flags, method {method}", flush=True)
        f, bc, flags = U.perturbed_state(desc, seed=21)  # LDC walls, moving lid
        run = U.DeviceRun(desc, f, bc, flags)
        a, b = f.copy(), f.copy()

        def step(k, bc_, flags_):
            nonlocal a, b
            if method == 0:
                run.step_db(0.7)
                O.step_db(desc, b, a, bc_, flags_, 0.7)
                a, b = b, a
            else:
                run.step_ep(0.7, k % 2 == 0)
                O.step_ep(desc, a, bc_, flags_, 0.7, k % 2 == 0)
            assert np.array_equal(run.f_host().view(np.uint8), a.view(np.uint8)), f"differs after step {k + 1}"

        step(0, bc, flags)
        flags2 = flags.copy()
        amap = U.index_map(desc, 1)[..., 0]
        flags2[amap[5:9, 4:8, 6:10].ravel()] = 1
        run.flags.copy_(torch.from_numpy(flags2))  # obstacle, same buffer
        step(1, bc, flags2)
        bc2 = bc.copy()
        side = U.index_map(desc, 4)[0, 1:15, 1:15, 1].ravel()  # the -X wall starts moving along y
        bc2[side] = 0.02
        run.bc.copy_(torch.from_numpy(bc2))
        step(2, bc2, flags2)
        step(3, bc2, flags2)  # nothing changed


ADDR64_CASES = [("D2Q9", 0, 0, 0), ("D2Q9", 1, 1, 1), ("D3Q19", 0, 0, 0), ("D3Q19", 1, 1, 1), ("D3Q19", 2, 0, 1),
                ("D3Q27", 0, 1, 0), ("D3Q27", 2, 0, 0)]


@pytest.mark.parametrize("lat,storage,ddf,method", ADDR64_CASES)
def test_addr64_kernels_match_oracle(gpu, lat, storage, ddf, method):
    """The 64-bit addressing kernels (the ones 512^3 Float16C and 256^3 FP64 buffers use) on small
    grids: populations bit-identical to the oracle after every step, macroscopic fields equal."""
    print(f"case addr64 {lat} storage={storage} ddf={ddf} method={method}", flush=True)
    n = (64, 32, 1) if lat == "D2Q9" else (16, 16, 16)
    desc = U.make_desc(lat, n, storage=storage, ddf=ddf, method=method, les=1 if lat != "D2Q9" else 0)
    f, bc, flags = U.perturbed_state(desc, seed=3 + storage)
    run = U.DeviceRun(desc, f, bc, flags)
    _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_ADDR64, 1))
    tau = 0.6
    a, b = f.copy(), f.copy()
    for k in range(4):
        if method == 0:
            run.step_db(tau)
            O.step_db(desc, b, a, bc, flags, tau)
            a, b = b, a
        else:
            run.step_ep(tau, k % 2 == 0)
            O.step_ep(desc, a, bc, flags, tau, k % 2 == 0)
        assert np.array_equal(run.f_host().view(np.uint8), a.view(np.uint8)), f"step {k + 1}"
    rho, u = run.macroscopic(after_odd=True)
    orho, ou = O.macroscopic(desc, a, bc, flags, after_odd=True)
    assert np.array_equal(rho, orho) and np.array_equal(u, ou)
    tot = run.reduce(after_odd=True)
    assert np.allclose(tot, O.reduce_moments(desc, a, bc, flags, after_odd=True), rtol=1e-12, atol=1e-12)


STEP_LANES_CASES = [("D3Q19", 1, 1, 0, (8, 8, 8), 256, (16, 16, 16)), ("D3Q19", 1, 1, 1, (8, 8, 8), 256, (16, 16, 16)),
                    ("D3Q19", 0, 0, 0, (8, 8, 8), 64, (16, 16, 16)), ("D3Q19", 0, 0, 1, (8, 4, 8), 128, (16, 16, 16)),
                    ("D3Q27", 2, 0, 0, (8, 8, 8), 128, (16, 16, 16)), ("D3Q27", 1, 1, 1, (16, 8, 4), 256, (16, 16, 16)),
                    # 12 tiles: no XCD remap (the tile count is not a multiple of 8)
                    ("D3Q19", 1, 1, 0, (8, 8, 8), 256, (24, 16, 16)), ("D3Q19", 0, 0, 1, (8, 8, 8), 128, (24, 16, 16)),
                    # one-layer slabs of a 16x16x2 tile (the Float16C default for this tile shape)
                    ("D3Q19", 1, 1, 1, (16, 16, 2), 256, (32, 16, 16))]


@pytest.mark.parametrize("lat,storage,ddf,method,tile,lanes,grid", STEP_LANES_CASES)
def test_step_lanes_match_oracle(gpu, lat, storage, ddf, method, tile, lanes, grid):
    """Several workgroups per tile (one z slab each, LBM_OPT_STEP_LANES): populations bit-identical
    to the oracle after every step, with walls, a moving lid and Smagorinsky LES."""
    print(f"case step lanes {lat} storage={storage} ddf={ddf} method={method} tile={tile} lanes={lanes} "
          f"grid={grid}", flush=True)
    desc = U.make_desc(lat, grid, tile=tile, storage=storage, ddf=ddf, method=method, les=1)
    f, bc, flags = U.perturbed_state(desc, seed=5 + storage)
    run = U.DeviceRun(desc, f, bc, flags)
    _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_STEP_LANES, lanes))
    tau = 0.6
    a, b = f.copy(), f.copy()
    for k in range(4):
        if method == 0:
            run.step_db(tau)
            O.step_db(desc, b, a, bc, flags, tau)
            a, b = b, a
        else:
            run.step_ep(tau, k % 2 == 0)
            O.step_ep(desc, a, bc, flags, tau, k % 2 == 0)
        assert np.array_equal(run.f_host().view(np.uint8), a.view(np.uint8)), f"step {k + 1}"


def test_step_lanes_with_links(gpu):
    """Interpolated links with two workgroups per tile: populations bit-identical, and the force
    records keep their whole-tile order (force to FP64 summation order)."""
    desc, f, bc, flags, nodes, qdist = U.obstacle_case("D3Q19", storage=1, ddf=1, seed=13)
    run = U.DeviceRun(desc, f, bc, flags)
    _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_STEP_LANES, desc.tx * desc.ty * desc.tz // 2))
    run.set_links(nodes, qdist)
    a, b = f.copy(), f.copy()
    for _ in range(3):
        run.step_db(0.6)
        force = O.step_db_links(desc, b, a, bc, flags, 0.6, nodes, qdist)
        a, b = b, a
    assert np.array_equal(run.f_host().view(np.uint8), a.view(np.uint8))
    assert np.allclose(run.link_force(), force, rtol=1e-12, atol=1e-15)


def test_step_lanes_errors(gpu):
    desc = U.make_desc("D3Q19", (16, 16, 16), tile=(16, 8, 4))
    f, bc, flags = U.perturbed_state(desc, seed=1)
    run = U.DeviceRun(desc, f, bc, flags)
    for bad in (32, 96, 64, 1024):  # below a wave, not a power of two, not whole layers, above the tile
        with pytest.raises(_lib.LbmError):
            _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_STEP_LANES, bad))
    _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_STEP_LANES, 256))
    _lib.check(run.L.lbm_set_option(run.h.ptr, _lib.OPT_STEP_LANES, 0))


# (lattice, collision, les, ddf, walls, grid, eq, tile): Float16C double buffer on 8 x 8 x tz tiles
# in 256-lane slabs, the shape the staged pull serves (two slabs per 8 x 8 x 8 tile, one per
# 8 x 8 x 4 tile)
STAGED_CASES = [("D3Q19", 0, 0, 1, True, (32, 16, 24), False, (8, 8, 8)),
                ("D3Q19", 0, 1, 0, True, (16, 32, 16), False, (8, 8, 8)),
                ("D3Q19", 2, 0, 1, False, (24, 16, 32), False, (8, 8, 8)),
                ("D3Q19", 1, 1, 1, True, (16, 16, 16), True, (8, 8, 8)),
                ("D3Q19", 3, 0, 0, True, (16, 24, 8), False, (8, 8, 8)),
                ("D3Q27", 0, 0, 1, True, (16, 16, 24), False, (8, 8, 8)),
                ("D3Q27", 2, 1, 1, False, (24, 16, 16), False, (8, 8, 8)),
                ("D3Q19", 0, 0, 1, True, (16, 16, 12), False, (8, 8, 4)),
                ("D3Q19", 0, 0, 1, False, (16, 8, 8), False, (8, 8, 4)),
                ("D3Q27", 1, 0, 1, False, (8, 16, 12), False, (8, 8, 4))]


@pytest.mark.parametrize("lat,coll,les,ddf,walls,grid,eq,tile", STAGED_CASES)
def test_staged_pull_matches_oracle(gpu, lat, coll, les, ddf, walls, grid, eq, tile):
    """The direct pull (LBM_OPT_STAGED_PULL 0), the staged pull per tile (1) and per x-adjacent tile
    pair (2) against the oracle: populations bit-identical after every step, periodic wrap in x, y
    and z included (periodic boxes, also with one tile per axis), walls with a moving lid,
    equilibrium nodes on one face, tiles 8 and 4 nodes high; an odd tile count in x falls back to
    one tile per workgroup."""
    print(f"case staged pull {lat} coll={coll} les={les} ddf={ddf} walls={walls} grid={grid} eq={eq} tile={tile}",
          flush=True)
    desc = U.make_desc(lat, grid, storage=1, ddf=ddf, collision=coll, les=les, eq=eq, tile=tile)
    f, bc, flags = U.perturbed_state(desc, seed=31 + coll, walls=walls)
    if eq:  # the +X face an equilibrium outlet (rho given, u free)
        O.set_exterior_wall(desc, flags, bc, "+X", 2, None, 1.0)
    runs = {}
    for mode in (0, 1, 2):  # direct, one tile per workgroup, x-adjacent tile pairs (even Tx)
        runs[mode] = U.DeviceRun(desc, f, bc, flags)
        _lib.check(runs[mode].L.lbm_set_option(runs[mode].h.ptr, _lib.OPT_STAGED_PULL, mode))
    a, b = f.copy(), f.copy()
    for k in range(4):
        for r in runs.values():
            r.step_db(0.56)
        O.step_db(desc, b, a, bc, flags, 0.56)
        a, b = b, a
        for mode, r in runs.items():
            assert np.array_equal(r.f_host().view(np.uint8), a.view(np.uint8)), f"staged {mode}, step {k + 1}"
    with pytest.raises(_lib.LbmError):
        _lib.check(runs[0].L.lbm_set_option(runs[0].h.ptr, _lib.OPT_STAGED_PULL, 3))

"""Pins the oracle's interpolated bounce-back (linear Bouzidi) and momentum-exchange force.

The reference's link-wise kernel (src/lbm/geometry/interpolated_BB.mojo:119-140) has no
golden vectors, so the oracle is pinned by what the mathematics fixes:
  * steady Couette flow between a stationary wall at an arbitrary link fraction qd and a
    halfway moving wall: the linear scheme reproduces the linear profile exactly (its
    interpolation is exact for linear fields and the bounce-back is exact at the wall), and
    the momentum exchange equals the exact wall shear nu*U/H per unit length in x and the
    pressure rho*cs^2 per unit length in y. Both branches (qd <= 1/2, qd > 1/2), both storages of
    the rest value (DDF shift on/off) and several tau are covered; a wrong coefficient,
    branch, direction or sign breaks the profile or the force by O(U);
  * qd = 1/2 reduces to halfway bounce-back bit for bit (2*0.5*a + 0*b == a);
  * the geometry's link fractions lie on the sphere.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U


def _couette(qd, tau, ddf, steps, U_wall=0.05, n=16):
    desc = U.make_desc("D2Q9", (n, n, 1), storage=2, ddf=ddf)
    f, bc, flags = O.alloc(desc)
    O.init_equilibrium(desc, f)
    O.set_exterior_wall(desc, flags, bc, "+Y", 1, [U_wall, 0.0], 1.0)
    O.set_exterior_wall(desc, flags, bc, "-Y", 1, [0.0, 0.0], 1.0)
    c, _, _ = O.lattice(2, 9)
    nodes = np.array(sorted(O.addr(desc, x, 1, 0, 0, 1) for x in range(n)), np.uint32)
    qdist = np.full((n, 8), -1.0, np.float32)
    qdist[:, c[1:, 1] == 1] = qd  # links towards the wall row y = 0
    flags[nodes] = 3
    a, b = f.copy(), f.copy()
    node_force = np.zeros(len(nodes) * 2)
    for _ in range(steps):
        force = O.step_db_links(desc, b, a, bc, flags, tau, nodes, qdist, node_force=node_force)
        a, b = b, a
    _couette.node_force = node_force.reshape(-1, 2)
    rho, u = O.macroscopic(desc, a, bc, flags, links=(nodes, qdist))
    y = np.arange(n)
    y_wall = 1.0 - float(np.float32(qd))
    y_top = n - 1.5
    exact = U_wall * (y - y_wall) / (y_top - y_wall)
    ux = u.reshape(n, n, 2)[:, :, 0]
    nu = (tau - 0.5) / 3.0
    return ux[:, 1:n - 1], exact[1:n - 1], force, nu * U_wall / (y_top - y_wall) * n, n / 3.0


@pytest.mark.parametrize("qd,tau,ddf,steps", [(0.2, 1.0, 0, 4000), (0.5, 1.0, 1, 4000), (0.7, 0.7, 1, 12000),
                                              (0.95, 1.4, 0, 6000), (0.35, 0.8, 0, 8000)])
def test_bouzidi_couette_is_exact(qd, tau, ddf, steps):
    ux, exact, force, fx, fy = _couette(qd, tau, ddf, steps)
    assert np.max(np.abs(ux - exact[None, :])) < 1e-10 * 0.05
    assert abs(force[0] - fx) < 1e-9 * fx
    assert abs(force[1] + fy) < 1e-9 * fy


def test_node_forces_of_couette_flow():
    """Per-node momentum exchange (drag.mojo:104-105): in steady Couette flow every one of the n wall
    nodes carries the same force, nu U / H in x and -rho cs^2 in y per node, and the node forces sum
    to the total."""
    _, _, force, fx, fy = _couette(0.35, 0.8, 1, 8000)
    nf = _couette.node_force
    n = len(nf)
    assert np.allclose(nf.sum(axis=0), force, rtol=1e-13, atol=0)
    assert np.all(np.abs(nf[:, 0] - fx / n) < 1e-9 * fx / n)
    assert np.all(np.abs(nf[:, 1] + fy / n) < 1e-9 * fy / n)


def test_bouzidi_couette_detects_wrong_fraction():
    # the same run read against a wall at the halfway position fails: the check is sensitive to qd
    ux, exact, _, _, _ = _couette(0.2, 1.0, 0, 4000)
    halfway = 0.05 * (np.arange(1, 15) - 0.5) / (14.5 - 0.5)
    assert np.max(np.abs(ux[0] - halfway)) > 1e-4


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_half_fraction_is_halfway_bounce_back(lat):
    desc, f, bc, flags, nodes, qdist = U.obstacle_case(lat, seed=3)
    half = np.where(qdist >= 0, np.float32(0.5), np.float32(-1.0))
    a = f.copy()
    b = f.copy()
    O.step_db(desc, a, f, bc, flags, 0.6)
    O.step_db_links(desc, b, f, bc, flags, 0.6, nodes, half)
    assert np.array_equal(a, b)
    # and a fraction away from 1/2 changes exactly the link nodes' rebuilt populations
    c = f.copy()
    O.step_db_links(desc, c, f, bc, flags, 0.6, nodes, qdist)
    assert not np.array_equal(a, c)


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19"])
def test_sphere_link_fractions_lie_on_the_sphere(lat):
    desc, f, bc, flags, nodes, qdist = U.obstacle_case(lat)
    grid = U.grid_of(desc)
    center = (16.3, 15.6, 0.0) if lat == "D2Q9" else (10.4, 7.7, 8.1)
    radius = 5.3 if lat == "D2Q9" else 4.2
    assert np.all(np.diff(nodes.astype(np.int64)) > 0)
    c, _, _ = O.lattice(desc.D, desc.Q)
    amap = U.index_map(desc, 1)[..., 0]
    pos = {int(amap[x, y, z]): (x, y, z) for x, y, z in zip(*np.nonzero(np.ones(grid.shape, bool)))}
    checked = 0
    for node, qs in zip(nodes, qdist):
        x = np.array(pos[int(node)], np.float64)
        assert flags[node] == 3
        for q in range(1, desc.Q):
            s = x - c[q]
            solid = np.sum((s - center) ** 2) <= radius ** 2
            assert (qs[q - 1] >= 0) == solid
            if solid:
                p = x + qs[q - 1] * (s - x)
                assert abs(np.linalg.norm(p - center) - radius) < 1e-4
                checked += 1
    assert checked > 20


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19"])
def test_bouzidi_conserves_mass_in_a_closed_box(lat):
    # FP64, walls on every face, an immersed obstacle with links, seeded non-equilibrium state:
    # halfway bounce-back conserves mass exactly, and the mass fix (the interpolation's mass change
    # taken from the rest population) keeps that; without it, or with the change added instead of
    # subtracted, the total drifts at O(1e-6) per step
    n = (32, 32, 1) if lat == "D2Q9" else (16, 16, 16)
    desc = U.make_desc(lat, n, storage=2)
    f, bc, flags = U.perturbed_state(desc, seed=21, amp=1e-2, u_lid=0.0)
    grid = U.grid_of(desc)
    from mojo_lbm_amd import geometry as G

    c = (15.3, 16.6, 0.0) if lat == "D2Q9" else (7.4, 8.3, 7.9)
    r = 5.2 if lat == "D2Q9" else 3.6
    inside = G.add_sphere(grid, flags, c, r)
    nodes, qdist = G.sphere_links(grid, flags, c, r, inside)
    qdist = np.where(qdist >= 0, np.float32(0.23), qdist)  # all links on the qd < 1/2 branch ...
    qdist[::2] = np.where(qdist[::2] >= 0, np.float32(0.81), qdist[::2])  # ... and half on the other
    m0 = O.reduce_moments(desc, f, bc, flags, links=(nodes, qdist))[0]
    a, b = f.copy(), f.copy()
    for _ in range(20):
        O.step_db_links(desc, b, a, bc, flags, 0.7, nodes, qdist)
        a, b = b, a
    m1 = O.reduce_moments(desc, a, bc, flags, links=(nodes, qdist))[0]
    assert abs(m1 - m0) < 1e-12 * m0


def _gap_case(qd1, qd2, ddf=0, seed=4):
    """D2Q9 FP64, node (5, 6) between SOLID nodes (4, 6) and (6, 6): links for q = 1 (pull neighbour
    x - c_1 = (4, 6)) and q = 2 (pull neighbour (6, 6)), every other neighbour fluid."""
    desc = U.make_desc("D2Q9", (16, 16, 1), storage=2, ddf=ddf)
    f, bc, flags = O.alloc(desc)
    f[:] = np.random.default_rng(seed).uniform(0.02, 0.2, f.size)
    bc[:] = 0.0
    for x in (4, 6):
        flags[O.addr(desc, x, 6, 0, 0, 1)] = 1
        bc[O.addr(desc, x, 6, 0, 2, 3)] = 1.0
    node = O.addr(desc, 5, 6, 0, 0, 1)
    flags[node] = 3
    qdist = np.full((1, 8), -1.0, np.float32)
    qdist[0, 0], qdist[0, 1] = qd1, qd2
    return desc, f, bc, flags, np.array([node], np.uint32), qdist


@pytest.mark.parametrize("qd1,qd2", [(0.3, 0.4), (0.2, 0.7), (0.8, 0.25), (0.6, 0.9)])
def test_bouzidi_one_node_gap_follows_reference(qd1, qd2):
    """In a one-node gap both links of the pair act, in the order of the reference's link bits
    (nodewise_interpolated_BB.mojo:140-178): bit q_ref = 1 rebuilds f_vec[2] first, with
    f_at_xff = f_vec[1] (the halfway bounce-back of the other wall); bit 2 then rebuilds f_vec[1]
    with f_at_xff = the just rebuilt f_vec[2]. qd > 1/2 reads f_out_of_wall = f(x, opp q_ref)
    from memory. The rest population gives back the interpolated mass (dm subtracted)."""
    desc, f, bc, flags, nodes, qdist = _gap_case(qd1, qd2)
    out = f.copy()
    O.step_db_links(desc, out, f, bc, flags, 1e30, nodes, qdist)
    c, w, opp = O.lattice(2, 9)
    x = (5, 6)
    fin = lambda xx, yy, q: f[O.addr(desc, xx % 16, yy % 16, 0, q, 9)]
    fv = np.array([fin(x[0] - c[q][0], x[1] - c[q][1], q) for q in range(9)])
    fv[1], fv[2] = fin(*x, 2), fin(*x, 1)  # halfway bounce-back from both walls
    dm = 0.0
    for q_ref in (1, 2):  # reference bit order; bit q_ref rebuilds f_vec[opp q_ref] with link qd[opp q_ref]
        o = opp[q_ref]
        qd = float(np.float32(qd1 if o == 1 else qd2))
        a = fv[o]
        if qd > 0.5:
            fb = 0.5 / qd * a + (2 * qd - 1) / (2 * qd) * fin(*x, o)
        else:
            fb = 2 * qd * a + (1 - 2 * qd) * fv[q_ref]
        dm += fb - a
        fv[o] = fb
    fv[0] -= dm
    got = np.array([out[O.addr(desc, *x, 0, q, 9)] for q in range(9)])
    np.testing.assert_allclose(got, fv, rtol=0, atol=1e-16)


def test_q_clamp_raises_small_fractions():
    """q_clamp (nodewise_interpolated_BB.mojo:151, q_dist = max(q_dist, q_clamp)): links below the
    clamp act as links at the clamp, links above it are untouched."""
    desc, f, bc, flags, nodes, qdist = _gap_case(0.05, 0.6)
    ref = f.copy()
    q2 = qdist.copy()
    q2[0, 0] = 0.2
    O.step_db_links(desc, ref, f, bc, flags, 0.8, nodes, q2)
    O.set_q_clamp(0.2)
    try:
        out = f.copy()
        O.step_db_links(desc, out, f, bc, flags, 0.8, nodes, qdist)
    finally:
        O.set_q_clamp(0.0)
    np.testing.assert_array_equal(out, ref)
    plain = f.copy()
    O.step_db_links(desc, plain, f, bc, flags, 0.8, nodes, qdist)
    assert not np.array_equal(plain, ref)

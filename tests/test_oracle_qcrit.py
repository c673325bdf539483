"""Pins of the oracle's Q-criterion (calculate_Q_criterion, src/lbm/output/Q_criterion.mojo:90-215,
and get_velocity_gradient, src/lbm/kernels/utils/finite_difference.mojo:15-136) against what
the mathematics fixes:

* solid-body rotation u = Omega x (r - r0) over populations at rest: interior Q = |Omega|^2;
  nodes on the domain edge drop the derivative across that edge (a one-sided difference
  carries weight 0 in the reference's blend), which the expected field applies per axis;
* uniform populations with a prescribed Pi_neq and no velocity: Q = -0.5 |S|^2 with
  S = Pi_neq * (-1 / (2 rho cs^2 tau)), i.e. Q = -9 eps^2 / (4 tau^2) per unit component;
* a uniform flow in equilibrium: Q = 0 (the strain uses Pi - Pi_eq, not Pi);
* a linear shear between SOLID walls sitting half a node away: Q = 0.25 gamma^2 up to the
  wall (a full-distance wall would give 0.25 (3 gamma / 4)^2 next to it).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U

CS4 = 1.0 / 9.0


def _rest_state(desc):
    f, bc, flags = O.alloc(desc)
    O.init_equilibrium(desc, f)
    return f, bc, flags


def _coords(desc):
    x, y, z = np.meshgrid(np.arange(desc.nx), np.arange(desc.ny), np.arange(desc.nz), indexing="ij")
    return x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)


def _rotation(desc, omega, r0):
    x, y, z = _coords(desc)
    r = np.stack([x - r0[0], y - r0[1], z - r0[2]], axis=-1)
    u = np.cross(np.broadcast_to(np.asarray(omega, np.float64), r.shape), r)
    return u[..., : desc.D]


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("u_layout", [0, 1])
def test_solid_body_rotation(lat, u_layout):
    n = (32, 16, 1) if lat == "D2Q9" else (16, 16, 16)
    desc = U.make_desc(lat, n, storage=2)
    f, bc, flags = _rest_state(desc)
    omega = np.array([0.0, 0.0, 1e-3]) if desc.D == 2 else np.array([3e-4, -7e-4, 5e-4])
    u = _rotation(desc, omega, (11.3, 9.6, 7.2))
    uin = u.reshape(-1) if u_layout == 0 else np.moveaxis(u, -1, 0).reshape(-1)
    q = O.q_criterion(desc, uin, f, bc, flags, 0.7, u_layout=u_layout).reshape(desc.nx, desc.ny, desc.nz)
    # expected: the exact gradient A[i, a] = d u_i / d x_a of the linear field, with the
    # derivative along axis a dropped on the two domain faces normal to a
    D = desc.D
    A = np.zeros((3, 3))
    for a in range(3):
        e = np.zeros(3)
        e[a] = 1.0
        A[:, a] = np.cross(omega, e)
    x, y, z = _coords(desc)
    pos, size = [x, y, z], [desc.nx, desc.ny, desc.nz]
    keep = [((pos[a] > 0) & (pos[a] < size[a] - 1)).astype(np.float64) for a in range(3)]
    g = lambda i, a: A[i, a] * keep[a]
    if D == 2:
        w2 = (g(1, 0) - g(0, 1)) ** 2
    else:
        w2 = (g(2, 1) - g(1, 2)) ** 2 + (g(0, 2) - g(2, 0)) ** 2 + (g(1, 0) - g(0, 1)) ** 2
    expect = 0.25 * w2
    np.testing.assert_allclose(q, expect, rtol=1e-9, atol=1e-18)
    inner = tuple(slice(1, s - 1) if s > 1 else slice(None) for s in size)
    np.testing.assert_allclose(q[inner], float(omega @ omega), rtol=1e-9)


def _uniform_strain(desc, eps, kind, ddf):
    """Populations w_q (1 + eps H_q / cs^4) with H = c_x c_y ("xy") or (c_x^2 - c_y^2) / 2 ("diag"):
    rho = 1, u = 0, Pi_neq_xy = eps (resp. Pi_neq_xx = -Pi_neq_yy = eps)."""
    c, w, _ = O.lattice(desc.D, desc.Q)
    H = c[:, 0] * c[:, 1] if kind == "xy" else 0.5 * (c[:, 0] ** 2 - c[:, 1] ** 2)
    fq = w * (1.0 + eps * H / CS4) - (w if ddf else 0.0)
    f, bc, flags = O.alloc(desc)
    idx = U.index_map(desc, desc.Q)
    f[idx.reshape(-1, desc.Q)] = fq
    return f, bc, flags


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("kind", ["xy", "diag"])
@pytest.mark.parametrize("ddf", [0, 1])
@pytest.mark.parametrize("method", [0, 1])
def test_uniform_strain(lat, kind, ddf, method):
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    desc = U.make_desc(lat, n, storage=2, ddf=ddf, method=method)
    eps, tau = 1e-3, 0.8
    f, bc, flags = _uniform_strain(desc, eps, kind, ddf)
    u0 = np.zeros(O.n_cells(desc) * desc.D)
    q = O.q_criterion(desc, u0, f, bc, flags, tau, after_odd=True)
    np.testing.assert_allclose(q, -9.0 * eps**2 / (4.0 * tau**2), rtol=1e-9)


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19"])
def test_uniform_flow_has_no_strain(lat):
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    desc = U.make_desc(lat, n, storage=2)
    f, bc, flags = O.alloc(desc)
    uvec = [0.05, -0.03, 0.02][: desc.D]
    O.init_equilibrium(desc, f, 1.02, uvec)
    uin = np.tile(np.asarray(uvec), O.n_cells(desc))
    q = O.q_criterion(desc, uin, f, bc, flags, 0.6)
    assert np.max(np.abs(q)) < 1e-20


@pytest.mark.parametrize("storage", [0, 2])
def test_shear_against_half_distance_walls(storage):
    desc = U.make_desc("D2Q9", (16, 16, 1), storage=storage)
    f, bc, flags = _rest_state(desc)
    fl = flags.copy()
    fidx = U.index_map(desc, 1)[..., 0]
    fl[fidx[:, 0, 0]] = 1    # SOLID rows y = 0 and y = 15
    fl[fidx[:, 15, 0]] = 1
    gamma = 2e-3
    _, y, _ = _coords(desc)
    ux = gamma * (y - 0.5)
    ux[:, 0, :] = 0.0                  # wall at y = 0.5 moves with u = 0
    ux[:, 15, :] = gamma * 14.0        # wall at y = 14.5
    u = np.stack([ux, np.zeros_like(ux)], axis=-1).reshape(-1)
    q = O.q_criterion(desc, u, f, bc, fl, 0.7).reshape(16, 16)
    rt = 1e-6 if storage == 0 else 1e-12
    np.testing.assert_allclose(q[:, 1:15], 0.25 * gamma**2, rtol=rt)
    assert np.all(q[:, 0] == 0) and np.all(q[:, 15] == 0)
    # the blend next to a wall: the half-distance side gets weight 2/3
    assert O.velocity_gradient(desc, u, fl, 3, 1, 0, 0, 1) == pytest.approx(gamma, rel=1e-12)


def test_one_sided_difference_has_zero_weight():
    desc = U.make_desc("D2Q9", (16, 16, 1), storage=2)
    _, _, flags = _rest_state(desc)
    x, y, _ = _coords(desc)
    u = np.stack([0.01 * x + 0.002 * y**2, 0.003 * x * y], axis=-1).reshape(-1)
    for (px, py) in [(0, 5), (15, 5), (5, 0), (5, 15)]:
        axis = 0 if px in (0, 15) else 1
        for comp in (0, 1):
            assert O.velocity_gradient(desc, u, flags, px, py, 0, comp, axis) == 0.0
    # interior: central difference of the quadratic is exact
    assert O.velocity_gradient(desc, u, flags, 6, 7, 0, 0, 1) == pytest.approx(0.004 * 7, rel=1e-12)

"""Pins of the oracle's initialize_f_from_func restatement (src/lbm/preprocess/initial_condition.mojo:86-255):

* without a gradient, each node holds f_eq(rho, u(x)): its moments give back rho and u(x),
  and a uniform field reproduces lbm_init_equilibrium's fill bit for bit;
* the gradient term carries no mass and no momentum and has the second moment
  Pi_neq = -2 rho cs^2 (tau - 1) S, S = (g + g^T) / 2 (4th-order lattice isotropy), so an
  antisymmetric gradient adds nothing;
* DDF shifting stores f - w; esoteric pull stores the same values with the odd pattern.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U


def _node_f(desc, f):
    """f as [nx, ny, nz, Q] (double buffer layout), float64."""
    idx = U.index_map(desc, desc.Q)
    v = f[idx]
    return (O.f16c_decode(v) if desc.storage == 1 else v).astype(np.float64)


def _fields(desc, seed, grad_scale=1e-3):
    rng = np.random.default_rng(seed)
    n = (desc.nx, desc.ny, desc.nz)
    u = 0.05 * rng.standard_normal(n + (desc.D,))
    g = grad_scale * rng.standard_normal(n + (desc.D, desc.D))
    return u, g


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_equilibrium_moments(lat):
    n = (32, 16, 1) if lat == "D2Q9" else (8, 8, 16)
    desc = U.make_desc(lat, n, storage=2)
    u, _ = _fields(desc, 1)
    f, _, _ = O.alloc(desc)
    O.init_f_from_fields(desc, f, 1.03, u.reshape(-1))
    fn = _node_f(desc, f)
    c, _, _ = O.lattice(desc.D, desc.Q)
    rho = fn.sum(axis=-1)
    mom = fn @ c[:, : desc.D].astype(np.float64)
    np.testing.assert_allclose(rho, 1.03, rtol=1e-14)
    np.testing.assert_allclose(mom / rho[..., None], u, atol=1e-15)


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_uniform_field_matches_uniform_fill(storage):
    desc = U.make_desc("D3Q19", (8, 8, 8), storage=storage, ddf=1)
    uvec = np.array([0.03, -0.02, 0.01])
    a, _, _ = O.alloc(desc)
    b, _, _ = O.alloc(desc)
    O.init_equilibrium(desc, a, 1.0, uvec)
    O.init_f_from_fields(desc, b, 1.0, np.tile(uvec, desc.nx * desc.ny * desc.nz))
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("tau", [0.6, 0.9, 1.7])
def test_gradient_term_moments(lat, tau):
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    desc = U.make_desc(lat, n, storage=2)
    _, g = _fields(desc, 2)
    rho = 0.98
    f0, _, _ = O.alloc(desc)
    f1, _, _ = O.alloc(desc)
    zero = np.zeros((desc.nx * desc.ny * desc.nz, desc.D))
    O.init_f_from_fields(desc, f0, rho, zero)
    O.init_f_from_fields(desc, f1, rho, zero, g.reshape(-1), tau)
    neq = _node_f(desc, f1) - _node_f(desc, f0)
    c = O.lattice(desc.D, desc.Q)[0][:, : desc.D].astype(np.float64)
    np.testing.assert_allclose(neq.sum(axis=-1), 0.0, atol=1e-15)
    np.testing.assert_allclose(neq @ c, 0.0, atol=1e-15)
    pi = np.einsum("...q,qa,qb->...ab", neq, c, c)
    S = 0.5 * (g + np.swapaxes(g, -1, -2))
    np.testing.assert_allclose(pi, -2.0 * rho * (1.0 / 3.0) * (tau - 1.0) * S, rtol=1e-12, atol=1e-15)
    # an antisymmetric gradient (a rotation) adds nothing
    f2, _, _ = O.alloc(desc)
    O.init_f_from_fields(desc, f2, rho, zero, (g - np.swapaxes(g, -1, -2)).reshape(-1), tau)
    np.testing.assert_allclose(_node_f(desc, f2), _node_f(desc, f0), atol=1e-16)


def test_ddf_and_esoteric_storage():
    desc = U.make_desc("D3Q19", (8, 8, 16), storage=2)
    u, g = _fields(desc, 3)
    plain, _, _ = O.alloc(desc)
    O.init_f_from_fields(desc, plain, 1.0, u.reshape(-1), g.reshape(-1), 0.8)
    dd = U.make_desc_like(desc, ddf_shift=1)
    shifted, _, _ = O.alloc(dd)
    O.init_f_from_fields(dd, shifted, 1.0, u.reshape(-1), g.reshape(-1), 0.8)
    w = O.lattice(3, 19)[1]
    np.testing.assert_allclose(_node_f(dd, shifted) + w, _node_f(desc, plain), atol=1e-15)
    ep = U.make_desc_like(desc, method=1)
    fe, _, _ = O.alloc(ep)
    O.init_f_from_fields(ep, fe, 1.0, u.reshape(-1), g.reshape(-1), 0.8)
    assert np.array_equal(fe, U.db_to_ep_odd(desc, plain))

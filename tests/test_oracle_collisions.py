"""Pins the oracle's TRT and RLBM collision operators (src/lbm/kernels/ops/collisions/collision.mojo:
49-113 and 168-223) with closed forms. Uniform states in periodic boxes, so one step is one
collision (streaming a uniform field changes nothing), in FP64:

TRT: a perturbation that is symmetric in every opposite pair (and carries no mass) relaxes by
(1 - 1/tau), an antisymmetric one (carrying no momentum) by (1 - 1/tau_a) with tau_a =
1/2 + (3/16)/(tau - 1/2); at tau = 1/2 + sqrt(3)/4, tau_a = tau and TRT equals SRT.
RLBM: a non-equilibrium part of the regularised form w_q/(2 cs^4) H_q : P is reproduced and
relaxed by (1 - 1/tau); a ghost mode (no mass, momentum or second moment) is removed entirely.
Both conserve mass and momentum.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U

LATS = {"D2Q9": (16, 16, 1), "D3Q19": (16, 16, 16), "D3Q27": (16, 16, 16)}


def _uniform(desc, fq):
    amap = U.index_map(desc, desc.Q)
    f = np.zeros(desc.nx * desc.ny * desc.nz * desc.Q, np.float64)
    f[amap.reshape(-1, desc.Q)] = np.broadcast_to(fq, (amap.size // desc.Q, desc.Q))
    return f


def _step_uniform(desc, fq, tau):
    f = _uniform(desc, fq)
    _, bc, flags = O.alloc(desc)
    out = f.copy()
    O.step_db(desc, out, f, bc, flags, tau)
    amap = U.index_map(desc, desc.Q)
    return out[amap[3, 2, 1 if desc.D == 3 else 0]]


def _feq(desc, rho, u):
    c, w, _ = O.lattice(desc.D, desc.Q)
    cu = c[:, :desc.D] @ np.asarray(u, np.float64)
    return w * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * np.dot(u, u))


def _moments(desc):
    c, _, _ = O.lattice(desc.D, desc.Q)
    c = c[:, :desc.D].astype(np.float64)
    rows = [np.ones(desc.Q)] + [c[:, a] for a in range(desc.D)]
    rows += [c[:, a] * c[:, b] for a in range(desc.D) for b in range(a, desc.D)]
    return np.array(rows)


def _pair_perturbation(desc, symmetric, seed):
    """A perturbation symmetric (g_q = g_q+1) or antisymmetric in each opposite pair, carrying no mass
    and no momentum, and (symmetric case) no second moment either is not required."""
    rng = np.random.default_rng(seed)
    g = np.zeros(desc.Q)
    for q in range(1, desc.Q, 2):
        a = rng.standard_normal()
        g[q], g[q + 1] = a, (a if symmetric else -a)
    # remove the mass and momentum components (keep the pair symmetry)
    M = _moments(desc)[: 1 + desc.D]
    coef = np.linalg.lstsq(M.T, g, rcond=None)[0]
    g = g - M.T @ coef
    assert np.allclose(M @ g, 0, atol=1e-14)
    return 1e-3 * g


@pytest.mark.parametrize("lat", list(LATS))
@pytest.mark.parametrize("symmetric", [True, False])
def test_trt_relaxes_pair_parts_at_their_rates(lat, symmetric):
    desc = U.make_desc(lat, LATS[lat], storage=2, collision=1)
    tau = 0.8
    rho, u = 1.02, [0.03, -0.02, 0.01][: desc.D]
    feq = _feq(desc, rho, u)
    g = _pair_perturbation(desc, symmetric, 3)
    if symmetric:
        # an antisymmetric part could only come from the momentum removal; check it is symmetric
        assert np.allclose(g[1::2], g[2::2])
    got = _step_uniform(desc, feq + g, tau)
    tau_a = 0.5 + 0.1875 / (tau - 0.5)
    rate = (1 - 1 / tau) if symmetric else (1 - 1 / tau_a)
    assert np.allclose(got, feq + rate * g, rtol=0, atol=1e-15)


@pytest.mark.parametrize("lat", list(LATS))
def test_trt_equals_srt_at_the_magic_tau(lat):
    tau = 0.5 + np.sqrt(3.0) / 4.0  # tau_a = 1/2 + (3/16)/(tau - 1/2) = tau
    d_trt = U.make_desc(lat, LATS[lat], storage=2, collision=1)
    d_srt = U.make_desc_like(d_trt, collision=0)
    f, bc, flags = U.perturbed_state(d_trt, seed=5)
    a, b = f.copy(), f.copy()
    O.step_db(d_trt, a, f, bc, flags, tau)
    O.step_db(d_srt, b, f, bc, flags, tau)
    assert np.max(np.abs(a - b)) < 1e-15


@pytest.mark.parametrize("lat", list(LATS))
def test_rlbm_reproduces_a_regularised_part_and_removes_a_ghost(lat):
    desc = U.make_desc(lat, LATS[lat], storage=2, collision=3)
    c, w, _ = O.lattice(desc.D, desc.Q)
    c = c[:, :desc.D].astype(np.float64)
    tau = 0.7
    rho, u = 0.98, [-0.02, 0.04, 0.015][: desc.D]
    feq = _feq(desc, rho, u)
    rng = np.random.default_rng(8)
    P = rng.standard_normal((desc.D, desc.D)) * 1e-3
    P = P + P.T
    H = np.einsum("qa,qb->qab", c, c) - np.eye(desc.D)[None] / 3.0
    g = w / (2 / 9) * np.einsum("qab,ab->q", H, P)
    got = _step_uniform(desc, feq + g, tau)
    assert np.allclose(got, feq + (1 - 1 / tau) * g, rtol=0, atol=1e-15)
    # a ghost: orthogonal to mass, momentum and second moments
    M = _moments(desc)
    h = rng.standard_normal(desc.Q)
    h = h - M.T @ np.linalg.lstsq(M.T, h, rcond=None)[0]
    h *= 1e-3 / np.max(np.abs(h))
    got = _step_uniform(desc, feq + h, tau)
    assert np.allclose(got, feq, rtol=0, atol=1e-15)
    # the same ghost survives SRT: the check distinguishes the operators
    got_srt = _step_uniform(U.make_desc_like(desc, collision=0), feq + h, tau)
    assert np.max(np.abs(got_srt - feq)) > 1e-5


@pytest.mark.parametrize("coll", [1, 3])
@pytest.mark.parametrize("lat", list(LATS))
def test_conservation(lat, coll):
    desc = U.make_desc(lat, LATS[lat], storage=2, collision=coll, ddf=1)
    f, bc, flags = U.perturbed_state(desc, seed=9, walls=False)
    m0 = O.reduce_moments(desc, f, bc, flags)
    a, b = f.copy(), f.copy()
    for _ in range(6):
        O.step_db(desc, b, a, bc, flags, 0.6)
        a, b = b, a
    m1 = O.reduce_moments(desc, a, bc, flags)
    assert np.allclose(m1, m0, rtol=1e-14, atol=1e-13)

"""Pins the CPU oracle (oracle/lbm_oracle.cpp) before it is trusted as the checker.

Each test fixes a part of the oracle by something other than its own formula:
the reference's velocity tables and published Float16C format, closed forms
(f_eq moments, Couette flow, Smagorinsky tau on a synthetic shear), invariants
(conservation, permutation property of streaming, esoteric == double buffer in
periodic boxes) and the Ghia et al. cavity fixture the reference uses.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U

# Velocity sets as listed in the reference (src/lbm/lattice.mojo:240-295 D3Q19,
# :141-213 D3Q27, :321-347 D2Q9) and their weights.
REF_D2Q9 = [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (-1, 1), (1, -1)]
REF_D3Q19 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
             (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1),
             (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)]
REF_D3Q27 = REF_D3Q19 + [(1, 1, 1), (-1, -1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, 1), (-1, 1, -1),
                         (-1, 1, 1), (1, -1, -1)]


@pytest.mark.parametrize("lat,ref", [("D2Q9", REF_D2Q9), ("D3Q19", REF_D3Q19), ("D3Q27", REF_D3Q27)])
def test_lattice_tables(lat, ref):
    D, Q = U.LATTICES[lat]
    c, w, opp = O.lattice(D, Q)
    assert [tuple(r[:D]) for r in c] == ref
    assert np.isclose(w.sum(), 1.0, atol=1e-15)
    for a in range(D):
        assert abs((w * c[:, a]).sum()) < 1e-15
        for b in range(D):
            assert abs((w * c[:, a] * c[:, b]).sum() - (1 / 3 if a == b else 0)) < 1e-15
    # fourth-order isotropy: sum w c_a^4 = 1/3, sum w c_a^2 c_b^2 = 1/9
    assert abs((w * c[:, 0] ** 4).sum() - 1 / 3) < 1e-15
    assert abs((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() - 1 / 9) < 1e-15
    for q in range(1, Q):
        assert np.array_equal(c[opp[q]], -c[q])
        assert opp[q] == (q + 1 if q % 2 else q - 1)  # opposites adjacent (esoteric pull requirement)


# ------------------------------------------------------------------ Float16C
def _f16c_value(h):
    """Code -> value from the format definition of Lehmann et al. (2022): sign, 4-bit exponent, 11-bit
    mantissa, exponent bias 15, e' = 0 subnormal."""
    h = np.asarray(h, np.int64)
    s = np.where(h & 0x8000, -1.0, 1.0)
    e = (h >> 11) & 0xF
    m = h & 0x7FF
    mag = np.where(e == 0, m * 2.0 ** -25, (1.0 + m / 2048.0) * np.exp2(e.astype(np.float64) - 15))
    return s * mag


def test_f16c_decode_exhaustive():
    codes = np.arange(65536, dtype=np.uint16)
    got = O.f16c_decode(codes).astype(np.float64)
    np.testing.assert_array_equal(got, _f16c_value(codes))
    assert O.f16c_decode(np.array([0x7800], np.uint16))[0] == 1.0
    assert O.f16c_decode(np.array([0x7FFF], np.uint16))[0] == 2 - 2 ** -11
    assert O.f16c_decode(np.array([0x0001], np.uint16))[0] == 2.0 ** -25
    assert O.f16c_decode(np.array([0x0800], np.uint16))[0] == 2.0 ** -14


def test_f16c_encode_roundtrip_and_monotone():
    codes = np.arange(0x8000, dtype=np.uint16)
    vals = O.f16c_decode(codes)
    np.testing.assert_array_equal(O.f16c_encode(vals), codes)
    np.testing.assert_array_equal(O.f16c_encode(-vals[1:]), codes[1:] | 0x8000)
    assert np.all(np.diff(vals.astype(np.float64)) > 0)


def test_f16c_encode_nearest_ties_away():
    pos = _f16c_value(np.arange(0x8000))
    # every midpoint between adjacent codes rounds up (away from zero), just below it rounds down
    mids = ((pos[:-1] + pos[1:]) / 2).astype(np.float32)
    exact = mids.astype(np.float64) == (pos[:-1] + pos[1:]) / 2
    enc = O.f16c_encode(mids).astype(np.int64)
    assert np.all(enc[exact] == np.arange(1, 0x8000)[exact])
    below = np.nextafter(mids, np.float32(0))
    assert np.all(O.f16c_encode(below).astype(np.int64) == np.arange(0x7FFF))
    # brute force: random values encode to the nearest code
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(0, 2, 20000), np.exp2(rng.uniform(-30, 1, 20000))]).astype(np.float32)
    x = x[x < 2 - 2 ** -12]
    enc = O.f16c_encode(x).astype(np.int64)
    err = np.abs(pos[enc] - x)
    lo = np.abs(pos[np.maximum(enc - 1, 0)] - x)
    hi = np.abs(pos[np.minimum(enc + 1, 0x7FFF)] - x)
    assert np.all(err <= lo) and np.all(err <= hi)


def test_f16c_saturation_and_specials():
    sat = np.array([2 - 2 ** -12, 2.0, 3.5, np.inf, np.nan, 1e30], np.float32)
    assert np.all(O.f16c_encode(sat) == 0x7FFF)
    assert np.all(O.f16c_encode(-sat[:4]) == 0xFFFF)
    tiny = np.array([0.0, 2.0 ** -27, 2.0 ** -26 * 0.999], np.float32)
    assert np.all(O.f16c_encode(tiny) == 0)
    assert O.f16c_encode(np.array([2.0 ** -26], np.float32))[0] == 1  # tie rounds away from zero
    assert O.f16c_encode(np.array([-0.0], np.float32))[0] == 0x8000


# ------------------------------------------------------------------ layout
def _blocked_index(n, t, K):
    """(x,y,z,k) -> address built as a numpy reshape of a col-major tile (tx,ty,tz,K) inside a col-major
    tiler (Tx,Ty,Tz): independent of the oracle's arithmetic formula."""
    nx, ny, nz = n
    tx, ty, tz = t
    Tx, Ty, Tz = nx // tx, ny // ty, nz // tz
    addr = np.arange(nx * ny * nz * K).reshape(Tz, Ty, Tx, K, tz, ty, tx)  # slowest .. fastest
    # axes -> (x = (Tx, tx), y = (Ty, ty), z = (Tz, tz), k)
    a = addr.transpose(2, 6, 1, 5, 0, 4, 3).reshape(nx, ny, nz, K)
    return a


@pytest.mark.parametrize("n,t,K", [((16, 8, 4), (8, 4, 2), 19), ((32, 16, 1), (16, 16, 1), 9), ((8, 8, 8), (8, 8, 8), 4)])
def test_layout_matches_blocked_product(n, t, K):
    D = 2 if n[2] == 1 else 3
    desc = U.make_desc("D2Q9" if D == 2 else "D3Q19", n, tile=t)
    ref = _blocked_index(n, t, K)
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y, z, k = (int(rng.integers(0, n[0])), int(rng.integers(0, n[1])), int(rng.integers(0, n[2])),
                      int(rng.integers(0, K)))
        assert O.addr(desc, x, y, z, k, K) == ref[x, y, z, k]
    assert len(np.unique(ref)) == ref.size


# ------------------------------------------------------------------ streaming / bounce-back
def _periodic_f64(lat, n, tau=1e30, **kw):
    desc = U.make_desc(lat, n, storage=2, **kw)
    f, bc, flags = O.alloc(desc)
    O.init_equilibrium(desc, f)
    return desc, f, bc, flags


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_pull_streaming_moves_each_population_one_node(lat):
    n = (32, 16, 1) if lat == "D2Q9" else (16, 8, 8)
    desc, f, bc, flags = _periodic_f64(lat, n)
    D, Q = desc.D, desc.Q
    c, w, _ = O.lattice(D, Q)
    x0 = (0, 3, 0 if D == 2 else 7)  # on a wrap boundary in x and z
    for q in range(Q):
        g = f.copy()
        g[O.addr(desc, *x0, q, Q)] += 0.25
        out = g.copy()
        O.step_db(desc, out, g, bc, flags, 1e30)  # tau -> inf: collision leaves f unchanged
        dst = [(x0[i] + c[q][i]) % n[i] if i < D else 0 for i in range(3)]
        assert out[O.addr(desc, *dst, q, Q)] == w[q] + 0.25
        assert np.count_nonzero(out != f) == 1


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_streaming_is_a_permutation(lat, method):
    """With tau -> inf a step only moves values: every slot is written exactly once (esoteric slot
    ownership) and walls reflect instead of losing populations."""
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    desc = U.make_desc(lat, n, storage=2, method=method)
    f, bc, flags = O.alloc(desc)
    f[:] = np.random.default_rng(3).uniform(0.01, 0.2, f.size)
    if method == 0:
        out = f.copy()
        O.step_db(desc, out, f, bc, flags, 1e30)
    else:
        out = f.copy()
        O.step_ep(desc, out, bc, flags, 1e30, True)
    np.testing.assert_array_equal(np.sort(out), np.sort(f))


def test_halfway_bounce_back_returns_in_one_step():
    desc, f, bc, flags = _periodic_f64("D3Q19", (8, 8, 8))
    O.set_exterior_wall(desc, flags, bc, "-X", 1, [0.0, 0.0, 0.0], 1.0)
    c, w, opp = O.lattice(3, 19)
    x = (1, 4, 4)  # next to the wall at x = 0
    for q in range(1, 19):
        if c[q][0] != -1:
            continue
        g = f.copy()
        g[O.addr(desc, *x, q, 19)] += 0.125
        out = g.copy()
        O.step_db(desc, out, g, bc, flags, 1e30)
        assert out[O.addr(desc, *x, opp[q], 19)] == w[q] + 0.125


def test_moving_wall_momentum_injection():
    desc, f, bc, flags = _periodic_f64("D3Q19", (8, 8, 8))
    U_w = (0.07, 0.0, -0.02)
    O.set_exterior_wall(desc, flags, bc, "+Y", 1, list(U_w), 1.3)
    out = f.copy()
    O.step_db(desc, out, f, bc, flags, 1e30)
    c, w, opp = O.lattice(3, 19)
    x = (3, 6, 2)  # under the lid
    for q in range(19):
        v = out[O.addr(desc, *x, q, 19)]
        if c[q][1] == -1:  # pulled from the lid: bounce-back of w_opp plus 6 w rho_w (c . u_w)
            expect = w[opp[q]] + 6 * w[q] * 1.3 * float(np.dot(c[q], U_w))
        else:
            expect = w[q]
        assert abs(v - expect) < 1e-15


# ------------------------------------------------------------------ collision
@pytest.mark.parametrize("coll", [0, 2])
@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_equilibrium_is_fixed_point_and_moments(lat, coll):
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    desc = U.make_desc(lat, n, storage=2, collision=coll)
    f, bc, flags = O.alloc(desc)
    u0 = [0.05, -0.03, 0.02][: desc.D]
    O.init_equilibrium(desc, f, 1.07, u0)
    rho, u = O.macroscopic(desc, f, bc, flags)
    np.testing.assert_allclose(rho, 1.07, rtol=1e-14)
    np.testing.assert_allclose(u.reshape(-1, desc.D), np.tile(u0, (rho.size, 1)), atol=1e-15)
    out = f.copy()
    O.step_db(desc, out, f, bc, flags, 0.6)
    np.testing.assert_allclose(out, f, rtol=0, atol=1e-15)


@pytest.mark.parametrize("variant", [dict(collision=0), dict(collision=2), dict(collision=0, les=1),
                                     dict(collision=2, les=1, ddf=1)])
@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_periodic_conservation(lat, variant):
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    desc = U.make_desc(lat, n, storage=2, **variant)
    f, bc, flags = U.perturbed_state(desc, seed=5, amp=5e-3, walls=False)
    before = O.reduce_moments(desc, f, bc, flags)
    after = O.reduce_moments(desc, U.oracle_run_db(desc, f, bc, flags, 0.52, 10), bc, flags)
    np.testing.assert_allclose(after, before, rtol=1e-14, atol=1e-13)


def test_ddf_shift_agrees_with_plain_storage():
    d0 = U.make_desc("D3Q19", (8, 8, 8), storage=2, ddf=0, les=1)
    d1 = U.make_desc_like(d0, ddf_shift=1)
    f0, bc, flags = U.perturbed_state(d0, seed=9, amp=2e-3)
    c, w, _ = O.lattice(3, 19)
    wf = np.tile(np.repeat(w, 512), f0.size // (19 * 512))  # one 8^3 tile = 512 nodes per q plane
    f1 = f0 - wf
    a = U.oracle_run_db(d0, f0, bc, flags, 0.6, 20)
    b = U.oracle_run_db(d1, f1, bc, flags, 0.6, 20)
    np.testing.assert_allclose(b + wf, a, rtol=0, atol=1e-13)


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_esoteric_equals_double_buffer_periodic(lat):
    """In periodic all-fluid boxes the esoteric and double-buffer schemes give identical moments
    after every even step (FP32, bitwise)."""
    n = (32, 16, 1) if lat == "D2Q9" else (16, 8, 8)
    d_db = U.make_desc(lat, n, storage=0, method=0, les=1)
    d_ep = U.make_desc_like(d_db, method=1)
    f, bc, flags = U.perturbed_state(d_db, seed=2, amp=3e-3, walls=False)
    a = U.oracle_run_db(d_db, f, bc, flags, 0.56, 6)
    b = U.oracle_run_ep(d_ep, U.db_to_ep_odd(d_ep, f), bc, flags, 0.56, 6)
    ra, ua = O.macroscopic(d_db, a, bc, flags)
    rb, ub = O.macroscopic(d_ep, b, bc, flags, after_odd=True)
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(ua, ub)


def _uniform_state_with_stress(desc, rho, u, pi):
    """f = f_eq(rho, u) + f_neq with f_neq_q = w_q / (2 cs^4) H_q : Pi at every node (FP64)."""
    D, Q = desc.D, desc.Q
    c, w, _ = O.lattice(D, Q)
    cc = c[:, :D].astype(np.float64)
    u = np.asarray(u, np.float64)
    cu = cc @ u
    feq = w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
    H = np.einsum("qa,qb->qab", cc, cc) - np.eye(D) / 3
    fneq = w * 4.5 * np.einsum("qab,ab->q", H, pi)
    fv = feq + fneq
    f = np.repeat(fv, 1)
    tile = desc.tx * desc.ty * desc.tz
    ntiles = (desc.nx * desc.ny * desc.nz) // tile
    return np.tile(np.repeat(fv, tile), ntiles), feq, fneq


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19"])
def test_smagorinsky_tau_closed_form(lat):
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    desc = U.make_desc(lat, n, storage=2, les=1, cs=0.1)
    D = desc.D
    rho, u = 1.02, [0.03, -0.01, 0.02][:D]
    pi = np.array([[2e-4, 3e-4, -1e-4], [3e-4, -1e-4, 2e-4], [-1e-4, 2e-4, 5e-5]])[:D, :D]
    f, feq, fneq = _uniform_state_with_stress(desc, rho, u, pi)
    _, bc, flags = O.alloc(desc)
    tau = 0.52
    out = f.copy()
    O.step_db(desc, out, f, bc, flags, tau)
    S = -pi / (2 * rho * tau / 3)
    tau_total = tau + 3 * 0.1 ** 2 * math.sqrt(2 * np.sum(S * S))
    expect = feq + (1 - 1 / tau_total) * fneq
    np.testing.assert_allclose(out[:: desc.tx * desc.ty * desc.tz][: desc.Q], expect, rtol=0, atol=1e-14)


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_kbc_at_tau_one_is_srt(lat):
    """beta = 1/2 makes gamma = inv_beta - (2 - inv_beta) * (...) = 2 exactly, and KBC's
    f - beta (2 ds + gamma dh) is then the SRT update f - (f - f_eq)."""
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    d_kbc = U.make_desc(lat, n, storage=2, collision=2)
    d_srt = U.make_desc_like(d_kbc, collision=0)
    f, bc, flags = U.perturbed_state(d_kbc, seed=11, amp=4e-3, walls=False)
    a = U.oracle_run_db(d_kbc, f, bc, flags, 1.0, 3)
    b = U.oracle_run_db(d_srt, f, bc, flags, 1.0, 3)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_kbc_shear_and_bulk_split(lat):
    """f_neq = shear (traceless stress) + bulk (trace) Hermite parts at u = 0. The parts are
    orthogonal in the entropic product, so sp1 = 0, gamma = 1/beta and KBC returns
    f_eq + (1 - 1/tau) * shear: the shear part relaxes at the SRT rate, the bulk part to
    equilibrium. For D2Q9 this pins get_shear_D2Q9 (KBC_.mojo:137-175); for D3Q19 / D3Q27
    the same split."""
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    d = U.make_desc(lat, n, storage=2, collision=2)
    D = d.D
    S = np.array([[3e-4, 2e-4, -1e-4], [2e-4, -1e-4, 1e-4], [-1e-4, 1e-4, -2e-4]])[:D, :D]
    S = S - np.trace(S) / D * np.eye(D)
    B = 2.5e-4 * np.eye(D)
    f, feq, _ = _uniform_state_with_stress(d, 1.0, [0.0] * D, S + B)
    _, _, fshear = _uniform_state_with_stress(d, 1.0, [0.0] * D, S)
    _, bc, flags = O.alloc(d)
    tau = 0.6
    out = f.copy()
    O.step_db(d, out, f, bc, flags, tau)
    tile = d.tx * d.ty * d.tz
    np.testing.assert_allclose(out[::tile][: d.Q], feq + (1 - 1 / tau) * fshear, rtol=0, atol=1e-15)


def test_equilibrium_bc_resets_node():
    desc = U.make_desc("D3Q19", (8, 8, 8), storage=2, eq=True)
    f, bc, flags = U.perturbed_state(desc, seed=4, amp=1e-3, walls=False)
    O.set_exterior_wall(desc, flags, bc, "-X", 2, [0.04, 0.01, 0.0], 1.01)
    out = f.copy()
    O.step_db(desc, out, f, bc, flags, 0.7)
    c, w, _ = O.lattice(3, 19)
    u = np.array([0.04, 0.01, 0.0])
    cu = c.astype(float) @ u
    feq = w * 1.01 * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
    for y in range(8):
        got = np.array([out[O.addr(desc, 0, y, 3, q, 19)] for q in range(19)])
        np.testing.assert_allclose(got, feq, rtol=0, atol=1e-15)


def _feq_np(c, w, rho, u):
    cu = c @ u
    return w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)


@pytest.mark.parametrize("ddf", [0, 1])
@pytest.mark.parametrize("free", ["u", "rho", "both"])
@pytest.mark.parametrize("side", ["-X", "+X"])
@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19"])
def test_equilibrium_bc_free_moments(lat, side, free, ddf):
    """EQUILIBRIUM column with a free (NaN) velocity and/or density: the node is reset to
    f_eq(rho*, u*) where (rho*, u*) are the moments of the pulled populations, each pull that
    leaves the grid on any axis replaced by the rest value w_q (0 with DDF) and rho* += 1 with
    DDF (boundary_condition.mojo:87-204). Expected values are built here node by node from the
    stored field. The column on the opposite face holds a different state, so a pull that wraps
    round instead of taking the rest value, a rest value of the wrong kind, a pull in the wrong
    direction or a swapped NaN selection all change the result by O(1e-2)."""
    n = (16, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    desc = U.make_desc(lat, n, storage=2, eq=True, ddf=ddf)
    D, Q = desc.D, desc.Q
    c, w, _ = O.lattice(D, Q)
    cc = c[:, :D].astype(np.float64)
    rho0, u0 = 1.03, np.array([0.04, -0.02, 0.01][:D])
    rho1, u1 = 0.97, np.array([-0.03, 0.05, -0.02][:D])
    shift = w if ddf else 0.0
    fv0 = _feq_np(cc, w, rho0, u0) - shift
    fv1 = _feq_np(cc, w, rho1, u1) - shift
    f, bc, flags = O.alloc(desc)
    xo = desc.nx - 1 if side == "-X" else 0  # the column a wrapped pull would reach
    for z in range(n[2] if D == 3 else 1):
        for y in range(n[1]):
            for x in range(n[0]):
                fv = fv1 if x == xo else fv0
                for q in range(Q):
                    f[O.addr(desc, x, y, z, q, Q)] = fv[q]
    u_bc = np.array([0.02, 0.01, -0.015][:D])
    rho_bc = 1.05
    O.set_exterior_wall(desc, flags, bc, side, 2, None if free in ("u", "both") else list(u_bc),
                        float("nan") if free in ("rho", "both") else rho_bc)
    out = f.copy()
    O.step_db(desc, out, f, bc, flags, 0.7)
    xe = 0 if side == "-X" else desc.nx - 1
    rest = np.zeros(Q) if ddf else w
    nn = [desc.nx, desc.ny, desc.nz]
    n_oob_seen = set()
    for z in range(n[2] if D == 3 else 1):
        for y in range(n[1]):
            idx = (xe, y, z)
            fq = np.empty(Q)
            for q in range(Q):
                p = [idx[i] - (c[q][i] if i < D else 0) for i in range(3)]
                oob = any(p[i] < 0 or p[i] >= nn[i] for i in range(D))
                if oob:
                    fq[q] = rest[q]
                else:
                    fq[q] = (fv1 if p[0] == xo else fv0)[q]
            n_oob_seen.add(int(sum(any((idx[i] - c[q][i]) < 0 or (idx[i] - c[q][i]) >= nn[i]
                                        for i in range(D)) for q in range(Q))))
            rho_l = fq.sum() + (1.0 if ddf else 0.0)
            u_l = (fq @ cc) / rho_l
            rho_s = rho_l if free in ("rho", "both") else rho_bc
            u_s = u_l if free in ("u", "both") else u_bc
            expect = _feq_np(cc, w, rho_s, u_s) - shift
            got = np.array([out[O.addr(desc, xe, y, z, q, Q)] for q in range(Q)])
            np.testing.assert_allclose(got, expect, rtol=0, atol=2e-15)
    assert len(n_oob_seen) >= 2  # face nodes and edge/corner nodes both covered


def test_reduce_moments_uniform_totals():
    """A uniform equilibrium gives N * rho0 and N * rho0 * u0 (periodic box, rho0 != 1 so that
    rho * u cannot pass for u); at rest with walls only the N_valid non-SOLID, non-EXCLUDE nodes
    count, whatever the walls' bc holds."""
    desc = U.make_desc("D3Q19", (16, 8, 8), storage=2)
    f, bc, flags = O.alloc(desc)
    rho0, u0 = 1.07, [0.03, -0.05, 0.02]
    O.init_equilibrium(desc, f, rho0, u0)
    tot = O.reduce_moments(desc, f, bc, flags)
    N = 16 * 8 * 8
    np.testing.assert_allclose(tot, [N * rho0] + [N * rho0 * v for v in u0], rtol=1e-13)
    O.init_equilibrium(desc, f, rho0, [0.0, 0.0, 0.0])
    O.set_exterior_wall(desc, flags, bc, "+Y", 1, [0.1, 0.0, 0.0], 3.0)
    O.set_exterior_wall(desc, flags, bc, "-Y", 1, [0.0, 0.0, 0.0], 3.0)
    ids = [O.addr(desc, x, 3, 4, 0, 1) for x in range(5, 9)]
    flags[ids] |= 0x80  # EXCLUDE
    tot = O.reduce_moments(desc, f, bc, flags)
    n_valid = N - 2 * 16 * 8 - 4
    assert abs(tot[0] - n_valid * rho0) < 1e-12 * n_valid
    # the lid's moving-wall term carries momentum into the nodes under it: 6 w rho_w (c . u_w) per
    # pulled link, summed over the 5 links that reach the lid from each of the 16 * 8 nodes
    c, w, _ = O.lattice(3, 19)
    inj = sum(6 * w[q] * 3.0 * c[q][0] * 0.1 * c[q][0] for q in range(19) if c[q][1] == -1)
    mass_inj = sum(6 * w[q] * 3.0 * c[q][0] * 0.1 for q in range(19) if c[q][1] == -1)
    assert abs(mass_inj) < 1e-16  # the +-x links cancel in the mass
    assert abs(tot[1] - 16 * 8 * inj) < 1e-12
    assert abs(tot[2]) < 1e-12 and abs(tot[3]) < 1e-12


@pytest.mark.parametrize("method", [0, 1])
def test_reduce_moments_equal_field_sums(method):
    """The FP64 totals equal numpy sums of the (separately pinned) macroscopic fields over the
    nodes that are neither SOLID nor EXCLUDE, on a perturbed walled state."""
    desc = U.make_desc("D3Q19", (16, 8, 8), storage=2, method=method)
    f, bc, flags = U.perturbed_state(desc, seed=21, amp=2e-3)
    rng = np.random.default_rng(3)
    ids = rng.choice(f.size // 19, 40, replace=False)
    flags[ids] |= 0x80
    if method == 1:
        f = U.db_to_ep_odd(desc, f)
    rho, u = O.macroscopic(desc, f, bc, flags, after_odd=True)
    tot = O.reduce_moments(desc, f, bc, flags, after_odd=True)
    fl = np.zeros(rho.size, np.uint8)  # flags in row-major (x, y, z) node order
    for i in range(rho.size):
        x, y, z = i // 64, (i // 8) % 8, i % 8
        fl[i] = flags[O.addr(desc, x, y, z, 0, 1)]
    valid = ((fl & 0x0F) != 1) & ((fl & 0x80) == 0)
    assert 0 < valid.sum() < rho.size
    uu = u.reshape(-1, 3)
    assert abs(tot[0] - math.fsum(rho[valid])) < 1e-12
    for d in range(3):
        assert abs(tot[1 + d] - math.fsum(rho[valid] * uu[valid, d])) < 1e-13


def test_couette_profile():
    """D2Q9 channel, periodic in x, bottom wall at rest, top wall moving: the steady profile is linear
    between the halfway walls (y = 0.5 and y = ny - 1.5)."""
    ny, Uw = 18, 0.05
    d = U.make_desc("D2Q9", (16, ny, 1), tile=(16, 2, 1), storage=2)
    f, bc, flags = O.alloc(d)
    O.init_equilibrium(d, f)
    O.set_exterior_wall(d, flags, bc, "-Y", 1, [0.0, 0.0], 1.0)
    O.set_exterior_wall(d, flags, bc, "+Y", 1, [Uw, 0.0], 1.0)
    g = f.copy()
    for _ in range(3000):
        O.step_db(d, g, f, bc, flags, 1.0)
        f, g = g, f
    _, u = O.macroscopic(d, f, bc, flags)
    ux = u.reshape(16, ny, 2)[5, 1:-1, 0]
    y = np.arange(1, ny - 1)
    np.testing.assert_allclose(ux, Uw * (y - 0.5) / (ny - 2), rtol=0, atol=1e-9)


@pytest.mark.slow
def test_ldc_d2q9_re100_matches_ghia():
    """BL-cfg1 physics check: D2Q9 128^2 cavity, Re = 100, u_lid = 0.1, 10,000 steps of FP32 pull
    streaming; centreline profiles against Ghia et al. (loose: the flow is still developing and
    the grid coarse)."""
    N = 128
    d = U.make_desc("D2Q9", (N, N, 1), storage=0)
    f, bc, flags = O.alloc(d)
    O.init_equilibrium(d, f)
    U.ldc_walls(d, flags, bc, 0.1)
    tau = 0.5 + 3 * 0.1 * N / 100
    f = U.oracle_run_db(d, f, bc, flags, tau, 10000)
    _, u = O.macroscopic(d, f, bc, flags)
    u = u.reshape(N, N, 2) / 0.1
    import os

    here = os.path.dirname(__file__)
    gu = np.loadtxt(os.path.join(here, "golden", "ghia_u_centerline.txt"), skiprows=1)
    gv = np.loadtxt(os.path.join(here, "golden", "ghia_v_centerline.csv"), skiprows=1, delimiter=",")
    # cavity spans the halfway walls: fluid nodes 1..N-2, walls at 0.5 and N-1.5
    pos = (np.arange(N) - 0.5) / (N - 2)
    uc = 0.5 * (u[N // 2 - 1, :, 0] + u[N // 2, :, 0])
    vc = 0.5 * (u[:, N // 2 - 1, 1] + u[:, N // 2, 1])
    iu = np.interp(gu[1:-1, 0], pos[1:-1], uc[1:-1])
    iv = np.interp(gv[1:-1, 0], pos[1:-1], vc[1:-1])
    assert np.max(np.abs(iu - gu[1:-1, 1])) < 0.05
    assert np.max(np.abs(iv - gv[1:-1, 1])) < 0.05


@pytest.mark.parametrize("ddf", [0, 1])
@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
def test_capture_writes_pre_collision_moments(lat, ddf):
    """capture_density / capture_velocity (steps/collide.mojo:54-63): in a periodic FP64 box the
    populations a node collides are f_in(x - c_q, q), so the captured density and velocity equal
    numpy's moments of those rolled values; only flagged nodes write, only the enabled kind, and
    every other bc entry keeps its value. A capture of post-collision moments would agree here up to
    rounding (SRT conserves both), a missing DDF +1, a wrong component order or a capture at an
    unflagged node would not."""
    n = (32, 16, 1) if lat == "D2Q9" else (8, 8, 8)
    D = 2 if lat == "D2Q9" else 3
    for mode in ("density", "velocity", "both", "off"):
        desc = U.make_desc(lat, n, storage=2, ddf=ddf)
        bits = {"density": 0x40, "velocity": 0x20, "both": 0x60, "off": 0}[mode]
        desc.included_bcs |= bits
        f, bc, flags = O.alloc(desc)
        rng = np.random.default_rng(5 + ddf)
        f[:] = rng.uniform(-0.01, 0.05, f.size) if ddf else rng.uniform(0.02, 0.2, f.size)
        bc[:] = -7.0
        ids1 = U.index_map(desc, 1)[..., 0]
        marks = rng.integers(0, 4, n)  # 0 none, 1 density, 2 velocity, 3 both
        flag_of = np.array([0, 0x40, 0x20, 0x60], np.uint8)
        flags[ids1] = flag_of[marks]
        out = f.copy()
        O.step_db(desc, out, f, bc, flags, 0.8)
        c, _, _ = O.lattice(D, desc.Q)
        idsq = U.index_map(desc, desc.Q)
        fin = f[idsq]  # (nx, ny, nz, Q)
        cc = np.zeros((desc.Q, 3), int)
        cc[:, :D] = np.array(c)[:, :D]
        post = np.stack([np.roll(fin[..., q], shift=tuple(cc[q]), axis=(0, 1, 2)) for q in range(desc.Q)], -1)
        rho = post.sum(-1) + (1.0 if ddf else 0.0)
        u = np.einsum("xyzq,qd->xyzd", post, cc[:, :D].astype(float)) / rho[..., None]
        idsb = U.index_map(desc, D + 1)
        got = bc[idsb]
        want = np.full(got.shape, -7.0)
        if bits & 0x40:
            sel = (marks & 1) == 1
            want[..., D][sel] = rho[sel]
        if bits & 0x20:
            sel = (marks & 2) == 2
            want[..., :D][sel] = u[sel]
        assert np.allclose(got, want, rtol=1e-12, atol=1e-15), mode
        if mode == "both":
            assert np.any(got != -7.0)

"""Host-side mirror of the reference API (CPU parts): config toggles, lattice
tables, grid launch geometry and layout, tau from Re, and the host wall setup
checked byte for byte against the oracle's restatement of set_exterior_walls."""
import math

import numpy as np
import pytest

import mojo_lbm_amd as lbm
from mojo_lbm_amd.solver import make_desc
from oracle import oracle as O


def test_config_defaults_and_implications():
    c = lbm.DoubleBufferConfig()
    assert c.collision_op == "SRT" and c.include_moving_boundary and not c.LES
    assert c.INCLUDED_BCs == {lbm.Flags.FLUID, lbm.Flags.SOLID}
    assert c.implies_get_adjacent_flags()
    e = lbm.EsotericPullConfig(include_moving_boundary=False, LES=True, BCs={lbm.Flags.EQUILIBRIUM})
    assert not e.implies_get_adjacent_flags() and e.implies_f_noneq()
    assert e.included_bcs_mask == 0b111
    k = lbm.DoubleBufferConfig(collision_op="KBC")
    assert k.implies_f_noneq()
    f = lbm.DoubleBufferConfig(use_float16c=True, DDF_shift=True)
    assert f.f_dtype is np.uint16 and f.storage_id == 1
    assert lbm.DoubleBufferConfig(f_dtype=np.float64).storage_id == 2
    with pytest.raises(ValueError):
        lbm.DoubleBufferConfig(collision_op="MRT")


def test_lattices_match_oracle_tables():
    for lat, (D, Q) in [(lbm.get_D2Q9(), (2, 9)), (lbm.get_D3Q19(), (3, 19)), (lbm.get_D3Q27(), (3, 27))]:
        c, w, opp = O.lattice(D, Q)
        assert np.array_equal(lat.directions, c[:, :D])
        assert np.array_equal(lat.weights, w)
        assert np.array_equal(lat.opposite_indices, opp)
        assert lat.is_valid_for_esoteric_pull()


def test_grid_geometry():
    g = lbm.LBM_Grid(lbm.get_D3Q19(), 256, 256, 256, 8)
    assert g.BLOCK_SHAPE == (8, 8, 8) and g.GRID_DIM == (32, 32, 32)
    g2 = lbm.LBM_Grid(lbm.get_D2Q9(), 128, 128, 1, 16)
    assert g2.tile_shape == (16, 16, 1) and g2.GRID_DIM == (8, 8, 1)
    with pytest.raises(ValueError):
        lbm.LBM_Grid(lbm.get_D3Q19(), 20, 16, 16, 8)


def test_layout_index_matches_oracle():
    g = lbm.LBM_Grid(lbm.get_D3Q19(), 16, 8, 24, (4, 8, 2))
    d = make_desc(g, lbm.DoubleBufferConfig())
    rng = np.random.default_rng(1)
    for _ in range(100):
        x, y, z, k = (int(rng.integers(0, 16)), int(rng.integers(0, 8)), int(rng.integers(0, 24)),
                      int(rng.integers(0, 19)))
        assert g.layouts.index(x, y, z, k, 19) == O.addr(d, x, y, z, k, 19)


def test_tau_from_re():
    # BL-cfg1/2/3/5 relaxation times: tau = 0.5 + 3 u N / Re
    assert math.isclose(lbm.tau_from_re(0.1, 128, 100), 0.884)
    assert math.isclose(lbm.tau_from_re(0.1, 256, 100), 1.268)
    assert math.isclose(lbm.tau_from_re(0.1, 256, 1000), 0.5768)
    assert math.isclose(lbm.tau_from_re(0.05, 128, 1e4), 0.50192)
    g = lbm.LBM_Grid(lbm.get_D2Q9(), 128, 128, 1, 16, dx=1.0 / 127)
    us = g.get_UnitSystem_with_Re(1.0, 0.1, 1.0, Re=100)
    assert math.isclose(us.tau, 0.5 + 3 * 0.1 * 127 / 100)  # L_lat = N - 1 as in the reference scripts


class _HostOnly:
    """Host-side stand-in for a ContextTileTensor (set_exterior_walls only touches .cpu())."""

    def __init__(self, arr):
        self.arr = arr

    def cpu(self):
        return self.arr


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19"])
def test_set_exterior_walls_matches_oracle(lat):
    L = lbm.get_D2Q9() if lat == "D2Q9" else lbm.get_D3Q19()
    n = (32, 16, 1) if lat == "D2Q9" else (16, 8, 16)
    g = lbm.LBM_Grid(L, *n, 8 if lat == "D3Q19" else (16, 8, 1))
    cfg = lbm.DoubleBufferConfig(BCs={lbm.Flags.EQUILIBRIUM})
    d = make_desc(g, cfg)
    f, bc, fl = O.alloc(d)
    hb, hf = _HostOnly(bc.copy()), _HostOnly(fl.copy())
    D = g.D
    plan = [("+Y", 1, [0.1] + [0.0] * (D - 1), 1.0), ("-Y", 1, [0.0] * D, 1.0), ("+X", 2, [], 1.02),
            ("-X", 2, [0.03] + [0.0] * (D - 1), 1.0)]
    if D == 3:
        plan += [("-Z", 1, [0.0] * 3, 1.0), ("+Z", 1, [0.0, 0.0, 0.01], 1.0)]
    for side, t, u, rho in plan:
        lbm.set_exterior_walls(g, cfg, hf, hb, side, t, u, rho)
        O.set_exterior_wall(d, fl, bc, side, t, u if u else None, rho)
    assert np.array_equal(hf.arr, fl)
    assert np.array_equal(hb.arr.view(np.uint32), bc.view(np.uint32))  # NaNs compared bitwise


@pytest.mark.parametrize("kw,msg", [
    (dict(side="+W", boundary_type=1, u=[0, 0, 0], rho=1.0), "Side not valid"),
    (dict(side="+X", boundary_type=1, u=[], rho=float("nan")), "Either velocity or density"),
    (dict(side="+X", boundary_type=1, u=[0, 0], rho=1.0), "length"),
    (dict(side="+X", boundary_type=1, u=[0, float("nan"), 0], rho=1.0), "free or specified"),
    (dict(side="+X", boundary_type=1, u=[0, 0, 0], rho=float("nan")), "For Solid Type"),
    (dict(side="+X", boundary_type=2, u=[0.1, 0, 0], rho=1.0), "valid boundary types"),
])
def test_set_exterior_walls_validation(kw, msg):
    g = lbm.LBM_Grid(lbm.get_D3Q19(), 16, 16, 16, 8)
    cfg = lbm.DoubleBufferConfig()
    with pytest.raises(ValueError, match=msg):
        lbm.set_exterior_walls(g, cfg, _HostOnly(np.zeros(16 ** 3, np.uint8)), _HostOnly(np.zeros(16 ** 3 * 4, np.float32)),
                               **kw)


def test_moving_wall_requires_config_toggle():
    g = lbm.LBM_Grid(lbm.get_D3Q19(), 16, 16, 16, 8)
    cfg = lbm.DoubleBufferConfig(include_moving_boundary=False)
    with pytest.raises(ValueError, match="include_moving_boundary"):
        lbm.set_exterior_walls(g, cfg, _HostOnly(np.zeros(16 ** 3, np.uint8)),
                               _HostOnly(np.zeros(16 ** 3 * 4, np.float32)), "+Y", 1, [0.1, 0, 0], 1.0)


def test_unit_system_with_tau_bounds():
    """grid.mojo:238-277: u_lat = (tau - 1/2) / 3 * Re / L_lat, rejected outside [0.001, 0.1]."""
    g = lbm.LBM_Grid(lbm.get_D3Q19(), 64, 64, 64, 8, dx=1.0 / 64)
    us = g.get_UnitSystem_with_tau(1.0, 0.56, 1.0, 1e-2)
    assert math.isclose(us.U.lattice, 0.06 / 3 * 100 / 64)
    assert math.isclose(us.tau, 0.56)
    with pytest.raises(ValueError, match="does not satisfy bounds"):
        g.get_UnitSystem_with_tau(1.0, 0.9, 1.0, 1e-2)


def test_assembly_method_mismatch():
    """assembly.mojo:100-164: asking an Assembly for the other method's buffers raises."""
    g = lbm.LBM_Grid(lbm.get_D2Q9(), 32, 32, 1, 16)
    ep = lbm.Assembly(g, lbm.EsotericPullConfig())
    with pytest.raises(RuntimeError, match="esoteric pull"):
        ep.get_gpu_tensors_for_double_buffer()
    db = lbm.Assembly(g, lbm.DoubleBufferConfig())
    with pytest.raises(RuntimeError, match="double buffer"):
        db.get_gpu_tensors_for_esoteric_pull()

"""Host-side setup mirrors on the CPU: add_sphere / add_circle / add_box in physical coordinates
(src/lbm/geometry/primatives.mojo:110-266) against a brute-force scan of every node, and
set_exterior_walls_with_func (src/lbm/preprocess/boundary_condition.mojo:186-309)."""
import numpy as np
import pytest

import mojo_lbm_amd as lbm
from mojo_lbm_amd import geometry as G


def _brute(grid, test):
    out = np.zeros(grid.shape, bool)
    for i in range(grid.shape[0]):
        for j in range(grid.shape[1]):
            for k in range(grid.shape[2]):
                p = [np.float32(n) * np.float32(grid.dx) + np.float32(o) for n, o in zip((i, j, k), grid.origin)]
                out[i, j, k] = test(p)
    return out


def _flag_nodes(grid, fl):
    idx = grid.layouts.index_map(1)[..., 0]
    return fl[idx] == lbm.Flags.SOLID


def test_sphere_and_box_physical():
    grid = lbm.LBM_Grid(lbm.get_D3Q19(), 16, 16, 8, 8, dx=0.5, origin=(-2.0, 1.0, 0.25))
    fl = np.zeros(grid.num_points, np.uint8)
    c, r = (1.3, 4.6, 2.1), 1.7
    got = G.add_sphere(grid, fl, c, r)
    f32 = np.float32
    exp = _brute(grid, lambda p: sum((p[i] - f32(c[i])) ** 2 for i in range(3)) <= f32(r) * f32(r))
    assert got.sum() > 50 and np.array_equal(got, exp) and np.array_equal(_flag_nodes(grid, fl), exp)
    fl2 = np.zeros(grid.num_points, np.uint8)
    cb, hb = (0.0, 3.0, 2.0), (1.2, 0.75, 1.0)
    got = G.add_box(grid, fl2, cb, hb)
    exp = _brute(grid, lambda p: all(f32(cb[i]) - f32(hb[i]) <= p[i] <= f32(cb[i]) + f32(hb[i]) for i in range(3)))
    assert got.sum() > 20 and np.array_equal(got, exp) and np.array_equal(_flag_nodes(grid, fl2), exp)
    with pytest.raises(ValueError):
        G.add_sphere(grid, fl, (1.0, 2.0), 1.0)
    with pytest.raises(ValueError):
        G.add_circle(grid, fl, c, r)


def test_circle_2d():
    grid = lbm.LBM_Grid(lbm.get_D2Q9(), 32, 32, 1, 16)
    fl = np.zeros(grid.num_points, np.uint8)
    got = G.add_circle(grid, fl, (10.5, 12.25, 0.0), 4.5)
    x, y = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
    exp = ((x - 10.5) ** 2 + (y - 12.25) ** 2 <= 4.5 ** 2)[..., None]
    assert np.array_equal(got, exp)


def test_walls_with_func():
    grid = lbm.LBM_Grid(lbm.get_D2Q9(), 32, 16, 1, 16, dx=0.1)
    config = lbm.LBM_Config(lbm.DOUBLE_BUFFER, BCs={lbm.Flags.EQUILIBRIUM})
    asm = lbm.Assembly(grid, config)
    units = lbm.UnitSystem(1.0, 0.05, 1.5, 15.0, 1.2, 1e-4, D=2)
    H = 1.5

    def inlet(x, y, z):
        return [4.0 * y * (H - y) / H ** 2, np.zeros_like(y)]

    lbm.set_exterior_walls_with_func(grid, config, asm.flags, asm.bc, "-X", lbm.Flags.EQUILIBRIUM, inlet, units, 1.2)
    lay = grid.layouts
    ys = np.arange(16)
    fl = asm.flags.cpu()[lay.index(np.zeros(16, int), ys, np.zeros(16, int), 0, 1)]
    assert np.all(fl == lbm.Flags.EQUILIBRIUM)
    b = asm.bc.cpu()
    y = ys.astype(np.float32) * np.float32(0.1)
    ux = b[lay.index(np.zeros(16, int), ys, np.zeros(16, int), 0, 3)]
    rho = b[lay.index(np.zeros(16, int), ys, np.zeros(16, int), 2, 3)]
    np.testing.assert_allclose(ux, 4.0 * y * (H - y) / H ** 2 * 0.05, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(rho, 1.0, rtol=1e-6)
    with pytest.raises(ValueError):
        lbm.set_exterior_walls_with_func(grid, config, asm.flags, asm.bc, "-X", lbm.Flags.SOLID, inlet, units)

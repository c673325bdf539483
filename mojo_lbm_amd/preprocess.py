"""Setup: initial populations and exterior walls.

initialize_fluid_at_rest / initialize_equilibrium fill f on the device through
lbm_init_equilibrium (src/lbm/preprocess/initial_condition.mojo:47-178);
initialize_f_from_func evaluates a velocity function (and its gradient) over the
grid and fills f on the device through lbm_init_f_from_fields, slab by slab
(initial_condition.mojo:86-255). set_exterior_walls / set_exterior_walls_with_func
write flags and bc on the host side of the buffers, like the reference's host
loops (src/lbm/preprocess/boundary_condition.mojo:78-309), vectorised with numpy;
later calls overwrite shared edges.

The velocity functions are called with numpy arrays of physical coordinates
(x, y, z), one entry per node, instead of once per node: u(x, y, z) returns D
arrays (or an array with a trailing axis of D), deriv_u(x, y, z, velocity)
returns D rows of D arrays, row a holding d u_a / d x_b.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .config import LBM_Config
from .constants import Flags
from .grid import LBM_Grid


def _error_check(config: LBM_Config, D: int, side: str, boundary_type: int, u, rho: float) -> None:
    """_error_check, boundary_condition.mojo:20-70."""
    valid = {"-X", "+X", "-Y", "+Y", "-Z", "+Z"}
    if boundary_type not in config.INCLUDED_BCs:
        raise ValueError(f"Input Boundary Type was {boundary_type} but valid boundary types are: {config.INCLUDED_BCs}")
    if side not in valid:
        raise ValueError(f"Side not valid. Input was {side} but expects {valid}")
    u_is_empty = len(u) == 0
    if u_is_empty and math.isnan(rho):
        raise ValueError("Either velocity or density or both have to be specified. Both cant be left as None")
    if not u_is_empty:
        if len(u) != D:
            raise ValueError(f"Input velocity list was of length {len(u)} but Grid is {D} Dimensional")
        if any(math.isnan(v) for v in u):
            raise ValueError("either all components of velocity must be free or specified (cannot be NaN)")
    if boundary_type == Flags.SOLID:
        if u_is_empty or math.isnan(rho):
            raise ValueError("For Solid Type you must specify both u and rho")
        if not config.include_moving_boundary and any(v != 0.0 for v in u):
            raise ValueError("include_moving_boundary in config was set to False but a non zero velocity was "
                             "specified, Please set config.include_moving_boundary to True First")
    if boundary_type == Flags.EQUILIBRIUM and not u_is_empty and math.isnan(rho):
        raise ValueError("For Equilibrium BC, if u is specified then the density must also be specified")


def set_exterior_walls(grid: LBM_Grid, config: LBM_Config, flags, bc, side: str, boundary_type: int, u=(),
                       rho: float = float("nan"), unitSystem=None) -> None:
    """Writes flag = boundary_type and bc = (u, rho) on one face of the grid (host side)."""
    D = grid.D
    u = list(u)
    _error_check(config, D, side, boundary_type, u, rho)
    velocity = [float("nan")] * D if not u else [float(v) for v in u]
    density = float(rho)
    if unitSystem is not None:
        density *= unitSystem.density.C_phys_to_lat()
        velocity = [unitSystem.U.C_phys_to_lat() * v for v in velocity]
    axis = "XYZ".index(side[1])
    if axis >= D:
        raise ValueError("Side not valid for a 2-D grid")
    fixed = 0 if side[0] == "-" else grid.shape[axis] - 1
    rng = [np.arange(n) for n in grid.shape]
    rng[axis] = np.array([fixed])
    x, y, z = np.meshgrid(*rng, indexing="ij")
    lay = grid.layouts
    fl = flags.cpu()
    b = bc.cpu()
    fl[lay.index(x, y, z, 0, 1)] = np.uint8(boundary_type)
    for i in range(D):
        b[lay.index(x, y, z, i, D + 1)] = velocity[i]
    b[lay.index(x, y, z, D, D + 1)] = density


def _handle_for(grid: LBM_Grid, config: LBM_Config):
    from .solver import make_desc

    return _lib.Handle(make_desc(grid, config))


def initialize_equilibrium(grid: LBM_Grid, config: LBM_Config, f, rho: float = 1.0, u=None, solver=None) -> None:
    """f = f_eq(rho, u) at every node, stored with the method's pattern, on the device."""
    import torch

    h = solver.handle if solver is not None else _handle_for(grid, config)
    ua = _lib.dvec(u)
    _lib.check(_lib.lib().lbm_init_equilibrium(h.ptr, f.gpu().data_ptr(), float(rho), ua,
                                               C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    f.mark_device_written()


def initialize_fluid_at_rest(grid: LBM_Grid, config: LBM_Config, f, solver=None) -> None:
    """initial_condition.mojo:47-83: f_eq(rho = 1, u = 0)."""
    initialize_equilibrium(grid, config, f, 1.0, None, solver)


def _f32(v, dtype):
    return np.asarray(v, dtype=dtype)


def _grid_coords(grid: LBM_Grid, xs, ys, zs):
    """get_grid_coordinates (src/lbm/grid.mojo:148-168): index * dx + origin in the grid float type."""
    dt = grid.float_dtype
    dx = dt(grid.dx)
    return tuple(_f32(i, dt) * dx + dt(o) for i, o in zip((xs, ys, zs), grid.origin))


def _components(values, D, shape, dtype):
    """u / deriv_u results as an array [..., D] (a sequence of D arrays or an array with a trailing D)."""
    if isinstance(values, np.ndarray) and values.shape == tuple(shape) + (D,):
        return values.astype(dtype, copy=False)
    comps = [np.broadcast_to(np.asarray(v, dtype=dtype), shape) for v in values]
    if len(comps) != D:
        raise ValueError(f"the velocity function returned {len(comps)} components for a {D}-D grid")
    return np.stack(comps, axis=-1)


def initialize_f_from_func(grid: LBM_Grid, config: LBM_Config, f, u, rho: float = 1.0, tau: float = 1.0,
                           unitSystem=None, deriv_u=None, solver=None, slab_nodes: int = 1 << 24) -> None:
    """initial_condition.mojo:86-178 on the device: f = f_eq(rho, u(x)) (+ fi_neq of deriv_u, :181-255)
    with u and deriv_u evaluated at the nodes' physical coordinates; velocities are scaled by
    U.C_phys_to_lat() and gradients by U.C_phys_to_lat() / L.C_phys_to_lat() when a unit system
    is given. The grid is filled in x-slabs of about slab_nodes nodes."""
    import torch

    h = solver.handle if solver is not None else _handle_for(grid, config)
    dt = grid.float_dtype
    D = grid.D
    ny, nz = grid.shape[1], grid.shape[2]
    u_scale = g_scale = None
    if unitSystem is not None:
        u_scale = dt(dt(unitSystem.U.lattice) / dt(unitSystem.U.physical))
        g_scale = dt(u_scale / dt(dt(unitSystem.L.lattice) / dt(unitSystem.L.physical)))
    fd = f.gpu()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    step = max(1, slab_nodes // (ny * nz))
    for x0 in range(0, grid.shape[0], step):
        nxc = min(step, grid.shape[0] - x0)
        xs, ys, zs = np.meshgrid(np.arange(x0, x0 + nxc), np.arange(ny), np.arange(nz), indexing="ij")
        X, Y, Z = _grid_coords(grid, xs, ys, zs)
        vel = _components(u(X, Y, Z), D, xs.shape, dt)
        if u_scale is not None:
            vel = vel * u_scale
        grad = None
        if deriv_u is not None:
            rows = deriv_u(X, Y, Z, vel.copy())
            if len(rows) != D:
                raise ValueError(f"deriv_u returned {len(rows)} rows for a {D}-D grid")
            grad = np.stack([_components(r, D, xs.shape, dt) for r in rows], axis=-2)  # [..., a, b]
            if g_scale is not None:
                grad = grad * g_scale
        ud = torch.from_numpy(np.ascontiguousarray(vel)).cuda()
        gd = None if grad is None else torch.from_numpy(np.ascontiguousarray(grad)).cuda()
        _lib.check(_lib.lib().lbm_init_f_from_fields(h.ptr, fd.data_ptr(), float(rho), ud.data_ptr(),
                                                      None if gd is None else gd.data_ptr(), float(tau),
                                                      int(x0), int(nxc), st))
        torch.cuda.current_stream().synchronize()  # the slab's host arrays may go now
    f.mark_device_written()


def set_exterior_walls_with_func(grid: LBM_Grid, config: LBM_Config, flags, bc, side: str, boundary_type: int, u,
                                 unitSystem=None, rho: float = float("nan")) -> None:
    """boundary_condition.mojo:186-309: flag = boundary_type on one face and bc = (u(x) * U.C_phys_to_lat(),
    rho * density.C_phys_to_lat()) with u evaluated at the face nodes' physical coordinates."""
    D = grid.D
    _error_check(config, D, side, boundary_type, [1.0] * D, rho)
    dt = grid.float_dtype
    density = dt(rho)
    factor = dt(1.0)
    if unitSystem is not None:
        density = dt(density * dt(dt(unitSystem.density.lattice) / dt(unitSystem.density.physical)))
        factor = dt(dt(unitSystem.U.lattice) / dt(unitSystem.U.physical))
    axis = "XYZ".index(side[1])
    if axis >= D:
        raise ValueError("Side not valid for a 2-D grid")
    fixed = 0 if side[0] == "-" else grid.shape[axis] - 1
    rng = [np.arange(n) for n in grid.shape]
    rng[axis] = np.array([fixed])
    xs, ys, zs = np.meshgrid(*rng, indexing="ij")
    vel = _components(u(*_grid_coords(grid, xs, ys, zs)), D, xs.shape, dt)
    lay = grid.layouts
    fl = flags.cpu()
    b = bc.cpu()
    fl[lay.index(xs, ys, zs, 0, 1)] = np.uint8(boundary_type)
    for i in range(D):
        b[lay.index(xs, ys, zs, i, D + 1)] = vel[..., i] * factor
    b[lay.index(xs, ys, zs, D, D + 1)] = density

"""Setup: initial populations and exterior walls.

initialize_fluid_at_rest / initialize_equilibrium fill f on the device through
lbm_init_equilibrium (src/lbm/preprocess/initial_condition.mojo:47-178).
set_exterior_walls writes flags and bc on the host side of the buffers, like the
reference's host loop (src/lbm/preprocess/boundary_condition.mojo:78-183),
vectorised with numpy; later calls overwrite shared edges.
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

"""Solver: the host launcher of src/lbm/solver.mojo:35-187 over liblbm_hip.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .config import LBM_Config
from .constants import Collisions, LBM_method
from .grid import LBM_Grid


def make_desc(grid: LBM_Grid, config: LBM_Config) -> _lib.LbmDesc:
    """lbm_desc for a grid + config pair."""
    tx, ty, tz = grid.tile_shape
    return _lib.LbmDesc(
        D=grid.D, Q=grid.Q, nx=grid.nx, ny=grid.ny, nz=grid.nz, tx=tx, ty=ty, tz=tz,
        method=config.lbm_method, collision=Collisions.ids[config.collision_op], storage=config.storage_id,
        ddf_shift=int(config.DDF_shift), les=config.les_id, include_moving_boundary=int(config.include_moving_boundary),
        included_bcs=config.included_bcs_mask, smagorinsky_cs=config.Cs,
    )


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> int:
    return t.data_ptr() if hasattr(t, "data_ptr") else int(t)


class Solver:
    """Solver[grid, config](ctx): even_step / odd_step over an Assembly, or the raw
    double_buffer_step / esoteric_pull_step over device buffers. Launches are
    asynchronous on torch's current stream."""

    def __init__(self, grid: LBM_Grid, config: LBM_Config):
        self.grid = grid
        self.config = config
        self.lbm_method = config.lbm_method
        self.desc = make_desc(grid, config)
        self.handle = _lib.Handle(self.desc)
        self.BLOCK_SHAPE = grid.BLOCK_SHAPE
        self.GRID_DIM = grid.GRID_DIM
        self.L = _lib.lib()

    @property
    def compute_dtype(self):
        return np.float64 if self.desc.storage == 2 else np.float32

    def double_buffer_step(self, f_out, f_in, bc, flags, tau: float) -> None:
        """solver.mojo:81-111."""
        _lib.check(self.L.lbm_step_db(self.handle.ptr, _ptr(f_out), _ptr(f_in), _ptr(bc), _ptr(flags), float(tau),
                                      _stream()))

    def esoteric_pull_step(self, f, bc, flags, tau: float, is_even_step: bool) -> None:
        """solver.mojo:115-154."""
        _lib.check(self.L.lbm_step_ep(self.handle.ptr, _ptr(f), _ptr(bc), _ptr(flags), float(tau),
                                      int(bool(is_even_step)), _stream()))

    def even_step(self, assembly, tau: float) -> None:
        """solver.mojo:159-172: double buffer f -> f2; esoteric pull even pattern."""
        if self.lbm_method == LBM_method.DOUBLE_BUFFER:
            f_in, f_out, bc, flags = assembly.get_gpu_tensors_for_double_buffer()
            self.double_buffer_step(f_out, f_in, bc, flags, tau)
            assembly.f2.mark_device_written()
        else:
            f, bc, flags = assembly.get_gpu_tensors_for_esoteric_pull()
            self.esoteric_pull_step(f, bc, flags, tau, True)
            assembly.f.mark_device_written()

    def odd_step(self, assembly, tau: float) -> None:
        """solver.mojo:174-187: double buffer f2 -> f; esoteric pull odd pattern."""
        if self.lbm_method == LBM_method.DOUBLE_BUFFER:
            f_in, f_out, bc, flags = assembly.get_gpu_tensors_for_double_buffer()
            self.double_buffer_step(f_in, f_out, bc, flags, tau)
            assembly.f.mark_device_written()
        else:
            f, bc, flags = assembly.get_gpu_tensors_for_esoteric_pull()
            self.esoteric_pull_step(f, bc, flags, tau, False)
            assembly.f.mark_device_written()

    def set_links(self, nodes, qdist, q_clamp: float = 0.0) -> None:
        """Interpolated bounce-back links (geometry.sphere_links records) with each link fraction raised
        to q_clamp, the argument of nodewise_bounceback_kernel (src/lbm/geometry/nodewise_interpolated_BB.mojo:69,151)."""
        nodes = np.ascontiguousarray(nodes, np.uint32)
        qdist = np.ascontiguousarray(qdist, np.float32)
        _lib.check(self.L.lbm_set_links_clamped(self.handle.ptr, nodes.ctypes.data, qdist.ctypes.data, len(nodes),
                                                float(q_clamp)))

    def link_force(self, out) -> None:
        """Momentum-exchange force of the last step into out (D device doubles): sum_force,
        src/lbm/geometry/rigidstationary.mojo:155-169."""
        _lib.check(self.L.lbm_link_force(self.handle.ptr, _ptr(out), _stream()))

    def synchronize(self) -> None:
        _lib.check(self.L.lbm_synchronize(_stream()))

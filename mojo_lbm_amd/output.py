"""Output: rho / u fields (src/lbm/output/velocity.mojo:25-125, OutputRequest.velocity_frame
src/lbm/outputrequest.mojo:162-188) and FP64 mass / momentum totals."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .constants import LBM_method


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def calculate_rho_and_velocity(solver, assembly, after_odd_step: bool = True, u_layout: str = "xyzd"):
    """Returns (rho[nx,ny,nz], u[nx,ny,nz,D] or u[D,nx,ny,nz]) as numpy arrays.

    For double buffer the current populations are in assembly.f after odd steps
    and in assembly.f2 after even steps; after_odd_step picks that buffer (and,
    for esoteric pull, the load pattern of the next step)."""
    import torch

    g = solver.grid
    n = g.num_points
    dt = torch.float64 if solver.desc.storage == 2 else torch.float32
    rho = torch.empty(n, dtype=dt, device="cuda")
    u = torch.empty(n * g.D, dtype=dt, device="cuda")
    if solver.lbm_method == LBM_method.DOUBLE_BUFFER:
        f = assembly.f.gpu() if after_odd_step else assembly.f2.gpu()
    else:
        f = assembly.f.gpu()
    lay = _lib.U_XYZD if u_layout == "xyzd" else _lib.U_DXYZ
    _lib.check(_lib.lib().lbm_macroscopic(solver.handle.ptr, rho.data_ptr(), u.data_ptr(), f.data_ptr(),
                                          assembly.bc.gpu().data_ptr(), assembly.flags.gpu().data_ptr(),
                                          int(after_odd_step), lay, _stream()))
    torch.cuda.synchronize()
    rho = rho.cpu().numpy().reshape(g.shape)
    u = u.cpu().numpy()
    u = u.reshape(*g.shape, g.D) if u_layout == "xyzd" else u.reshape(g.D, *g.shape)
    return rho, u


def reduce_moments(solver, assembly, after_odd_step: bool = True) -> np.ndarray:
    """FP64 [sum rho, sum rho u_d] over valid nodes."""
    import torch

    out = torch.zeros(solver.grid.D + 1, dtype=torch.float64, device="cuda")
    if solver.lbm_method == LBM_method.DOUBLE_BUFFER:
        f = assembly.f.gpu() if after_odd_step else assembly.f2.gpu()
    else:
        f = assembly.f.gpu()
    _lib.check(_lib.lib().lbm_reduce_moments_f64(solver.handle.ptr, out.data_ptr(), f.data_ptr(),
                                                 assembly.bc.gpu().data_ptr(), assembly.flags.gpu().data_ptr(),
                                                 int(after_odd_step), _stream()))
    torch.cuda.synchronize()
    return out.cpu().numpy()

"""Output: rho / u fields (src/lbm/output/velocity.mojo:25-125), the Q-criterion
(src/lbm/output/Q_criterion.mojo:90-215), OutputRequest (src/lbm/outputrequest.mojo:22-287)
and FP64 mass / momentum totals. Every field is computed by liblbm_hip.so on the GPU."""
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


def check_health(solver, assembly, after_odd_step: bool = True, mass0: float | None = None,
                 max_mass_drift: float = 1e-3) -> np.ndarray:
    """Divergence watchdog (the reference has none): the FP64 totals of reduce_moments must be
    finite and, when mass0 is given, the total mass within max_mass_drift (relative) of it.
    Returns the totals; raises FloatingPointError otherwise."""
    tot = reduce_moments(solver, assembly, after_odd_step)
    if not np.all(np.isfinite(tot)):
        raise FloatingPointError(f"the run diverged: mass / momentum totals {tot}")
    if mass0 is not None and abs(tot[0] - mass0) > max_mass_drift * abs(mass0):
        raise FloatingPointError(f"total mass drifted from {mass0} to {tot[0]} (more than {max_mass_drift:g})")
    return tot


def _current_f(solver, assembly, after_odd_step: bool):
    if solver.lbm_method == LBM_method.DOUBLE_BUFFER:
        return assembly.f.gpu() if after_odd_step else assembly.f2.gpu()
    return assembly.f.gpu()


def calculate_Q_criterion(solver, assembly, velocity, tau: float, after_odd_step: bool = True,
                          u_layout: str = "xyzd"):
    """Q[nx, ny, nz] (numpy) of the current state; `velocity` is the field calculate_rho_and_velocity
    returned for the same state and u_layout (a numpy array or a device tensor)."""
    import torch

    g = solver.grid
    dt = torch.float64 if solver.desc.storage == 2 else torch.float32
    u = velocity if isinstance(velocity, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(velocity))
    u = u.to(device="cuda", dtype=dt).contiguous()
    if u.numel() != g.num_points * g.D:
        raise ValueError(f"velocity holds {u.numel()} values, the grid needs {g.num_points * g.D}")
    q = torch.empty(g.num_points, dtype=dt, device="cuda")
    lay = _lib.U_XYZD if u_layout == "xyzd" else _lib.U_DXYZ
    f = _current_f(solver, assembly, after_odd_step)
    _lib.check(_lib.lib().lbm_q_criterion(solver.handle.ptr, q.data_ptr(), u.data_ptr(), f.data_ptr(),
                                          assembly.bc.gpu().data_ptr(), assembly.flags.gpu().data_ptr(),
                                          float(tau), int(after_odd_step), lay, _stream()))
    torch.cuda.synchronize()
    return q.cpu().numpy().reshape(g.shape)


class OutputRequest:
    """OutputRequest[grid, config] (src/lbm/outputrequest.mojo:22-287): device buffers for the
    density, velocity and Q-criterion fields, filled from an Assembly by velocity_frame and
    Q_criterion_frame and returned as numpy arrays shaped (nx, ny, nz[, D]).

    The device buffers are row-major (x, y, z[, d]); the arrays returned are the same
    logical fields the reference returns from its column-major buffers reshaped in
    Fortran order. Q_criterion_frame takes tau explicitly: the reference's
    get_Q_criterion launch (outputrequest.mojo:136-158) passes no tau to a kernel that
    needs one."""

    def __init__(self, solver, unit_system=None, *, velocity: bool = True, density: bool = True,
                 Q_criterion: bool = False):
        import torch

        if Q_criterion and not velocity:
            raise ValueError("Q criterion, the velocity must be on")
        self.solver = solver
        self.grid = solver.grid
        self.unitSystem = unit_system
        dt = torch.float64 if solver.desc.storage == 2 else torch.float32
        n = self.grid.num_points
        # velocity_frame writes density and velocity together, so both buffers exist when either is asked for
        self.density = torch.zeros(n, dtype=dt, device="cuda") if (density or velocity) else None
        self.velocity = torch.zeros(n * self.grid.D, dtype=dt, device="cuda") if (density or velocity) else None
        self.Q_criterion = torch.zeros(n, dtype=dt, device="cuda") if Q_criterion else None
        self._want = {"density": density, "velocity": velocity}

    def velocity_frame(self, assembly, after_odd_step: bool) -> None:
        """outputrequest.mojo:162-188: density and velocity of the current state (asynchronous)."""
        if self.velocity is None:
            raise ValueError("this OutputRequest holds no velocity or density buffer")
        f = _current_f(self.solver, assembly, after_odd_step)
        _lib.check(_lib.lib().lbm_macroscopic(self.solver.handle.ptr, self.density.data_ptr(),
                                              self.velocity.data_ptr(), f.data_ptr(), assembly.bc.gpu().data_ptr(),
                                              assembly.flags.gpu().data_ptr(), int(after_odd_step), _lib.U_XYZD,
                                              _stream()))

    def Q_criterion_frame(self, assembly, after_odd_step: bool, tau: float) -> None:
        """outputrequest.mojo:192-223: Q of the current state from the velocity buffer, which
        velocity_frame must have filled for the same state (asynchronous)."""
        if self.Q_criterion is None:
            raise ValueError("this OutputRequest was made without Q_criterion")
        f = _current_f(self.solver, assembly, after_odd_step)
        _lib.check(_lib.lib().lbm_q_criterion(self.solver.handle.ptr, self.Q_criterion.data_ptr(),
                                              self.velocity.data_ptr(), f.data_ptr(), assembly.bc.gpu().data_ptr(),
                                              assembly.flags.gpu().data_ptr(), float(tau), int(after_odd_step),
                                              _lib.U_XYZD, _stream()))

    def _host(self, t, shape, factor, squeeze):
        import torch

        torch.cuda.synchronize()
        a = t.cpu().numpy().reshape(shape)
        if factor is not None:
            a = a * factor
        return a.squeeze() if squeeze else a

    def _factor(self, convert, kind):
        if not convert:
            return None
        if self.unitSystem is None:
            raise ValueError("converting to physical units needs the OutputRequest's unit system")
        U = self.unitSystem.U.C_lat_to_phys()
        # the reference scales Q by U / L (outputrequest.mojo:281-282)
        return U if kind == "u" else U / self.unitSystem.L.C_lat_to_phys()

    def velocity_as_numpy(self, convert_from_lattice_units: bool, squeeze: bool = True) -> np.ndarray:
        """outputrequest.mojo:226-254: velocity shaped (nx, ny, nz, D)."""
        if not self._want["velocity"]:
            raise ValueError("this OutputRequest was made without velocity")
        return self._host(self.velocity, (*self.grid.shape, self.grid.D),
                          self._factor(convert_from_lattice_units, "u"), squeeze)

    def density_as_numpy(self, squeeze: bool = True) -> np.ndarray:
        """Density shaped (nx, ny, nz) in lattice units."""
        if not self._want["density"]:
            raise ValueError("this OutputRequest was made without density")
        return self._host(self.density, tuple(self.grid.shape), None, squeeze)

    def Q_criterion_as_numpy(self, convert_from_lattice_units: bool, squeeze: bool = True) -> np.ndarray:
        """outputrequest.mojo:257-287: Q shaped (nx, ny, nz)."""
        if self.Q_criterion is None:
            raise ValueError("this OutputRequest was made without Q_criterion")
        return self._host(self.Q_criterion, tuple(self.grid.shape), self._factor(convert_from_lattice_units, "q"),
                          squeeze)

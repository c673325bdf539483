"""Shared helpers for the parity tests: run descriptions, seeded inputs, and a
small driver that runs the HIP library through its C ABI on torch-allocated
device buffers."""
from __future__ import annotations

import ctypes as C

import numpy as np

from mojo_lbm_amd import _lib
from mojo_lbm_amd._lib import LbmDesc
from oracle import oracle as O

LATTICES = {"D2Q9": (2, 9), "D3Q19": (3, 19), "D3Q27": (3, 27)}


def make_desc(lat="D3Q19", n=(16, 16, 16), tile=None, method=0, collision=0, storage=0, ddf=0, les=0,
              moving=1, eq=False, cs=0.1) -> LbmDesc:
    D, Q = LATTICES[lat]
    if D == 2:
        n = (n[0], n[1], 1)
        tile = tile or (16, 16, 1)
    else:
        tile = tile or (8, 8, 8)
    bcs = (1 << 0) | (1 << 1) | ((1 << 2) if eq else 0)
    return LbmDesc(D=D, Q=Q, nx=n[0], ny=n[1], nz=n[2], tx=tile[0], ty=tile[1], tz=tile[2], method=method,
                   collision=collision, storage=storage, ddf_shift=ddf, les=les, include_moving_boundary=moving,
                   included_bcs=bcs, smagorinsky_cs=cs)


def ldc_walls(desc, flags, bc, u_lid=0.1):
    """Lid-driven cavity walls in the reference scripts' order: +Y lid, then the other faces
    (examples/LDC.mojo:92-95, benchmarks/LDC/LDC_3D_Timed.mojo:88-93)."""
    D = desc.D
    lid = [u_lid] + [0.0] * (D - 1)
    zero = [0.0] * D
    O.set_exterior_wall(desc, flags, bc, "+Y", 1, lid, 1.0)
    sides = ["-Y", "+X", "-X"] + (["-Z", "+Z"] if D == 3 else [])
    for s in sides:
        O.set_exterior_wall(desc, flags, bc, s, 1, zero, 1.0)


def perturbed_state(desc, seed=0, amp=1e-3, walls=True, u_lid=0.1):
    """Equilibrium at rest plus seeded noise on every population (so every population
    carries a distinct value), with LDC walls."""
    f, bc, flags = O.alloc(desc)
    rng = np.random.default_rng(seed)
    if desc.storage == 1:
        ff = np.zeros(f.size, np.float32)
        d32 = make_desc_like(desc, storage=0)
        O.init_equilibrium(d32, ff)
        ff += (amp * rng.standard_normal(ff.size)).astype(np.float32)
        f[:] = O.f16c_encode(ff)
    else:
        O.init_equilibrium(desc, f)
        f += (amp * rng.standard_normal(f.size)).astype(f.dtype)
    if walls:
        ldc_walls(desc, flags, bc, u_lid)
    return f, bc, flags


def make_desc_like(desc, **kw) -> LbmDesc:
    d = LbmDesc()
    C.memmove(C.byref(d), C.byref(desc), C.sizeof(LbmDesc))
    for k, v in kw.items():
        setattr(d, k, v)
    return d


class DeviceRun:
    """Device buffers + handle for one description, driven through the C ABI."""

    def __init__(self, desc, f, bc, flags):
        import torch

        self.torch = torch
        self.desc = desc
        self.h = _lib.Handle(desc)
        self.f = torch.from_numpy(f.copy()).cuda()
        self.f2 = torch.from_numpy(f.copy()).cuda()
        self.bc = torch.from_numpy(bc.copy()).cuda()
        self.flags = torch.from_numpy(flags.copy()).cuda()
        self.L = _lib.lib()

    @property
    def stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def step_db(self, tau):
        _lib.check(self.L.lbm_step_db(self.h.ptr, self.f2.data_ptr(), self.f.data_ptr(), self.bc.data_ptr(),
                                      self.flags.data_ptr(), float(tau), self.stream))
        self.f, self.f2 = self.f2, self.f

    def step_ep(self, tau, is_even):
        _lib.check(self.L.lbm_step_ep(self.h.ptr, self.f.data_ptr(), self.bc.data_ptr(), self.flags.data_ptr(),
                                      float(tau), int(is_even), self.stream))

    def macroscopic(self, after_odd=False, u_layout=0, misalign=False):
        """rho and u on the host; misalign: output buffers one element past a 16-byte boundary."""
        t = self.torch
        dt = t.float64 if self.desc.storage == 2 else t.float32
        n = self.desc.nx * self.desc.ny * self.desc.nz
        k = 1 if misalign else 0
        rho = t.empty(n + k, dtype=dt, device="cuda")[k:]
        u = t.empty(n * self.desc.D + k, dtype=dt, device="cuda")[k:]
        _lib.check(self.L.lbm_macroscopic(self.h.ptr, rho.data_ptr(), u.data_ptr(), self.f.data_ptr(),
                                          self.bc.data_ptr(), self.flags.data_ptr(), int(after_odd), int(u_layout),
                                          self.stream))
        t.cuda.synchronize()
        return rho.cpu().numpy(), u.cpu().numpy()

    def q_criterion(self, u, tau, after_odd=False, u_layout=0, misalign=False):
        """Q-criterion of the current state from the velocity field u (host array in u_layout);
        misalign: the output one element past a 16-byte boundary."""
        t = self.torch
        dt = t.float64 if self.desc.storage == 2 else t.float32
        n = self.desc.nx * self.desc.ny * self.desc.nz
        ud = t.from_numpy(np.ascontiguousarray(u)).to(device="cuda", dtype=dt)
        k = 1 if misalign else 0
        q = t.full((n + k,), float("nan"), dtype=dt, device="cuda")[k:]
        _lib.check(self.L.lbm_q_criterion(self.h.ptr, q.data_ptr(), ud.data_ptr(), self.f.data_ptr(),
                                          self.bc.data_ptr(), self.flags.data_ptr(), float(tau), int(after_odd),
                                          int(u_layout), self.stream))
        t.cuda.synchronize()
        return q.cpu().numpy()

    def reduce(self, after_odd=False):
        t = self.torch
        out = t.zeros(self.desc.D + 1, dtype=t.float64, device="cuda")
        _lib.check(self.L.lbm_reduce_moments_f64(self.h.ptr, out.data_ptr(), self.f.data_ptr(), self.bc.data_ptr(),
                                                 self.flags.data_ptr(), int(after_odd), self.stream))
        t.cuda.synchronize()
        return out.cpu().numpy()

    def set_links(self, nodes, qdist, q_clamp=None):
        nodes = np.ascontiguousarray(nodes, np.uint32)
        qdist = np.ascontiguousarray(qdist, np.float32)
        if q_clamp is None:
            _lib.check(self.L.lbm_set_links(self.h.ptr, nodes.ctypes.data, qdist.ctypes.data, len(nodes)))
        else:
            _lib.check(self.L.lbm_set_links_clamped(self.h.ptr, nodes.ctypes.data, qdist.ctypes.data, len(nodes),
                                                    float(q_clamp)))

    def link_force(self):
        t = self.torch
        out = t.zeros(3, dtype=t.float64, device="cuda")
        _lib.check(self.L.lbm_link_force(self.h.ptr, out.data_ptr(), self.stream))
        t.cuda.synchronize()
        return out.cpu().numpy()[: self.desc.D]

    def f_host(self):
        self.torch.cuda.synchronize()
        return self.f.cpu().numpy()


def oracle_run_db(desc, f, bc, flags, tau, steps):
    a, b = f.copy(), f.copy()
    for _ in range(steps):
        O.step_db(desc, b, a, bc, flags, tau)
        a, b = b, a
    return a


def oracle_run_ep(desc, f, bc, flags, tau, steps, start_even=True):
    a = f.copy()
    even = start_even
    for _ in range(steps):
        O.step_ep(desc, a, bc, flags, tau, even)
        even = not even
    return a


def max_abs_diff(desc, a, b):
    if desc.storage == 1:
        a, b = O.f16c_decode(a), O.f16c_decode(b)
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))))


def index_map(desc, K):
    """addr[x, y, z, k] of the tiled layout, built with numpy reshapes (a col-major tile (tx,ty,tz,K)
    inside a col-major tiler (Tx,Ty,Tz))."""
    nx, ny, nz = desc.nx, desc.ny, desc.nz
    tx, ty, tz = desc.tx, desc.ty, desc.tz
    a = np.arange(nx * ny * nz * K).reshape(nz // tz, ny // ty, nx // tx, K, tz, ty, tx)
    return a.transpose(2, 6, 1, 5, 0, 4, 3).reshape(nx, ny, nz, K)


def db_to_ep_odd(desc, f):
    """Stores a double-buffer population array with the esoteric odd store pattern
    (the pattern initialize_f_from_func uses for esoteric pull): f_p -> slot p at x + c_p,
    f_n -> slot n at x, rest -> slot 0 at x."""
    c, _, _ = O.lattice(desc.D, desc.Q)
    idx = index_map(desc, desc.Q)
    out = f.copy()
    for p in range(1, desc.Q, 2):
        shifted = np.roll(idx[..., p], shift=(-c[p][0], -c[p][1], -c[p][2]), axis=(0, 1, 2))
        out[shifted] = f[idx[..., p]]
    return out


def grid_of(desc):
    """Host LBM_Grid with the description's lattice, sizes and tile."""
    import mojo_lbm_amd as lbm

    lat = {9: lbm.get_D2Q9, 19: lbm.get_D3Q19, 27: lbm.get_D3Q27}[desc.Q]()
    return lbm.LBM_Grid(lat, desc.nx, desc.ny, desc.nz, (desc.tx, desc.ty, desc.tz))


def obstacle_case(lat="D2Q9", storage=0, ddf=0, seed=0, amp=1e-3, u_in=0.05, collision=0, les=0):
    """Channel with an immersed circle (2-D) or sphere (3-D) in the reference cylinder script's
    boundary order (benchmarks/Cylinder2D/CylinderFlow_Re20.mojo:147-150): +X EQUILIBRIUM outlet,
    -X SOLID inlet with velocity, -Y/+Y SOLID walls. Seeded noise on every population.
    Returns desc, f, bc, flags, nodes, qdist."""
    from mojo_lbm_amd import geometry as G

    if lat == "D2Q9":
        n, center, radius = (64, 32, 1), (16.3, 15.6, 0.0), 5.3
    else:
        n, center, radius = (32, 16, 16), (10.4, 7.7, 8.1), 4.2
    desc = make_desc(lat, n, storage=storage, ddf=ddf, eq=True, collision=collision, les=les)
    f, bc, flags = perturbed_state(desc, seed=seed, amp=amp, walls=False)
    D = desc.D
    O.set_exterior_wall(desc, flags, bc, "+X", 2, None, 1.0)
    O.set_exterior_wall(desc, flags, bc, "-X", 1, [u_in] + [0.0] * (D - 1), 1.0)
    O.set_exterior_wall(desc, flags, bc, "-Y", 1, [0.0] * D, 1.0)
    O.set_exterior_wall(desc, flags, bc, "+Y", 1, [0.0] * D, 1.0)
    grid = grid_of(desc)
    inside = G.add_sphere(grid, flags, center, radius)
    nodes, qdist = G.sphere_links(grid, flags, center, radius, inside)
    return desc, f, bc, flags, nodes, qdist

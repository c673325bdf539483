"""Batched independent cases (the lbm_batch_* entry points of include/lbm_hip.h).

Many double-buffer grids of different sizes and relaxation times advance together:
one kernel launch per step for every case, runs of steps replayed from a captured
hipGraph, per-case momentum-exchange forces. The reference runs such cases one
script at a time (benchmarks/Cylinder2D/CylinderFlow_Re20.mojo:120-228); there is
no batched counterpart.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .config import LBM_Config
from .constants import Collisions
from .lattice import Lattice


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class CaseBatch:
    """Owns one lbm_batch. sizes: [(nx, ny, nz)] per case; taus: relaxation time per case.
    Case arrays are row-major (x, y, z) host arrays (x slowest)."""

    def __init__(self, lattice: Lattice, sizes, taus, config: LBM_Config, tile=None):
        self.lattice = lattice
        self.D, self.Q = lattice.D, lattice.Q
        self.sizes = np.ascontiguousarray(np.asarray(sizes, np.int32).reshape(-1, 3))
        self.taus = np.ascontiguousarray(np.asarray(taus, np.float64).reshape(-1))
        if len(self.sizes) != len(self.taus):
            raise ValueError("one tau per case")
        tile = tile or ((16, 16, 1) if self.D == 2 else (8, 8, 8))
        self.proto = _lib.LbmDesc(
            D=self.D, Q=self.Q, nx=0, ny=0, nz=0, tx=tile[0], ty=tile[1], tz=tile[2], method=config.lbm_method,
            collision=Collisions.ids[config.collision_op], storage=config.storage_id, ddf_shift=int(config.DDF_shift),
            les=config.les_id, include_moving_boundary=int(config.include_moving_boundary),
            included_bcs=config.included_bcs_mask, smagorinsky_cs=config.Cs)
        self.L = _lib.lib()
        ptr = C.c_void_p()
        _lib.check(self.L.lbm_batch_create(C.byref(self.proto), len(self.taus), self.sizes.ctypes.data,
                                           self.taus.ctypes.data, C.byref(ptr)))
        self.ptr = ptr
        self.n_cases = len(self.taus)

    # -- setup
    def set_case(self, c: int, flags=None, bc=None, links=None, q_clamp: float = 0.0) -> None:
        """flags [nx, ny, nz] uint8, bc [nx, ny, nz, D+1] float64, links = (row-major node ids, qdist),
        each link fraction raised to q_clamp as lbm_set_links_clamped does."""
        if not 0 <= c < self.n_cases:
            raise IndexError(f"case {c} outside 0..{self.n_cases - 1}")
        n = int(np.prod(self.sizes[c]))
        fl = None if flags is None else np.ascontiguousarray(flags, np.uint8).reshape(-1)
        b = None if bc is None else np.ascontiguousarray(bc, np.float64).reshape(-1)
        if fl is not None and fl.size != n:
            raise ValueError("flags size does not match the case")
        if b is not None and b.size != n * (self.D + 1):
            raise ValueError("bc size does not match the case")
        nodes = qd = None
        nl = 0
        if links is not None:
            nodes = np.ascontiguousarray(links[0], np.uint32)
            qd = np.array(links[1], np.float32)
            if not 0.0 <= q_clamp <= 1.0:
                raise ValueError("q_clamp must lie in [0, 1]")
            qd[qd >= 0] = np.maximum(qd[qd >= 0], np.float32(q_clamp))
            nl = len(nodes)
        _lib.check(self.L.lbm_batch_set_case(self.ptr, int(c), None if fl is None else fl.ctypes.data,
                                             None if b is None else b.ctypes.data,
                                             None if nodes is None else nodes.ctypes.data,
                                             None if qd is None else qd.ctypes.data, nl))

    def set_tile(self, tile) -> None:
        _lib.check(self.L.lbm_batch_set_tile(self.ptr, *[int(t) for t in tile]))

    def set_option(self, grid_blocks: int | None = None, graph_steps: int | None = None) -> None:
        if grid_blocks is not None:
            _lib.check(self.L.lbm_batch_set_option(self.ptr, _lib.BATCH_OPT_GRID_BLOCKS, int(grid_blocks)))
        if graph_steps is not None:
            _lib.check(self.L.lbm_batch_set_option(self.ptr, _lib.BATCH_OPT_GRAPH_STEPS, int(graph_steps)))

    def autotune(self, steps: int = 200):
        """Picks the fastest tile shape and grid; returns (tile, grid_blocks). Resets the state to rest."""
        best = (C.c_int32 * 4)()
        _lib.check(self.L.lbm_batch_autotune(self.ptr, int(steps), _stream(), best))
        return (best[0], best[1], best[2]), best[3]

    # -- state
    def init_equilibrium(self) -> None:
        _lib.check(self.L.lbm_batch_init_equilibrium(self.ptr, _stream()))

    def run(self, steps: int) -> None:
        _lib.check(self.L.lbm_batch_run(self.ptr, int(steps), _stream()))

    def force(self) -> np.ndarray:
        """Momentum-exchange force of the last step, [n_cases, D] float64."""
        import torch

        out = torch.zeros(self.n_cases * self.D, dtype=torch.float64, device="cuda")
        _lib.check(self.L.lbm_batch_force(self.ptr, out.data_ptr(), _stream()))
        return out.cpu().numpy().reshape(self.n_cases, self.D)

    def case_desc(self, c: int) -> _lib.LbmDesc:
        d = _lib.LbmDesc()
        _lib.check(self.L.lbm_batch_case_desc(self.ptr, int(c), C.byref(d)))
        return d

    def export_case(self, c: int):
        """Padded tiled (flags, bc, link nodes, qdist) of case c, for a single-grid rerun."""
        d = self.case_desc(c)
        cells = d.nx * d.ny * d.nz
        flags = np.zeros(cells, np.uint8)
        bc = np.zeros(cells * (self.D + 1), np.float64 if d.storage == _lib.STORE_F64 else np.float32)
        n = C.c_uint32(0)
        _lib.check(self.L.lbm_batch_export_case(self.ptr, int(c), None, None, None, None, C.byref(n)))
        nodes = np.zeros(n.value, np.uint32)
        qd = np.zeros(n.value * (self.Q - 1), np.float32)
        _lib.check(self.L.lbm_batch_export_case(self.ptr, int(c), flags.ctypes.data, bc.ctypes.data,
                                                nodes.ctypes.data if n.value else None,
                                                qd.ctypes.data if n.value else None, C.byref(n)))
        return d, flags, bc, nodes, qd.reshape(-1, self.Q - 1)

    def _f_dtype(self):
        import torch

        return {0: torch.float32, 1: torch.int16, 2: torch.float64}[self.proto.storage]

    def get_f(self, c: int):
        """Current populations of case c (its tiled layout) as a device tensor."""
        import torch

        d = self.case_desc(c)
        out = torch.empty(d.nx * d.ny * d.nz * self.Q, dtype=self._f_dtype(), device="cuda")
        _lib.check(self.L.lbm_batch_get_f(self.ptr, int(c), out.data_ptr(), _stream()))
        return out

    def set_f(self, c: int, f) -> None:
        _lib.check(self.L.lbm_batch_set_f(self.ptr, int(c), f.data_ptr(), _stream()))

    def info(self) -> dict:
        out = (C.c_int64 * 9)()
        _lib.check(self.L.lbm_batch_info(self.ptr, out))
        keys = ("cases", "real_nodes", "padded_nodes", "tiles", "tx", "ty", "tz", "grid_blocks", "graph_steps")
        return dict(zip(keys, list(out)))

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.L.lbm_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Immersed circles / spheres with interpolated bounce-back links (host side).

add_sphere marks nodes with |x - c|^2 <= r^2 SOLID (src/lbm/geometry/primatives.mojo:110-161,
in lattice units). sphere_links marks the fluid nodes that pull from such a node
FLUID_SOLID_BOUNDARY and returns their link records for lbm_set_links: the
fraction of each fluid -> solid link lying in the fluid, the root in [0, 1] of
|x_f + t (x_s - x_f) - c|^2 = r^2 (get_q_for_sphere, src/lbm/geometry/rigidSphere.mojo:132-170).
"""
from __future__ import annotations

import numpy as np

from .constants import Flags
from .grid import LBM_Grid


def _coords(grid: LBM_Grid):
    return np.meshgrid(*[np.arange(n, dtype=np.float64) for n in grid.shape], indexing="ij")


def add_sphere(grid: LBM_Grid, flags_host: np.ndarray, center, radius: float) -> np.ndarray:
    """Marks nodes inside the sphere (circle for D = 2) SOLID; returns the boolean node mask [nx, ny, nz]."""
    x, y, z = _coords(grid)
    c = list(center) + [0.0] * (3 - len(center))
    d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + ((z - c[2]) ** 2 if grid.D == 3 else 0.0)
    inside = d2 <= radius * radius
    idx = grid.layouts.index(x[inside].astype(np.int64), y[inside].astype(np.int64), z[inside].astype(np.int64), 0, 1)
    flags_host[idx] = Flags.SOLID
    return inside


def sphere_links(grid: LBM_Grid, flags_host: np.ndarray, center, radius: float, inside: np.ndarray):
    """Link records of the fluid nodes next to the sphere: (nodes uint32 sorted, qdist float32 [n, Q-1]).
    Marks those nodes FLUID_SOLID_BOUNDARY in flags_host (nodes whose flag is FLUID)."""
    lat = grid.lattice
    nx, ny, nz = grid.shape
    c = np.array(list(center) + [0.0] * (3 - len(center)), np.float64)
    dirs = np.zeros((lat.Q, 3), np.int64)
    dirs[:, : lat.D] = lat.directions
    x, y, z = _coords(grid)
    qd = np.full((nx, ny, nz, lat.Q - 1), -1.0)
    fluid = ~inside
    for q in range(1, lat.Q):
        # pull neighbour x - c_q (periodic wrap, as the kernels do)
        nb_inside = np.roll(inside, shift=tuple(dirs[q]), axis=(0, 1, 2))
        link = fluid & nb_inside
        if not link.any():
            continue
        xf = np.stack([x[link], y[link], z[link]], axis=1)
        xs = xf - dirs[q][None, :]
        a = np.sum((xs - xf) ** 2, axis=1)
        b = 2.0 * np.sum((xf - c) * (xs - xf), axis=1)
        cc = np.sum((xf - c) ** 2, axis=1) - radius * radius
        disc = np.sqrt(np.maximum(b * b - 4 * a * cc, 0.0))
        t1 = (-b + disc) / (2 * a)
        t2 = (-b - disc) / (2 * a)
        t = np.where((t1 >= 0) & (t1 <= 1), t1, t2)
        qd[link, q - 1] = np.clip(t, 0.0, 1.0)
    has = (qd >= 0).any(axis=3)
    xs_, ys_, zs_ = np.nonzero(has)
    nodes = grid.layouts.index(xs_, ys_, zs_, 0, 1)
    order = np.argsort(nodes)
    nodes = nodes[order].astype(np.uint32)
    q = qd[xs_, ys_, zs_][order].astype(np.float32)
    sel = flags_host[nodes] == Flags.FLUID
    flags_host[nodes[sel]] = Flags.FLUID_SOLID_BOUNDARY
    return nodes, np.ascontiguousarray(q)

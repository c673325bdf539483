"""Immersed circles / spheres with interpolated bounce-back links (host side).

add_sphere marks nodes with |x - c|^2 <= r^2 SOLID (src/lbm/geometry/primatives.mojo:110-161,
in lattice units). sphere_links marks the fluid nodes that pull from such a node
FLUID_SOLID_BOUNDARY and returns their link records for lbm_set_links: the
fraction of each fluid -> solid link lying in the fluid, the root in [0, 1] of
|x_f + t (x_s - x_f) - c|^2 = r^2 (get_q_for_sphere, src/lbm/geometry/rigidSphere.mojo:132-170).
The *_rowmajor variants work on plain (nx, ny, nz) arrays with row-major node ids,
the form the batched-case API takes.
"""
from __future__ import annotations

import numpy as np

from .constants import Flags
from .grid import LBM_Grid


def _coords(shape):
    return np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")


def _center3(center):
    return np.array(list(center) + [0.0] * (3 - len(center)), np.float64)


def sphere_mask(shape, center, radius: float, D: int) -> np.ndarray:
    """Boolean [nx, ny, nz] mask of the nodes with |x - c|^2 <= r^2 (z ignored for D = 2)."""
    x, y, z = _coords(shape)
    c = _center3(center)
    d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + ((z - c[2]) ** 2 if D == 3 else 0.0)
    return d2 <= radius * radius


def link_fractions(directions: np.ndarray, inside: np.ndarray, center, radius: float):
    """Per node and direction q >= 1, the fraction of the link to the pull neighbour x - c_q lying in
    the fluid when that neighbour is inside the sphere and the node is not; -1 elsewhere.
    Returns qd [nx, ny, nz, Q-1] (the pull neighbour wraps periodically, as the kernels do)."""
    Q, D = directions.shape
    dirs = np.zeros((Q, 3), np.int64)
    dirs[:, :D] = directions
    c = _center3(center)
    x, y, z = _coords(inside.shape)
    qd = np.full(inside.shape + (Q - 1,), -1.0)
    fluid = ~inside
    for q in range(1, Q):
        nb_inside = np.roll(inside, shift=tuple(dirs[q]), axis=(0, 1, 2))  # nb_inside[x] = inside[x - c_q]
        link = fluid & nb_inside
        if not link.any():
            continue
        xf = np.stack([x[link], y[link], z[link]], axis=1)
        step = -dirs[q][None, :].astype(np.float64)  # x_s - x_f
        a = np.sum(step * step, axis=1)
        b = 2.0 * np.sum((xf - c) * step, axis=1)
        cc = np.sum((xf - c) ** 2, axis=1) - radius * radius
        disc = np.sqrt(np.maximum(b * b - 4 * a * cc, 0.0))
        t1 = (-b + disc) / (2 * a)
        t2 = (-b - disc) / (2 * a)
        t = np.where((t1 >= 0) & (t1 <= 1), t1, t2)
        qd[link, q - 1] = np.clip(t, 0.0, 1.0)
    return qd


def add_sphere(grid: LBM_Grid, flags_host: np.ndarray, center, radius: float) -> np.ndarray:
    """Marks nodes inside the sphere (circle for D = 2) SOLID in a tiled flag array; returns the
    boolean node mask [nx, ny, nz]."""
    inside = sphere_mask(grid.shape, center, radius, grid.D)
    xs, ys, zs = np.nonzero(inside)
    flags_host[grid.layouts.index(xs, ys, zs, 0, 1)] = Flags.SOLID
    return inside


def sphere_links(grid: LBM_Grid, flags_host: np.ndarray, center, radius: float, inside: np.ndarray):
    """Link records of the fluid nodes next to the sphere in the tiled layout: (nodes uint32
    increasing, qdist float32 [n, Q-1]). Marks those nodes FLUID_SOLID_BOUNDARY where they are FLUID."""
    qd = link_fractions(grid.lattice.directions, inside, center, radius)
    xs, ys, zs = np.nonzero((qd >= 0).any(axis=3))
    nodes = grid.layouts.index(xs, ys, zs, 0, 1)
    order = np.argsort(nodes)
    nodes = nodes[order].astype(np.uint32)
    q = qd[xs, ys, zs][order].astype(np.float32)
    sel = flags_host[nodes] == Flags.FLUID
    flags_host[nodes[sel]] = Flags.FLUID_SOLID_BOUNDARY
    return nodes, np.ascontiguousarray(q)


def sphere_links_rowmajor(directions: np.ndarray, flags: np.ndarray, center, radius: float, inside: np.ndarray):
    """sphere_links for a plain [nx, ny, nz] flag array: row-major node ids (x slowest)."""
    qd = link_fractions(directions, inside, center, radius)
    has = (qd >= 0).any(axis=3)
    xs, ys, zs = np.nonzero(has)  # row-major order already
    nodes = np.ravel_multi_index((xs, ys, zs), flags.shape).astype(np.uint32)
    q = qd[xs, ys, zs].astype(np.float32)
    sel = flags[xs, ys, zs] == Flags.FLUID
    flags[xs[sel], ys[sel], zs[sel]] = Flags.FLUID_SOLID_BOUNDARY
    return nodes, np.ascontiguousarray(q)

"""Immersed circles / spheres / boxes and interpolated bounce-back links (host side).

add_sphere / add_circle / add_box mark SOLID the nodes whose physical coordinates
(index * dx + origin, in the grid float type) lie in the sphere or box, scanning the
same bounding box as the reference (src/lbm/geometry/primatives.mojo:110-266), with
numpy instead of a per-node loop. sphere_links marks the fluid nodes that pull from such a node
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


def _flags_host(flags):
    return flags.cpu() if hasattr(flags, "cpu") and not isinstance(flags, np.ndarray) else flags


def _bbox(grid: LBM_Grid, center, half):
    """Index ranges of the bounding box, primatives.mojo:138-146 (floor division, clipped to the grid)."""
    out = []
    for i in range(3):
        if i < grid.D:
            lo, hi = center[i] - half[i], center[i] + half[i]
            n_min = max(0, int((lo - grid.origin[i]) // grid.dx))
            n_max = min(grid.shape[i], int((hi - grid.origin[i]) // grid.dx + 1))
            out.append(np.arange(n_min, max(n_min, n_max)))
        else:
            out.append(np.arange(0, 1))
    return out


def _box_coords(grid: LBM_Grid, rng):
    xs, ys, zs = np.meshgrid(*rng, indexing="ij")
    dt = grid.float_dtype
    dx = dt(grid.dx)
    pts = [dt(1) * np.asarray(i, dt) * dx + dt(o) for i, o in zip((xs, ys, zs), grid.origin)]
    return xs, ys, zs, pts


def add_sphere(grid: LBM_Grid, flags, center, radius: float) -> np.ndarray:
    """primatives.mojo:110-161: SOLID where sum_i (x_i - c_i)^2 <= r^2 (all three axes, in the grid
    float type). flags is a ContextTileTensor or a tiled host array; returns the boolean node
    mask [nx, ny, nz] of the nodes it marked."""
    if len(center) != 3:
        raise ValueError(f"centre must be a list of 3 floats got a len of {len(center)} instead")
    dt = grid.float_dtype
    c = [dt(v) for v in center]
    r = dt(radius)
    xs, ys, zs, pts = _box_coords(grid, _bbox(grid, [float(v) for v in c], [float(r)] * 3))
    d2 = (pts[0] - c[0]) ** 2
    d2 = d2 + (pts[1] - c[1]) ** 2
    d2 = d2 + (pts[2] - c[2]) ** 2
    sel = d2 <= r * r
    fl = _flags_host(flags)
    fl[grid.layouts.index(xs[sel], ys[sel], zs[sel], 0, 1)] = Flags.SOLID
    inside = np.zeros(grid.shape, bool)
    inside[xs[sel], ys[sel], zs[sel]] = True
    return inside


def add_circle(grid: LBM_Grid, flags, center, radius: float) -> np.ndarray:
    """primatives.mojo:163-183: add_sphere on a 2-D grid."""
    if grid.D != 2:
        raise ValueError("Circle only valid for 2D")
    return add_sphere(grid, flags, center, radius)


def add_box(grid: LBM_Grid, flags, center, box_radius) -> np.ndarray:
    """primatives.mojo:186-243 / check_box_axis :246-266: SOLID where c - h <= x <= c + h on every axis."""
    if len(center) != 3:
        raise ValueError(f"centre must be a list of 3 floats got a len of {len(center)} instead")
    if len(box_radius) != 3:
        raise ValueError(f"lengths must be a list of 3 floats got a len of {len(box_radius)} instead")
    dt = grid.float_dtype
    c = [dt(v) for v in center]
    h = [dt(v) for v in box_radius]
    xs, ys, zs, pts = _box_coords(grid, _bbox(grid, [float(v) for v in c], [float(v) for v in h]))
    sel = np.ones(xs.shape, bool)
    for i in range(3):
        sel &= ((c[i] - h[i]) <= pts[i]) & (pts[i] <= (c[i] + h[i]))
    fl = _flags_host(flags)
    fl[grid.layouts.index(xs[sel], ys[sel], zs[sel], 0, 1)] = Flags.SOLID
    inside = np.zeros(grid.shape, bool)
    inside[xs[sel], ys[sel], zs[sel]] = True
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

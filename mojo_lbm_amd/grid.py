"""LBM_Grid and TiledLayouts (src/lbm/grid.mojo:55-277, src/lbm/layoutstruct.mojo:18-146).

The (x, y, z, k) fields use the blocked product of a column-major tile
(tx, ty, tz, K) and a column-major tiler (Tx, Ty, Tz, 1):
    addr = lx + tx*(ly + ty*(lz + tz*k)) + tx*ty*tz*K*(bx + Tx*(by + Ty*bz)).
"""
from __future__ import annotations

import numpy as np

from .lattice import Lattice
from .units import UnitSystem


def set_tile_shape(tile_shape, D: int):
    """set_tile_shape (grid.mojo:370-388): an int expands to (t, t|1, t|1)."""
    if isinstance(tile_shape, int):
        t = tile_shape
        return (t, t if D >= 2 else 1, t if D == 3 else 1)
    return tuple(int(v) for v in tile_shape)


def set_default_block_shape(tile_shape, D: int):
    """set_default_block_shape (grid.mojo:392-416)."""
    if tuple(tile_shape) == (1, 1, 1):
        return {1: (256, 1, 1), 2: (16, 16, 1)}.get(D, (8, 8, 4))
    for i in range(3):
        ok = tile_shape[i] > 1 if i < D else tile_shape[i] == 1
        if not ok:
            raise ValueError("tile_shape must be > 1 along lattice axes and 1 elsewhere")
    return tuple(tile_shape)


def set_grid_dims(nx, ny, nz, block_shape):
    """set_grid_dims (grid.mojo:419-440)."""
    if not ((nx % block_shape[0] == 0 or nx == 1) and (ny % block_shape[1] == 0 or ny == 1)
            and (nz % block_shape[2] == 0 or nz == 1)):
        raise ValueError("The grid shape along nx,ny and nz should be either equal to 1 or divisible by the block shape")
    return (nx // block_shape[0], ny // block_shape[1], nz // block_shape[2])


class TiledLayouts:
    """Tile / tiler sizes of the flag, f and bc fields (layoutstruct.mojo:18-146)."""

    def __init__(self, D: int, Q: int, grid_shape, tile_shape):
        self.D, self.Q = D, Q
        self.grid_shape = tuple(grid_shape)
        self.tile_shape = tuple(tile_shape)
        self.n_tiles_x = grid_shape[0] // tile_shape[0]
        self.n_tiles_y = grid_shape[1] // tile_shape[1] if D >= 2 else 1
        self.n_tiles_z = grid_shape[2] // tile_shape[2] if D == 3 else 1
        self.tile_cells = tile_shape[0] * tile_shape[1] * tile_shape[2]
        self.num_points = grid_shape[0] * grid_shape[1] * grid_shape[2]

    def index(self, x, y, z, k=0, K=1):
        """Element address of (x, y, z, k) in a field with K values per node (numpy broadcasting)."""
        tx, ty, tz = self.tile_shape
        x, y, z, k = (np.asarray(v, dtype=np.int64) for v in (x, y, z, k))
        loc = x % tx + tx * (y % ty + ty * (z % tz + tz * k))
        tile = x // tx + self.n_tiles_x * (y // ty + self.n_tiles_y * (z // tz))
        return loc + self.tile_cells * K * tile

    def index_map(self, K: int) -> np.ndarray:
        """addr[x, y, z, k] for every node and value."""
        nx, ny, nz = self.grid_shape
        x, y, z, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), np.arange(K), indexing="ij")
        return self.index(x, y, z, k, K)

    @property
    def f_size(self):
        return self.num_points * self.Q

    @property
    def bc_size(self):
        return self.num_points * (self.D + 1)

    @property
    def flag_size(self):
        return self.num_points


class LBM_Grid:
    """LBM_Grid[lattice, nx, ny, nz, tile_shape](dx, origin) with the launch geometry and layouts."""

    def __init__(self, lattice: Lattice, nx: int, ny: int, nz: int, tile_shape=8, dx: float = 1.0,
                 origin=(0.0, 0.0, 0.0), float_dtype=np.float32):
        self.lattice = lattice
        self.D, self.Q = lattice.D, lattice.Q
        if self.D == 2 and nz != 1:
            raise ValueError("A 2-D lattice needs nz = 1")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (nx, ny, nz)
        self.tile_shape = set_tile_shape(tile_shape, self.D)
        self.BLOCK_SHAPE = set_default_block_shape(self.tile_shape, self.D)
        self.GRID_DIM = set_grid_dims(nx, ny, nz, self.BLOCK_SHAPE)
        self.THREADS_PER_BLOCK = self.BLOCK_SHAPE[0] * self.BLOCK_SHAPE[1] * self.BLOCK_SHAPE[2]
        self.layouts = TiledLayouts(self.D, self.Q, self.shape, self.tile_shape)
        self.num_points = nx * ny * nz
        self.float_dtype = np.dtype(float_dtype).type
        self.dx = float(dx)
        self.area = self.dx ** 2
        self.volume = self.dx ** 3
        self.domain_size = ((nx - 1) * self.dx, (ny - 1) * self.dx, (nz - 1) * self.dx)
        self.origin = tuple(float(o) for o in origin)

    def get_grid_coordinates(self, i, j, k):
        return tuple(idx * self.dx + o for idx, o in zip((i, j, k), self.origin))

    def get_UnitSystem_with_Re(self, U_phys, U_lattice, L_phys, Re, density=1.0) -> UnitSystem:
        """grid.mojo:170-194: L_lattice = L_phys / dx, nu = U L / Re."""
        L_lattice = L_phys / self.dx
        return UnitSystem(U_phys, U_lattice, L_phys, L_lattice, density, U_phys * L_phys / Re, D=self.D)

    def get_UnitSystem(self, U_phys, U_lattice, L_phys, kinematic_viscosity, density=1.0) -> UnitSystem:
        return UnitSystem(U_phys, U_lattice, L_phys, L_phys / self.dx, density, kinematic_viscosity, D=self.D)

    def get_UnitSystem_with_tau(self, U_phys, tau, L_phys, kinematic_viscosity, density=1.0, *, min_u=0.001,
                                max_u=0.1) -> UnitSystem:
        """grid.mojo:238-277: the lattice velocity that gives relaxation time tau at this Re;
        raises when it falls outside [min_u, max_u]."""
        L_lattice = L_phys / self.dx
        v_lat = (tau - 0.5) * (1 / 3.0)
        re = U_phys * L_phys / kinematic_viscosity
        U_lattice = v_lat * re / L_lattice
        if U_lattice < min_u or U_lattice > max_u:
            raise ValueError(f"Target Tau of {tau} does not satisfy bounds. Caclulated u lattice = {U_lattice} "
                             f"which is outside the given bounds of [{min_u} {max_u}]")
        return UnitSystem(U_phys, U_lattice, L_phys, L_lattice, density, kinematic_viscosity, D=self.D)

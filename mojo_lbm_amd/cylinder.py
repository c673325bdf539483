"""2-D cylinder-in-channel cases (the DFG / Schaefer-Turek set-up of
benchmarks/Cylinder2D/CylinderFlow_Re20.mojo:96-150) as row-major host arrays.

Channel of nx x ny nodes: the -X face is a SOLID inlet with the parabolic
velocity u(j) = 4 u_max j (ny-1-j) / (ny-1)^2 (the reference's inlet function with
y = j dx, L = (N-1) dx), the +X face an EQUILIBRIUM outlet with rho = 1 and free
velocity, the -Y/+Y faces SOLID walls, set in the reference's order (+X, -X, -Y, +Y).
The cylinder (|x - c| <= D/2) is SOLID with interpolated bounce-back links on its
fluid neighbours. Re = u_mean D / nu with u_mean = 2/3 u_max, and the force
coefficients are C = 2 F / (u_mean^2 D) (CylinderFlow_Re20.mojo:207-208).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .constants import Flags
from .geometry import sphere_links_rowmajor, sphere_mask
from .lattice import Lattice


@dataclass
class CylinderCase:
    nx: int
    ny: int
    diameter: float
    cx: float
    cy: float
    re: float
    u_max: float = 0.05

    @property
    def u_mean(self) -> float:
        return 2.0 / 3.0 * self.u_max

    @property
    def nu(self) -> float:
        return self.u_mean * self.diameter / self.re

    @property
    def tau(self) -> float:
        return 0.5 + 3.0 * self.nu

    @property
    def nodes(self) -> int:
        return self.nx * self.ny

    def arrays(self, lattice: Lattice):
        """flags [nx, ny, 1] uint8, bc [nx, ny, 1, 3] float64, (link nodes, qdist)."""
        nx, ny = self.nx, self.ny
        flags = np.zeros((nx, ny, 1), np.uint8)
        bc = np.zeros((nx, ny, 1, 3), np.float64)
        bc[..., 2] = 1.0
        flags[nx - 1] = Flags.EQUILIBRIUM
        bc[nx - 1, :, :, 0:2] = np.nan
        j = np.arange(ny, dtype=np.float64)
        flags[0] = Flags.SOLID
        bc[0, :, 0, 0] = 4.0 * self.u_max * j * (ny - 1 - j) / float(ny - 1) ** 2
        bc[0, :, 0, 1] = 0.0
        for y in (0, ny - 1):
            flags[:, y] = Flags.SOLID
            bc[:, y, :, 0:2] = 0.0
            bc[:, y, :, 2] = 1.0
        inside = sphere_mask((nx, ny, 1), (self.cx, self.cy), self.diameter / 2.0, 2)
        flags[inside] = Flags.SOLID
        links = sphere_links_rowmajor(lattice.directions, flags, (self.cx, self.cy), self.diameter / 2.0, inside)
        return flags, bc, links

    def coefficients(self, force) -> tuple[float, float]:
        """(C_drag, C_lift) from the momentum-exchange force on the cylinder."""
        s = 2.0 / (self.u_mean ** 2 * self.diameter)
        return float(force[0] * s), float(force[1] * s)


def random_cases(n: int = 64, seed: int = 2024, u_max: float = 0.05) -> list[CylinderCase]:
    """Seeded case set: nx ~ U{200..800}, ny = round(nx U[0.2, 0.35]), D ~ U[10, min(40, 0.4 ny)],
    centre (nx/5, ny/2 + U[-0.05, 0.05] ny), Re ~ U[20, 200]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(200, 801))
        ny = int(round(nx * rng.uniform(0.2, 0.35)))
        d = float(rng.uniform(10.0, min(40.0, 0.4 * ny)))
        cy = ny / 2.0 + float(rng.uniform(-0.05, 0.05)) * ny
        re = float(rng.uniform(20.0, 200.0))
        out.append(CylinderCase(nx=nx, ny=ny, diameter=d, cx=nx / 5.0, cy=cy, re=re, u_max=u_max))
    return out

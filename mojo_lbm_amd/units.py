"""Physical <-> lattice units and tau from Re (src/lbm/units.mojo:46-210)."""
from __future__ import annotations


class Unit_:
    def __init__(self, physical: float, lattice: float):
        self.physical = float(physical)
        self.lattice = float(lattice)

    def C_lat_to_phys(self) -> float:
        return self.physical / self.lattice

    def C_phys_to_lat(self) -> float:
        return self.lattice / self.physical


class UnitSystem:
    """UnitSystem(u_physical, u_lattice, L_physical, L_lattice, density, kinematic_viscosity).

    tau = nu_lat / cs^2 + 0.5 with nu_lat = U_lat * L_lat / Re (units.mojo:116-126).
    """

    def __init__(self, u_physical, u_lattice, L_physical, L_lattice, density, kinematic_viscosity, D=3):
        if not (0.001 <= u_lattice <= 0.1 + 1e-12):
            import warnings

            warnings.warn("u_lattice outside the standard range [0.001, 0.1]")
        self.D = D
        self.U = Unit_(u_physical, u_lattice)
        self.L = Unit_(L_physical, L_lattice)
        self.t = Unit_((L_physical / u_physical) / (L_lattice / u_lattice), 1.0)
        self.density = Unit_(density, 1.0)
        self.Re = (self.U.physical * self.L.physical) / kinematic_viscosity
        v_lat = self.U.lattice * self.L.lattice / self.Re
        self.dt = self.t.physical
        self.tau = v_lat / (1 / 3.0) + 0.5
        self.kinematic_viscosity = Unit_(kinematic_viscosity, v_lat)
        self.mass = Unit_(self.density.C_lat_to_phys() * (self.L.C_lat_to_phys() ** D), 1.0)


def tau_from_re(u_lattice: float, L_lattice: float, Re: float) -> float:
    """tau = 3 u L / Re + 0.5 (the closed form UnitSystem evaluates)."""
    return 3.0 * u_lattice * L_lattice / Re + 0.5

"""MI355X (gfx950) lattice Boltzmann stream-collide engine.

The host API mirrors mojo_lbm's LBM_Config / lattice / LBM_Grid / Solver /
Assembly names; every step runs in liblbm_hip.so (hand-written HIP kernels
behind a C ABI, include/lbm_hip.h).
"""
from .config import DoubleBufferConfig, EsotericPullConfig, LBM_Config
from .constants import (DOUBLE_BUFFER, ESOTERIC_PULL, FLUID_NODE, SOLID_NODE, Collisions, Flags, LBM_method, cs,
                        cs_squared)
from .grid import LBM_Grid, TiledLayouts
from .lattice import Lattice, get_D2Q9, get_D3Q19, get_D3Q27
from .units import UnitSystem, tau_from_re

__all__ = [
    "LBM_Config", "DoubleBufferConfig", "EsotericPullConfig", "Flags", "LBM_method", "Collisions",
    "DOUBLE_BUFFER", "ESOTERIC_PULL", "FLUID_NODE", "SOLID_NODE", "cs", "cs_squared", "LBM_Grid",
    "TiledLayouts", "Lattice", "get_D2Q9", "get_D3Q19", "get_D3Q27", "UnitSystem", "tau_from_re",
    "Solver", "Assembly", "ContextTileTensor", "set_exterior_walls", "set_exterior_walls_with_func",
    "initialize_fluid_at_rest", "initialize_equilibrium", "initialize_f_from_func", "calculate_rho_and_velocity",
    "calculate_Q_criterion", "reduce_moments", "OutputRequest", "check_health",
]


def __getattr__(name):
    # Modules that need torch / the HIP library load on first use.
    if name == "Solver":
        from .solver import Solver
        return Solver
    if name == "Assembly":
        from .assembly import Assembly
        return Assembly
    if name == "ContextTileTensor":
        from .tensor import ContextTileTensor
        return ContextTileTensor
    if name in ("set_exterior_walls", "set_exterior_walls_with_func", "initialize_fluid_at_rest",
                "initialize_equilibrium", "initialize_f_from_func"):
        from . import preprocess
        return getattr(preprocess, name)
    if name in ("calculate_rho_and_velocity", "calculate_Q_criterion", "reduce_moments", "OutputRequest",
                "check_health"):
        from . import output
        return getattr(output, name)
    if name == "snapshot":
        import importlib
        return importlib.import_module(".snapshot", __name__)
    raise AttributeError(name)

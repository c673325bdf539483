"""Flags, method and collision names (src/lbm/constants.mojo:20-123)."""
from __future__ import annotations

cs_squared = 1.0 / 3.0
cs = cs_squared ** 0.5


class LBM_method:
    """Streaming methods (constants.mojo:20-31)."""

    DOUBLE_BUFFER = 0
    ESOTERIC_PULL = 1
    valid_set = {DOUBLE_BUFFER, ESOTERIC_PULL}


DOUBLE_BUFFER = LBM_method.DOUBLE_BUFFER
ESOTERIC_PULL = LBM_method.ESOTERIC_PULL


class Collisions:
    """Collision operator names (constants.mojo:41-55)."""

    SRT = "SRT"
    TRT = "TRT"
    KBC = "KBC"
    RLBM = "RLBM"
    valid_set = {SRT, TRT, KBC, RLBM}
    that_need_fneq = {KBC, RLBM}
    # ids of include/lbm_hip.h
    ids = {SRT: 0, TRT: 1, KBC: 2, RLBM: 3}


class Flags:
    """Node flags: an enum in the low 4 bits and bit masks above (constants.mojo:58-117)."""

    FLUID = 0
    SOLID = 1
    EQUILIBRIUM = 2
    FLUID_SOLID_BOUNDARY = 3
    EXCLUDE = 0b10000000
    CAPTURE_DENSITY = 0b01000000
    CAPTURE_VELOCITY = 0b00100000
    valid_enums = {FLUID, SOLID, EQUILIBRIUM, FLUID_SOLID_BOUNDARY}
    valid_bitmasks = {EXCLUDE, CAPTURE_DENSITY, CAPTURE_VELOCITY}

    @staticmethod
    def get_enum_flag(flag: int) -> int:
        return flag & 0x0F

    @staticmethod
    def get_bitmask_flags(flag: int) -> int:
        return flag >> 4

    @classmethod
    def is_(cls, enum_flag: int, flag: int) -> bool:
        assert enum_flag in cls.valid_enums
        return cls.get_enum_flag(flag) == enum_flag

    @classmethod
    def has(cls, bitmask_flag: int, flag: int) -> bool:
        assert bitmask_flag in cls.valid_bitmasks
        return (bitmask_flag & flag) != 0

    @classmethod
    def add_exclude_to_flag(cls, flag: int) -> int:
        return cls.EXCLUDE | flag

    @classmethod
    def is_excluded(cls, flag: int) -> bool:
        return bool(cls.EXCLUDE & flag)


class Bounceback_method:
    MID_GRID = 0
    BOUZIDI = 1


_FlagSet = {Flags.FLUID, Flags.SOLID, Flags.EQUILIBRIUM}
FLUID_NODE = Flags.FLUID
SOLID_NODE = Flags.SOLID

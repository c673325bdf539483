"""LBM_Config: the run toggles of src/lbm/config.mojo:50-190, mapped onto lbm_desc."""
from __future__ import annotations

import warnings

import numpy as np

from .constants import Collisions, Flags, LBM_method, _FlagSet


class LBM_Config:
    """Run toggles with the reference's keyword names (config.mojo:85-99).

    ``LBM_Config(LBM_method.DOUBLE_BUFFER, collision_op='SRT', LES=False, BCs=set(),
    DDF_shift=False, use_float16c=False, f_dtype=None, include_moving_boundary=True)``.
    ``f_dtype=np.float64`` selects FP64 populations (the FP64 reference runs).
    """

    def __init__(self, lbm_method=LBM_method.DOUBLE_BUFFER, *, collision_op: str = "SRT", LES: bool = False,
                 BCs=frozenset(), DDF_shift: bool = False, use_float16c: bool = False, f_dtype=None,
                 include_moving_boundary: bool = True, capture_density: bool = False,
                 capture_velocity: bool = False, Cs: float = 0.1, LES_model: str = "smagorinsky"):
        if lbm_method not in LBM_method.valid_set:
            raise ValueError(f"lbm_method not valid: {lbm_method}")
        self.lbm_method = lbm_method
        self.collision_op = collision_op
        self.LES = bool(LES)
        if LES_model not in ("smagorinsky", "mlp"):
            raise ValueError(f"LES_model must be 'smagorinsky' or 'mlp', not {LES_model!r}")
        self.LES_model = LES_model
        self.DDF_shift = bool(DDF_shift)
        self.use_float16c = bool(use_float16c)
        self.include_moving_boundary = bool(include_moving_boundary)
        self.capture_density = bool(capture_density)
        self.capture_velocity = bool(capture_velocity)
        self.second_moment = self.LES
        self.Cs = float(Cs)
        if use_float16c and f_dtype is not None:
            warnings.warn("Float16c set to True and f_dtype was also specified. Float16c overides the given dtype")
        self.f_dtype = np.uint16 if use_float16c else (np.dtype(f_dtype).type if f_dtype is not None else None)
        BCs = set(BCs)
        if BCs and not BCs.issubset(_FlagSet):
            warnings.warn(f"Some Specified BC types are not standard: {BCs - _FlagSet}")
        self.INCLUDED_BCs = {Flags.FLUID, Flags.SOLID} | BCs
        if not self.collision_op_is_valid():
            raise ValueError(f"Invalid collision operator {collision_op!r}")

    @property
    def les_id(self) -> int:
        """lbm_desc.les: 0 none, 1 Smagorinsky (the reference's LES), 2 the learned closure."""
        if not self.LES:
            return 0
        return 2 if self.LES_model == "mlp" else 1

    def implies_f_noneq(self, *, force: bool = False) -> bool:
        return self.LES or self.collision_op in Collisions.that_need_fneq or force

    def implies_get_adjacent_flags(self, *, force: bool = False) -> bool:
        return self.lbm_method == LBM_method.DOUBLE_BUFFER or self.include_moving_boundary or force

    def implies_bc_is_mutable(self) -> bool:
        return self.capture_density or self.capture_velocity

    def collision_op_is_valid(self) -> bool:
        return self.collision_op in Collisions.valid_set

    def set_f_dtype(self, float_dtype_for_math_ops):
        return self.f_dtype if self.f_dtype is not None else float_dtype_for_math_ops

    @property
    def storage_id(self) -> int:
        if self.use_float16c:
            return 1
        if self.f_dtype is not None and np.dtype(self.f_dtype) == np.float64:
            return 2
        return 0

    @property
    def included_bcs_mask(self) -> int:
        """lbm_desc.included_bcs: bit (1 << flag) per BC type, plus the capture flag values
        (Flags.CAPTURE_DENSITY / CAPTURE_VELOCITY) when capture_density / capture_velocity are on."""
        m = 0
        for b in self.INCLUDED_BCs:
            if 0 <= b < 32:
                m |= 1 << b
        if self.capture_density:
            m |= Flags.CAPTURE_DENSITY
        if self.capture_velocity:
            m |= Flags.CAPTURE_VELOCITY
        return m


def DoubleBufferConfig(**kw) -> LBM_Config:
    """LBM_Config[LBM_method.DOUBLE_BUFFER] (config.mojo:19)."""
    return LBM_Config(LBM_method.DOUBLE_BUFFER, **kw)


def EsotericPullConfig(**kw) -> LBM_Config:
    """LBM_Config[LBM_method.ESOTERIC_PULL] (config.mojo:23)."""
    return LBM_Config(LBM_method.ESOTERIC_PULL, **kw)

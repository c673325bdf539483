"""Assembly: owns the f (and f2), bc and flags buffer pairs (src/lbm/assembly.mojo:43-269)."""
from __future__ import annotations

import numpy as np

from .config import LBM_Config
from .constants import LBM_method
from .grid import LBM_Grid
from .tensor import ContextTileTensor


class Assembly:
    def __init__(self, grid: LBM_Grid, config: LBM_Config):
        self.grid = grid
        self.config = config
        lay = grid.layouts
        f_dtype = config.set_f_dtype(grid.float_dtype)
        bc_dtype = np.float64 if config.storage_id == 2 else grid.float_dtype
        self.flags = ContextTileTensor(np.uint8, lay.flag_size)
        self.bc = ContextTileTensor(bc_dtype, lay.bc_size)
        self.f = ContextTileTensor(f_dtype, lay.f_size)
        self.f2 = ContextTileTensor(f_dtype, lay.f_size) if config.lbm_method == LBM_method.DOUBLE_BUFFER else None

    def get_gpu_tensors_for_double_buffer(self):
        """(f, f2, bc, flags) device buffers; raises for an esoteric-pull config (assembly.mojo:141-164)."""
        if self.config.lbm_method != LBM_method.DOUBLE_BUFFER or self.f2 is None:
            raise RuntimeError("Assembly was built for esoteric pull: there is no second f buffer")
        return self.f.gpu(), self.f2.gpu(), self.bc.gpu(), self.flags.gpu()

    def get_gpu_tensors_for_esoteric_pull(self):
        """(f, bc, flags) device buffers; raises for a double-buffer config (assembly.mojo:100-139)."""
        if self.config.lbm_method != LBM_method.ESOTERIC_PULL:
            raise RuntimeError("Assembly was built for double buffer, not esoteric pull")
        return self.f.gpu(), self.bc.gpu(), self.flags.gpu()

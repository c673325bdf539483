"""ContextTileTensor: a host numpy array paired with a device buffer, copied only
when the side in use changes (src/utils/contextTileTensor.mojo:55-293).

Device memory is a torch CUDA tensor (allocation and stream plumbing only); all
compute goes through liblbm_hip.so.
"""
from __future__ import annotations

import numpy as np


class ContextTileTensor:
    def __init__(self, dtype, size: int, fill=0):
        import torch

        self._torch = torch
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        self._host = np.full(self.size, fill, dtype=self.dtype)
        self._dev = None
        self._last = "cpu"  # side holding the current data

    def cpu(self) -> np.ndarray:
        """Host view; copies device -> host first when the device side was used last."""
        if self._last == "gpu":
            self._torch.cuda.synchronize()
            self._host[...] = self._dev.cpu().numpy().view(self.dtype)
            self._last = "cpu"
        return self._host

    def gpu(self):
        """Device tensor (uint8 view of the buffer); copies host -> device first when needed."""
        t = self._torch
        if self._dev is None:
            self._dev = t.empty(self.size * self.dtype.itemsize, dtype=t.uint8, device="cuda")
            self._last = "cpu"
        if self._last == "cpu":
            self._dev.copy_(t.from_numpy(self._host.view(np.uint8)))
            self._last = "gpu"
        return self._dev

    def mark_device_written(self) -> None:
        """The device buffer was written by a kernel: the host copy is stale."""
        if self._dev is None:
            self.gpu()
        self._last = "gpu"

    def data_ptr(self) -> int:
        return self.gpu().data_ptr()

    def buffer_to_numpy(self) -> np.ndarray:
        """Host array (contextTileTensor.mojo:296-349)."""
        return self.cpu()

"""Host<->device helpers for the GPU parity tests."""
from __future__ import annotations

import numpy as np
import torch

_SIGNED = {1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}
_NP_SIGNED = {1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}
TORCH_OF = {
    np.dtype(np.uint8): torch.uint8, np.dtype(np.uint16): torch.uint16, np.dtype(np.uint32): torch.uint32,
    np.dtype(np.uint64): torch.uint64, np.dtype(np.int8): torch.int8, np.dtype(np.int16): torch.int16,
    np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64, np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
}


def to_dev(x: np.ndarray, dev) -> torch.Tensor:
    """Copy a numpy array to the GPU bit for bit (through a signed view)."""
    x = np.ascontiguousarray(x)
    w = x.dtype.itemsize
    t = torch.from_numpy(x.view(_NP_SIGNED[w]).copy()).to(dev)
    return t.view(TORCH_OF[x.dtype])


def to_host(t: torch.Tensor, np_dtype) -> np.ndarray:
    w = t.element_size()
    return t.view(_SIGNED[w]).cpu().numpy().view(np_dtype)


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize])

"""Tensor plumbing: torch device tensors are the handles passed to the C ABI."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

_DT = {
    torch.float32: _lib.PC_F32, torch.float64: _lib.PC_F64,
    torch.complex64: _lib.PC_C64, torch.complex128: _lib.PC_C128,
}
_REAL = {torch.float32: torch.float32, torch.float64: torch.float64,
         torch.complex64: torch.float32, torch.complex128: torch.float64}


def pc_dtype(t: torch.Tensor) -> int:
    try:
        return _DT[t.dtype]
    except KeyError:
        raise TypeError(f"unsupported dtype {t.dtype}; use float32, float64, complex64 or complex128") from None


def real_dtype(dt: torch.dtype) -> torch.dtype:
    return _REAL[dt]


class Arg:
    """A device matrix view: (tensor, leading dimension) plus how to give results back."""

    def __init__(self, x, name="A", dtype=None):
        self.was_numpy = isinstance(x, np.ndarray)
        if self.was_numpy:
            arr = np.ascontiguousarray(x)
            if arr.dtype.kind in "iub":
                arr = arr.astype(np.float64)
            t = torch.from_numpy(arr).cuda()
        elif isinstance(x, torch.Tensor):
            t = x if x.is_cuda else x.cuda()
        else:
            raise TypeError(f"{name} must be a numpy array or a torch tensor")
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        if t.dtype not in _DT:
            t = t.to(torch.float64)
        if t.dim() == 1:
            t = t.reshape(-1, 1) if name != "v" else t
        if t.dim() == 2 and (t.stride(1) != 1 or t.stride(0) < t.shape[1]):
            t = t.contiguous()
        self.t = t

    @property
    def ld(self) -> int:
        return self.t.stride(0) if self.t.dim() == 2 else 1

    def ptr(self) -> int:
        return self.t.data_ptr()

    def give(self, out: torch.Tensor):
        return out.cpu().numpy() if self.was_numpy else out


def empty(shape, dtype, like: torch.Tensor):
    return torch.empty(shape, dtype=dtype, device=like.device)

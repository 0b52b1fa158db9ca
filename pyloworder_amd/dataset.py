"""``pyLOM.Dataset`` subset on the device: the fields of a dataset and the
assembly of the snapshot matrix X (pyLOM/dataset.py:252-292) that POD, SPOD
and DMD take.  The mesh, partition table and HDF5 I/O of the reference are
not part of this library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._tensor import Arg, empty, pc_dtype
from .cr import cr
from .vmmath import _ctx

__all__ = ["Dataset"]


class Dataset:
    """Fields on ``npoints`` points: ``fields[name] = {'ndim': k, 'value': V}``
    with V of shape (k * npoints, ...) (components of a point consecutive), as
    pyLOM/dataset.py keeps them."""

    def __init__(self, npoints: int, fields: dict | None = None):
        self.npoints = int(npoints)
        self.fields = {}
        for name, f in (fields or {}).items():
            self.add_field(name, f["ndim"], f["value"])

    def __len__(self):
        return self.npoints

    @property
    def fieldnames(self):
        return list(self.fields)

    def add_field(self, name: str, ndim: int, value):
        a = Arg(value, name)
        if a.t.shape[0] != ndim * self.npoints:
            raise ValueError(f"field {name}: {ndim * self.npoints} rows expected, got {a.t.shape[0]}")
        self.fields[name] = {"ndim": int(ndim), "value": value, "_arg": a}

    @cr("Dataset.X")
    def X(self, *args):
        """Snapshot matrix of the given fields (all if none), rows interleaved
        per point: row p * nc + c is component c of point p, c running over the
        fields in the order given and their components
        (pyLOM/dataset.py:252-292).  One gather kernel."""
        names = list(args) if args else self.fieldnames
        if not names:
            raise ValueError("Dataset.X: no fields")
        args_ = [self.fields[n]["_arg"] for n in names]
        dt = args_[0].t.dtype
        trail = tuple(args_[0].t.shape[1:])
        for a in args_:
            if a.t.dtype != dt or tuple(a.t.shape[1:]) != trail:
                raise ValueError("Dataset.X: fields differ in dtype or in their trailing dimensions")
        flat = [a.t.reshape(a.t.shape[0], -1) for a in args_]
        nt = flat[0].shape[1]
        nc = sum(self.fields[n]["ndim"] for n in names)
        X = empty((self.npoints * nc, nt), dt, flat[0])
        k = len(names)
        srcs = (C.c_void_p * k)(*[f.data_ptr() for f in flat])
        ndim = (C.c_int64 * k)(*[self.fields[n]["ndim"] for n in names])
        lds = (C.c_int64 * k)(*[f.stride(0) for f in flat])
        ctx = _ctx()
        ctx.check(ctx.lib.pc_dataset_x(ctx.h, pc_dtype(X), k, srcs, ndim, lds, self.npoints, nt, X.data_ptr(),
                                       X.stride(0)))
        X = X.reshape((self.npoints * nc,) + trail)
        return X.cpu().numpy() if args_[0].was_numpy else X

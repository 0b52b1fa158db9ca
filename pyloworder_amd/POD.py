"""``pyLOM.POD`` counterparts: run / truncate / reconstruct (+ project).

Signatures and (PSI, S, V) outputs follow pyLOM/POD/wrapper.py:15-103.  The
mean (and variance) removal of ``run`` is fused into the MI355X passes; the
input is never copied or modified.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._tensor import Arg, empty, pc_dtype, real_dtype
from . import vmmath
from .cr import cr, cr_start, cr_stop
from .vmmath import _ctx, _prec, compute_truncation_residual


@cr("POD.run")
def run(X, remove_mean: bool = True, divide_variance: bool = False, randomized: bool = False, r: int = 1,
        q: int = 3, seed: int = -1, precision: str = "fp64"):
    """POD of X (m points x n snapshots) — pyLOM/POD/wrapper.py:15-51.

    Returns (U (m, k), S (k,), V (k, n)) with k = n (tsqr_svd) or r (randomized),
    V holding the right singular vectors as rows.
    """
    center = (2 if divide_variance else 1) if remove_mean else 0
    cr_start("POD.SVD", 0)
    try:
        if not randomized:
            return vmmath.tsqr_svd(X, precision=precision, center=center)
        if center:
            cr_start("POD.temporal_mean", 0)
            mu = vmmath.temporal_mean(X if isinstance(X, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(X)).cuda())
            Xt = X if isinstance(X, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(X)).cuda()
            if divide_variance:
                Y = vmmath.norm_variance(Xt, mu, vmmath.temporal_variance(Xt, mu))
            else:
                Y = vmmath.subtract_mean(Xt, mu)
            cr_stop("POD.temporal_mean", 0)
            U, S, V = vmmath.randomized_svd(Y, r, q, seed=seed, precision=precision)
            if isinstance(X, np.ndarray):
                return U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy()
            return U, S, V
        return vmmath.randomized_svd(X, r, q, seed=seed, precision=precision)
    finally:
        cr_stop("POD.SVD", 0)


@cr("POD.truncate")
def truncate(U, S, V, r: float = 1e-8):
    """Keep N modes — pyLOM/POD/wrapper.py:54-82 (views, no copies).

    r >= 1: N = int(r); 0 < r < 1: residual target; r < 0: cumulative energy.
    """
    N = int(r) if r >= 1 else compute_truncation_residual(S, r)
    return U[:, :N], S[:N], V[:N, :]


@cr("POD.reconstruct")
def reconstruct(U, S, V, precision: str = "fp64"):
    """X = U diag(S) V — pyLOM/POD/wrapper.py:85-103, one fused pass."""
    u = Arg(U, "U")
    s = Arg(S, "v", dtype=real_dtype(u.t.dtype)).t.contiguous().reshape(-1)
    v = Arg(V, "V", dtype=u.t.dtype)
    m, N = u.t.shape
    N2, n = v.t.shape
    if N2 != N or s.shape[0] != N:
        raise ValueError("reconstruct: U (m,N), S (N,), V (N,n) expected")
    X = empty((m, n), u.t.dtype, u.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_reconstruct(ctx.h, pc_dtype(u.t), u.ptr(), u.ld, s.data_ptr(), v.ptr(), v.ld, m, N, n,
                                     X.data_ptr(), X.stride(0), _prec(precision)))
    return u.give(X)


@cr("POD.project")
def project(U, X, precision: str = "fp64"):
    """Modal coefficients C = U^H X (N x n).  No reference function; it is the
    idiom matmulp(transpose(U), Y) of pyLOM/DMD/wrapper.py:85."""
    u = Arg(U, "U")
    x = Arg(X, "X", dtype=u.t.dtype)
    m, N = u.t.shape
    m2, n = x.t.shape
    if m2 != m:
        raise ValueError("project: U and X must have the same number of rows")
    Cm = empty((N, n), u.t.dtype, u.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_project(ctx.h, pc_dtype(u.t), u.ptr(), u.ld, x.ptr(), x.ld, m, N, n, Cm.data_ptr(),
                                 Cm.stride(0), _prec(precision)))
    return u.give(Cm)

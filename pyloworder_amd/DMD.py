"""``pyLOM.DMD`` counterparts: run / frequency_damping / reconstruction_jovanovic.

Signatures and outputs follow pyLOM/DMD/wrapper.py:49-147.  Every step runs
on the device: the mean removal, the tall-skinny SVD of the first n-1
snapshots, the projected operator, its eigen-decomposition (no host eig, the
reference's cupy path falls back to numpy there), the modes with the
Jovanovic amplitudes in the reference's order (one pass over the snapshots)
and the reconstruction.  FP64 arithmetic; float32 snapshots are widened.
"""
from __future__ import annotations

import numpy as np
import torch

from . import vmmath
from ._tensor import Arg, empty
from .cr import cr, cr_start, cr_stop
from .vmmath import _ctx, compute_truncation_residual

__all__ = ["run", "frequency_damping", "reconstruction_jovanovic"]


@cr("DMD.run")
def run(X, r, remove_mean: bool = True):
    """DMD of X (m points x n snapshots) — pyLOM/DMD/wrapper.py:49-117.

    ``r``: truncation as POD.truncate (an int >= 1 keeps r modes, 0 < r < 1 a
    residual, r < 0 an energy fraction).  Returns (muReal, muImag, Phi, bJov):
    the eigenvalues, the modes (m x N complex) and their amplitudes, ordered
    by decreasing |bJov| with each conjugate pair positive-imaginary first.
    """
    x = Arg(X, "X", dtype=torch.float64)
    m, n = x.t.shape
    if n < 3:
        raise ValueError("DMD.run: at least three snapshots")
    cr_start("DMD.temporal_mean", 0)
    Y = vmmath.subtract_mean(x.t, vmmath.temporal_mean(x.t)) if remove_mean else x.t
    cr_stop("DMD.temporal_mean", 0)
    cr_start("DMD.SVD", 0)
    U, S, VT = vmmath.tsqr_svd(Y[:, :-1])
    cr_stop("DMD.SVD", 0)
    N = int(r) if r >= 1 else compute_truncation_residual(S, r)
    N = max(1, min(N, S.shape[0]))
    U, S, VT = U[:, :N], S[:N].contiguous(), VT[:N, :]
    ns = n - 1
    Y1 = Y[:, 1:]
    ctx = _ctx()
    cr_start("DMD.linear_mapping", 0)
    At = empty((N, N), torch.float64, x.t)
    ctx.check(ctx.lib.pc_dmd_operator(ctx.h, U.data_ptr(), U.stride(0), Y1.data_ptr(), Y1.stride(0), m, N, ns,
                                      S.data_ptr(), VT.data_ptr(), VT.stride(0), At.data_ptr(), At.stride(0)))
    cr_stop("DMD.linear_mapping", 0)
    cr_start("DMD.modes", 0)
    wr = empty((N,), torch.float64, x.t)
    wi = empty((N,), torch.float64, x.t)
    W = empty((N, N), torch.complex128, x.t)
    ctx.check(ctx.lib.pc_eig(ctx.h, At.data_ptr(), At.stride(0), N, wr.data_ptr(), wi.data_ptr(), W.data_ptr(),
                             W.stride(0)))
    muR = empty((N,), torch.float64, x.t)
    muI = empty((N,), torch.float64, x.t)
    Phi = empty((m, N), torch.complex128, x.t)
    b = empty((N,), torch.complex128, x.t)
    ctx.check(ctx.lib.pc_dmd_modes(ctx.h, Y1.data_ptr(), Y1.stride(0), m, ns, N, S.data_ptr(), VT.data_ptr(),
                                   VT.stride(0), wr.data_ptr(), wi.data_ptr(), W.data_ptr(), muR.data_ptr(),
                                   muI.data_ptr(), Phi.data_ptr(), Phi.stride(0), b.data_ptr()))
    cr_stop("DMD.modes", 0)
    return x.give(muR), x.give(muI), x.give(Phi), x.give(b)


@cr("DMD.frequency_damping")
def frequency_damping(real, imag, dt):
    """delta = log|mu| / dt, omega = arg(mu) / dt — pyLOM/DMD/wrapper.py:120-130
    (the polar form of pyLOM/vmmath/maths.py:185-200), on the N eigenvalues."""
    if isinstance(real, torch.Tensor):
        mod = torch.sqrt(real * real + imag * imag)
        return torch.log(mod) / dt, torch.atan2(imag, real) / dt
    real, imag = np.asarray(real), np.asarray(imag)
    return np.log(np.sqrt(real * real + imag * imag)) / dt, np.arctan2(imag, real) / dt


@cr("DMD.reconstruction_jovanovic")
def reconstruction_jovanovic(Phi, real, imag, t, bJov):
    """Re(Phi diag(bJov) Vand(t)), Vand[k][j] = mu_k^{t_j} —
    pyLOM/DMD/wrapper.py:141-147 (vandermondeTime, maths.py:225-246): one
    pass writing the m x len(t) reconstruction."""
    p = Arg(Phi, "Phi", dtype=torch.complex128)
    m, N = p.t.shape
    dev = p.t
    mr = Arg(real, "v", dtype=torch.float64).t.contiguous().reshape(-1)
    mi = Arg(imag, "v", dtype=torch.float64).t.contiguous().reshape(-1)
    bb = Arg(bJov, "v", dtype=torch.complex128).t.contiguous().reshape(-1)
    tt = Arg(np.asarray(t, dtype=np.float64) if not isinstance(t, torch.Tensor) else t, "v",
             dtype=torch.float64).t.contiguous().reshape(-1)
    if mr.shape[0] != N or mi.shape[0] != N or bb.shape[0] != N:
        raise ValueError("reconstruction_jovanovic: Phi (m, N), real / imag / bJov (N,) expected")
    nt = tt.shape[0]
    X = empty((m, nt), torch.float64, dev)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_dmd_reconstruct(ctx.h, p.ptr(), p.ld, m, N, mr.data_ptr(), mi.data_ptr(), bb.data_ptr(),
                                         tt.data_ptr(), nt, X.data_ptr(), X.stride(0)))
    return p.give(X)

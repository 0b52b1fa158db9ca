"""``pyLOM.SPOD`` counterpart: run (spectral POD) on the MI355X.

Signature and (L, P, f) outputs follow pyLOM/SPOD/wrapper.py:24-95.  The Welch
blocks (mean removal, Hamming window, DFT, scaling) are one kernel
(pc_spod_blocks); the nf per-frequency SVDs of the M x nBlks complex64 block
matrices are one grouped call (pc_grouped_tsqr_svd) over a single workspace;
the sorted real parts of the modes are one more kernel (pc_spod_modes).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .cr import cr, cr_start, cr_stop
from .vmmath import _ctx, _prec, subtract_mean, temporal_mean

NDFT_MAX = 4096


def _sizes(N: int, nDFT: int, nolap: int):
    """Block geometry exactly as pyLOM/SPOD/wrapper.py:45-52."""
    if nDFT == 0:
        nDFT = int(np.power(2, np.floor(np.log2(N / 10))))
    if nolap == 0:
        nolap = int(np.floor(nDFT / 2))
    nBlks = int(np.floor((N - nolap) / (nDFT - nolap)))
    return nDFT, nolap, nBlks


def _prepare(X, t, nDFT, nolap, remove_mean):
    was_np = isinstance(X, np.ndarray)
    Xt = torch.from_numpy(np.ascontiguousarray(X)).cuda() if was_np else (X if X.is_cuda else X.cuda())
    if Xt.dim() != 2:
        raise ValueError("SPOD: X must be 2-D (points x time samples)")
    M, N = Xt.shape
    tt = np.asarray(t.cpu() if isinstance(t, torch.Tensor) else t)
    dt = tt[1] - tt[0]
    nDFT, nolap, nBlks = _sizes(N, nDFT, nolap)
    if not (2 <= nDFT <= min(N, NDFT_MAX)) or not (0 <= nolap < nDFT) or nBlks < 1:
        raise ValueError(f"SPOD: unsupported block geometry nDFT={nDFT} nolap={nolap} nBlks={nBlks} (N={N})")
    return was_np, Xt, M, N, dt, nDFT, nolap, nBlks


def _blocks(Xt, M, N, nDFT, nolap, nBlks, remove_mean):
    # float64 input: the mean is removed in float64 first (as the reference
    # does before its float64 FFT); the block kernel then reads float32
    rm = 1 if remove_mean else 0
    if Xt.dtype == torch.float64:
        if remove_mean:
            cr_start("SPOD.temporal_mean", 0)
            Xt = subtract_mean(Xt, temporal_mean(Xt))
            cr_stop("SPOD.temporal_mean", 0)
            rm = 0
        Xs = Xt.to(torch.float32)
    elif Xt.dtype == torch.float32:
        Xs = Xt if (Xt.stride(1) == 1 and Xt.stride(0) >= N) else Xt.contiguous()
    else:
        raise TypeError(f"SPOD: X must be float32 or float64, not {Xt.dtype}")
    nf = int(np.ceil(nDFT / 2)) + 1
    Q = torch.empty((nf, M, nBlks), dtype=torch.complex64, device=Xs.device)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_spod_blocks(ctx.h, Xs.data_ptr(), Xs.stride(0), M, N, nDFT, nolap, rm, Q.data_ptr()))
    return Q


@cr("SPOD.blocks")
def blocks(X, t, nDFT: int = 0, nolap: int = 0, remove_mean: bool = True):
    """The Welch block matrices of pyLOM/SPOD/wrapper.py:39-82 (no reference
    function of its own: the loop body at wrapper.py:66-76 plus the
    1/sqrt(nBlks) of wrapper.py:82).  Returns (Q, f, nBlks) with Q a
    (nf, M, nBlks) complex64 device tensor, Q[k] the operand of frequency f[k]."""
    was_np, Xt, M, N, dt, nDFT, nolap, nBlks = _prepare(X, t, nDFT, nolap, remove_mean)
    Q = _blocks(Xt, M, N, nDFT, nolap, nBlks, remove_mean)
    f = np.arange(np.ceil(nDFT / 2) + 1) / dt / nDFT
    return (Q.cpu().numpy() if was_np else Q), f, nBlks


@cr("SPOD.run")
def run(X, t, nDFT: int = 0, nolap: int = 0, remove_mean: bool = True, precision: str = "fp64"):
    """SPOD of X (M points x N time samples) — pyLOM/SPOD/wrapper.py:24-95.

    Returns L (nf, nBlks) modal energies, P (M*nBlks, nf) real parts of the
    SPOD modes (column-major flattening of each frequency's U), and f (nf,)
    frequencies, all sorted by L[:, 0] descending.  The blocks are complex64 as
    in the reference (wrapper.py:42 never selects complex128); ``precision``
    selects the arithmetic of the per-frequency SVD passes.
    """
    was_np, Xt, M, N, dt, nDFT, nolap, nBlks = _prepare(X, t, nDFT, nolap, remove_mean)
    if nBlks > M:
        raise ValueError(f"SPOD.run: nBlks={nBlks} exceeds the number of points M={M}")
    ctx = _ctx()
    f = np.arange(np.ceil(nDFT / 2) + 1) / dt / nDFT
    nf = f.shape[0]
    cr_start("SPOD.fft", 0)
    Q = _blocks(Xt, M, N, nDFT, nolap, nBlks, remove_mean)
    cr_stop("SPOD.fft", 0)
    dev = Q.device

    cr_start("SPOD.SVD", 0)
    U = torch.empty((nf, M, nBlks), dtype=torch.complex64, device=dev)
    S = torch.empty((nf, nBlks), dtype=torch.float32, device=dev)
    V = torch.empty((nf, nBlks, nBlks), dtype=torch.complex64, device=dev)
    probs = (_lib.PCProblem * nf)()
    qb, ub, sb, vb = Q.data_ptr(), U.data_ptr(), S.data_ptr(), V.data_ptr()
    for i in range(nf):
        probs[i] = _lib.PCProblem(qb + 8 * i * M * nBlks, nBlks, M, nBlks, 0, 0, ub + 8 * i * M * nBlks, nBlks,
                                  sb + 4 * i * nBlks, vb + 8 * i * nBlks * nBlks, nBlks)
    ctx.check(ctx.lib.pc_grouped_tsqr_svd(ctx.h, _lib.PC_C64, C.cast(probs, C.c_void_p), nf, _prec(precision)))
    del Q
    cr_stop("SPOD.SVD", 0)

    cr_start("SPOD.sort", 0)
    real = torch.float64 if Xt.dtype == torch.float64 else torch.float32
    L = (S.double() * S.double()).to(real)
    order = np.argsort(L[:, 0].cpu().numpy())[::-1].copy()
    oi = torch.from_numpy(order).to(dev)
    # P[j*M + i, k] = Re U_order[k][i, j]  (column-major flattening, wrapper.py:84)
    P = torch.empty((nBlks * M, nf), dtype=real, device=dev)
    ctx.check(ctx.lib.pc_spod_modes(ctx.h, U.data_ptr(), M, nBlks, nf, oi.data_ptr(), P.data_ptr(),
                                    1 if real == torch.float64 else 0))
    L = L.index_select(0, oi)
    f = f[order]
    cr_stop("SPOD.sort", 0)
    if was_np:
        return L.cpu().numpy(), P.cpu().numpy(), f
    return L, P, torch.from_numpy(f).to(dev)

"""Accuracy metrics used to compare a result with the FP64 oracle (test infra).

* sigma errors: max_k |s_k - s_ref_k| / s_ref_k over the gated modes
  (s_ref_k >= gate * s_ref_1), the same over every numerically non-zero mode,
  and the normwise max_k |s_k - s_ref_k| / s_ref_1;
* the largest principal angle between two k-dimensional column subspaces;
* relative reconstruction RMSE, as pyLOM/vmmath/stats.py:16-34.
"""
from __future__ import annotations

import numpy as np


def sigma_errors(S, Sref, gate=1e-3):
    S = np.asarray(S, dtype=np.float64)
    Sref = np.asarray(Sref, dtype=np.float64)
    k = min(S.shape[0], Sref.shape[0])
    S, Sref = S[:k], Sref[:k]
    s1 = Sref[0] if k else 1.0
    d = np.abs(S - Sref)
    gated = Sref >= gate * s1
    nz = Sref > 1e-10 * s1
    rel = np.where(nz, d / np.where(nz, Sref, 1.0), 0.0)
    return {
        "maxrel_gated": float(rel[gated].max()) if gated.any() else 0.0,
        "n_gated": int(gated.sum()),
        "maxrel_all": float(rel[nz].max()) if nz.any() else 0.0,
        "normwise": float(d.max() / s1) if k else 0.0,
    }


def max_principal_angle(U, Uref, k=None):
    """Largest principal angle (radians) between span(U[:, :k]) and span(Uref[:, :k]).

    Both bases are re-orthonormalised first so that the angle measures the
    subspace only; computed from the sines (accurate for tiny angles).
    """
    U = np.asarray(U)
    Uref = np.asarray(Uref)
    if k is None:
        k = min(U.shape[1], Uref.shape[1])
    if k == 0:
        return 0.0
    A, _ = np.linalg.qr(U[:, :k].astype(np.complex128 if np.iscomplexobj(U) else np.float64))
    B, _ = np.linalg.qr(Uref[:, :k].astype(np.complex128 if np.iscomplexobj(Uref) else np.float64))
    P = A - B @ (np.conj(B).T @ A)
    s = np.linalg.norm(P, 2)
    return float(np.arcsin(min(1.0, s)))


def orthogonality(U, k=None):
    """max |U_k^H U_k - I|."""
    U = np.asarray(U, dtype=np.complex128 if np.iscomplexobj(U) else np.float64)
    if k is None:
        k = U.shape[1]
    G = np.conj(U[:, :k]).T @ U[:, :k]
    return float(np.abs(G - np.eye(k)).max())


def recon_rmse(X, U, S, V):
    """Relative RMSE of X against U diag(S) V (stats.py:16-34 with A = X)."""
    X = np.asarray(X, dtype=np.complex128 if np.iscomplexobj(X) else np.float64)
    R = (np.asarray(U) * np.asarray(S)) @ np.asarray(V)
    d = X - R
    return float(np.sqrt(np.sum(np.abs(d) ** 2) / np.sum(np.abs(X) ** 2)))


def align_columns(U, Uref):
    """Per-column unit phase c_k with U[:, k] * c_k closest to Uref[:, k]."""
    U = np.asarray(U)
    Uref = np.asarray(Uref)
    k = min(U.shape[1], Uref.shape[1])
    dots = np.sum(np.conj(U[:, :k]) * Uref[:, :k], axis=0)
    mag = np.abs(dots)
    c = np.where(mag > 0, dots / np.where(mag > 0, mag, 1), 1)
    return c


def report(X, U, S, V, Sref, Uref, gate=1e-3, k=None):
    e = sigma_errors(S, Sref, gate)
    kk = e["n_gated"] if k is None else k
    e["angle"] = max_principal_angle(U, Uref, kk)
    e["orth"] = orthogonality(U, kk)
    e["rmse"] = recon_rmse(X, U, S, V) if X is not None else None
    return e

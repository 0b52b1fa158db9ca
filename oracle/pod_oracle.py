"""numpy restatement of the reference POD core (test oracle, CPU only).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The MPI ranks of the reference are simulated:
a matrix distributed over P ranks is a Python list of P row blocks split
exactly as pyLOM/utils/parall.py:24-48 (worksplit) splits rows, and the
butterfly messages of the TSQR tree are exchanged level by level in-process.
"""
from __future__ import annotations

import time

import numpy as np
import scipy.fft

__all__ = [
    "next_power_of_2", "worksplit", "split_rows", "qr", "svd", "tsqr",
    "tsqr_svd", "randomized_qr", "randomized_svd", "reference_omega",
    "temporal_mean", "subtract_mean", "temporal_variance", "norm_variance",
    "transpose", "matmul", "matmulp", "vecmat", "compute_truncation_residual",
    "energy", "RMSE", "pod_run", "pod_truncate", "pod_reconstruct",
    "pod_project", "hammwin", "spod_blocks", "spod_run", "init_qr_streaming", "update_qr_streaming",
    "eigen", "vandermonde", "vandermonde_time", "dmd_order_modes", "dmd_run", "dmd_frequency_damping",
    "dmd_reconstruction_jovanovic", "dataset_x",
]


# ---------------------------------------------------------------------------
# rank bookkeeping
# ---------------------------------------------------------------------------
def next_power_of_2(n: int) -> int:
    """pyLOM/vmmath/svd.py:17-25."""
    p = 1
    if n and not (n & (n - 1)):
        return n
    while p < n:
        p <<= 1
    return p


def worksplit(istart: int, iend: int, who: int, nworkers: int):
    """Row range owned by rank ``who`` — pyLOM/utils/parall.py:24-48."""
    irange = iend - istart
    if nworkers < irange:
        per = irange // nworkers
        lo = istart + who * per
        hi = lo + per
        rem = irange - per * nworkers
        if rem > who:
            lo += who
            hi += who + 1
        else:
            lo += rem
            hi += rem
        return lo, hi
    lo = who if who < iend else iend
    hi = who + 1 if who < iend else iend
    return lo, hi


def split_rows(A: np.ndarray, nranks: int):
    """Distribute the rows of A over ``nranks`` simulated ranks."""
    m = A.shape[0]
    out = []
    for r in range(nranks):
        lo, hi = worksplit(0, m, r, nranks)
        out.append(A[lo:hi])
    return out


# ---------------------------------------------------------------------------
# dense kernels (numpy LAPACK, as the reference's Python path uses)
# ---------------------------------------------------------------------------
def qr(A):
    """Reduced QR — pyLOM/vmmath/svd.py:28-36 (numpy geqrf + orgqr)."""
    return np.linalg.qr(A)


def svd(A):
    """Thin SVD, V returned as V^H — pyLOM/vmmath/svd.py:38-47."""
    return np.linalg.svd(A, full_matrices=False)


def tsqr(blocks):
    """Demmel TSQR with the reference's butterfly tree.

    pyLOM/vmmath/svd.py:49-118 (compiled twin: pyLOM/vmmath/src/svd.c:565-676).
    ``blocks`` is the list of per-rank row blocks A_i (each with m_i >= n).
    Returns (list of Q_i, R) where R is the rank-0 R after the down-sweep.
    Senders and receivers of one level are matched, so executing every rank's
    send before any receive of the same level reproduces the MPI schedule.
    """
    P = len(blocks)
    n = blocks[0].shape[1]
    dt = np.result_type(*[b.dtype for b in blocks])
    # svd.py:60 leaf QR on every rank
    Q1 = []
    R = []
    for b in blocks:
        q, r = qr(b)
        if r.shape[0] != n:
            raise ValueError("tsqr needs m_i >= n on every rank (svd.py:69)")
        Q1.append(q)
        R.append(r)
    nxt = next_power_of_2(P)
    nlev = int(np.log2(nxt))
    QW = [np.eye(n, dtype=dt) for _ in range(P)]
    C = [np.zeros((2 * n, n), dt) for _ in range(P)]
    Q2l = [np.zeros((2 * n * nlev, n), dt) for _ in range(P)]
    # up-sweep svd.py:67-84
    blevel = 1
    for ilev in range(nlev):
        mail = {}
        for rk in range(P):
            C[rk][:n, :] = R[rk]
            pr = rk ^ blevel
            if rk & blevel and pr < P:
                mail[pr] = R[rk].copy()
        for rk in range(P):
            pr = rk ^ blevel
            if not (rk & blevel) and pr < P:
                C[rk][n:, :] = mail[rk]
                q2, R[rk] = qr(C[rk])
                Q2l[rk][2 * n * ilev:2 * n * ilev + 2 * n, :] = q2
        blevel <<= 1
    # down-sweep svd.py:87-115
    if P > 1:
        blevel = 1 << (nlev - 1)
        mask = blevel - 1
    for ilev in reversed(range(nlev)):
        mail = {}
        active = [rk for rk in range(P) if (rk & mask) == 0]
        q2i = {}
        for rk in active:
            q2i[rk] = Q2l[rk][2 * n * ilev:2 * n * ilev + 2 * n, :] @ QW[rk]
        for rk in active:  # senders first
            pr = rk ^ blevel
            if not (rk & blevel) and pr < P:
                msg = np.empty((2 * n, n), dt)
                msg[:n, :] = R[rk]
                msg[n:, :] = q2i[rk][n:, :]
                QW[rk] = q2i[rk][:n, :]
                mail[pr] = msg
        for rk in active:
            pr = rk ^ blevel
            if rk & blevel and pr < P:
                msg = mail[rk]
                R[rk] = msg[:n, :]
                QW[rk] = msg[n:, :]
        blevel >>= 1
        mask >>= 1
    # svd.py:117
    Qi = [Q1[rk] @ QW[rk] for rk in range(P)]
    return Qi, R[0]


def tsqr_svd(blocks):
    """pyLOM/vmmath/svd.py:227-252: tsqr, svd(R), U_i = Q_i U_R."""
    Qi, R = tsqr(blocks)
    Ur, S, V = svd(R)
    return [q @ Ur for q in Qi], S, V


def reference_omega(n: int, r: int, seed: int, dtype):
    """Omega of pyLOM/vmmath/svd.py:131-133 (numpy legacy MT19937 stream)."""
    seed = int(time.time()) if seed < 0 else seed
    np.random.seed(seed=seed)
    return np.random.rand(n, r).astype(dtype)


def randomized_qr(blocks, r: int, q: int, seed: int = -1, omega=None):
    """pyLOM/vmmath/svd.py:120-144 on simulated ranks.

    Every rank seeds identically, so every rank draws the same Omega.
    ``matmulp`` (maths.py:92-110) is the sum over ranks of the local products.
    """
    n = blocks[0].shape[1]
    dt = blocks[0].dtype
    if omega is None:
        omega = reference_omega(n, r, seed, dt)
    Y = [b @ omega for b in blocks]
    for _ in range(q):
        Qi, _ = tsqr(Y)
        Q2 = matmulp([b.T for b in blocks], Qi)
        Y = [b @ Q2 for b in blocks]
    Qi, _ = tsqr(Y)
    B = matmulp([qq.T for qq in Qi], blocks)
    return Qi, B


def _qr_signs(blocks, positive):
    """tsqr of the blocks, LAPACK's signs as the reference
    (pyLOM/vmmath/svd.py:49-118 via np.linalg.qr); with ``positive`` the
    column signs are flipped so that diag(R) >= 0, a second convention that
    the oracle tests use to show the sketch depends on the choice."""
    Qi, R = tsqr(blocks)
    if positive:
        d = np.where(np.diag(R) < 0, -1.0, 1.0)
        Qi = [qq * d for qq in Qi]
    return Qi


def init_qr_streaming(blocks, r: int, q: int, seed=None, omega=None, positive=False):
    """pyLOM/vmmath/svd.py:175-197: randomized_qr that also returns the last
    sketch Y (per rank).  Returns (Qi, B, Yi)."""
    n = blocks[0].shape[1]
    dt = blocks[0].dtype
    if omega is None:
        omega = reference_omega(n, r, int(time.time()) if seed is None else seed, dt)
    Y = [b @ omega for b in blocks]
    for _ in range(q):
        Qi = _qr_signs(Y, positive)
        Q2 = matmulp([b.T for b in blocks], Qi)
        Y = [b @ Q2 for b in blocks]
    Qi = _qr_signs(Y, positive)
    B = matmulp([qq.T for qq in Qi], blocks)
    return Qi, B, Y


def update_qr_streaming(blocks, Q1, B1, Yo, r: int, q: int, omega=None, positive=False):
    """pyLOM/vmmath/svd.py:199-225: fold a new batch of snapshots into the
    sketch.  Yn = (A A^T)^q A Omega with Omega from numpy's global stream
    (unseeded, svd.py:212); Yo += Yn (in place, as the reference); Q2 = tsqr(Yo);
    B2 = [Q2^T Q1 B1, Q2^T A].  Returns (Q2, B2, Yo)."""
    n = blocks[0].shape[1]
    dt = blocks[0].dtype
    if omega is None:
        omega = np.random.rand(n, r).astype(dt)
    Yn = [b @ omega for b in blocks]
    for _ in range(q):
        Qp = _qr_signs(Yn, positive)
        O2 = matmulp([b.T for b in blocks], Qp)
        Yn = [b @ O2 for b in blocks]
    for yo, yn in zip(Yo, Yn):
        yo += yn
    Q2 = _qr_signs(Yo, positive)
    Q2Q1 = matmulp([qq.T for qq in Q2], Q1)
    B2o = Q2Q1 @ B1
    B2n = matmulp([qq.T for qq in Q2], blocks)
    return Q2, np.hstack((B2o, B2n)), Yo


def randomized_svd(blocks, r: int, q: int, seed: int = -1, omega=None):
    """pyLOM/vmmath/svd.py:254-273."""
    Qi, B = randomized_qr(blocks, r, q, seed=seed, omega=omega)
    Ur, S, V = svd(B)
    return [qq @ Ur for qq in Qi], S, V


# ---------------------------------------------------------------------------
# averaging.py, maths.py
# ---------------------------------------------------------------------------
def temporal_mean(X):
    """pyLOM/vmmath/averaging.py:17-29 (row mean over snapshots)."""
    return np.mean(X, axis=1)


def subtract_mean(X, X_mean):
    """pyLOM/vmmath/averaging.py:31-44."""
    return X - np.tile(X_mean, (X.shape[1], 1)).T


def temporal_variance(X, X_mean):
    """pyLOM/vmmath/averaging.py:46-59 (population variance, keepdims)."""
    return np.var(X, axis=1, keepdims=True)


def norm_variance(X, X_mean, X_var):
    """pyLOM/vmmath/averaging.py:61-75 — divides by the VARIANCE."""
    return subtract_mean(X, X_mean) / X_var


def transpose(A):
    """pyLOM/vmmath/maths.py:17-29."""
    return np.transpose(A)


def matmul(A, B):
    """pyLOM/vmmath/maths.py:76-90."""
    return np.matmul(A, B)


def matmulp(As, Bs):
    """pyLOM/vmmath/maths.py:92-110: local products summed over ranks."""
    out = None
    for a, b in zip(As, Bs):
        c = np.matmul(a, b)
        out = c if out is None else out + c
    return out


def vecmat(v, A):
    """pyLOM/vmmath/maths.py:112-129: C[i,:] = v[i] * A[i,:]."""
    C = np.zeros_like(A)
    for ii in range(v.shape[0]):
        C[ii, :] = v[ii] * A[ii, :]
    return C


# ---------------------------------------------------------------------------
# truncation.py, stats.py
# ---------------------------------------------------------------------------
def compute_truncation_residual(S, r):
    """pyLOM/vmmath/truncation.py:17-39 (r<0 returns i+1, as the Python path)."""
    N = 0
    if r > 0:
        normS = np.linalg.norm(S[0:], 2)
        for ii in range(S.shape[0]):
            acc = np.linalg.norm(S[ii:], 2) / normS
            if acc < r:
                break
            N += 1
    else:
        r = abs(r)
        normS = np.sum(S[0:])
        acc = 0
        for ii in range(S.shape[0]):
            acc += S[ii] / normS
            N += 1
            if acc > r:
                break
    return N


def energy(original, rec):
    """pyLOM/vmmath/truncation.py:41-60."""
    num = np.sum((original - rec) ** 2)
    den = np.sum(original ** 2)
    return 1 - num / den


def RMSE(A, B, relative=True):
    """pyLOM/vmmath/stats.py:16-34."""
    diff = A - B
    s1 = np.sum(diff * diff)
    s2 = np.sum(A * A) if relative else np.prod(np.array(A.shape))
    return np.sqrt(s1 / s2)


# ---------------------------------------------------------------------------
# POD
# ---------------------------------------------------------------------------
def pod_run(X, remove_mean=True, divide_variance=False, randomized=False,
            r=1, q=3, seed=-1, nranks=1, omega=None):
    """pyLOM/POD/wrapper.py:15-51 over ``nranks`` simulated ranks.

    Returns (U, S, V) with U stacked over the ranks and V = V^H.
    """
    if remove_mean:
        mu = temporal_mean(X)
        if divide_variance:
            Y = norm_variance(X, mu, temporal_variance(X, mu))
        else:
            Y = subtract_mean(X, mu)
    else:
        Y = X.copy()
    blocks = split_rows(Y, nranks)
    if randomized:
        Ui, S, V = randomized_svd(blocks, r, q, seed=seed, omega=omega)
    else:
        Ui, S, V = tsqr_svd(blocks)
    return np.concatenate(Ui, axis=0), S, V


def pod_truncate(U, S, V, r=1e-8):
    """pyLOM/POD/wrapper.py:54-82."""
    N = int(r) if r >= 1 else compute_truncation_residual(S, r)
    return U[:, :N], S[:N], V[:N, :]


def pod_reconstruct(U, S, V):
    """pyLOM/POD/wrapper.py:85-103: U @ vecmat(S, V)."""
    return matmul(U, vecmat(S, V))


def pod_project(U, X):
    """Projection coefficients U^H X (idiom of pyLOM/DMD/wrapper.py:85)."""
    return np.conj(U).T @ X


# ---------------------------------------------------------------------------
# SPOD
# ---------------------------------------------------------------------------
def hammwin(N):
    """pyLOM/vmmath/fft.py:16-27."""
    return np.transpose(0.54 - 0.46 * np.cos(2 * np.pi * np.arange(N) / (N - 1)))


def spod_blocks(X, t, nDFT=0, nolap=0, remove_mean=True):
    """Welch blocks of pyLOM/SPOD/wrapper.py:39-76.

    Returns (Q, f, nBlks) with Q of shape (nf, M, nBlks), already divided by
    sqrt(nBlks) as at wrapper.py:82.  complex64 blocks (quirk: wrapper.py:42
    never selects complex128 because a dtype object is never np.double).
    """
    M, N = X.shape
    dt = t[1] - t[0]
    cdtype = np.complex64
    if nDFT == 0:
        nDFT = int(np.power(2, np.floor(np.log2(N / 10))))
    window = hammwin(nDFT)
    if nolap == 0:
        nolap = int(np.floor(nDFT / 2))
    nBlks = int(np.floor((N - nolap) / (nDFT - nolap)))
    winWeight = 1 / np.mean(window)
    Y = subtract_mean(X, temporal_mean(X)) if remove_mean else X.copy()
    f = np.arange(np.ceil(nDFT / 2) + 1) / dt / nDFT
    nf = f.shape[0]
    Q = np.zeros((nf, M, nBlks), cdtype)
    for iblk in range(nBlks):
        i0 = iblk * (nDFT - nolap)
        ix = np.arange(nDFT) + i0
        Xf = Y[:, ix].copy() * window
        qk = (winWeight / nDFT) * scipy.fft.fft2(Xf, axes=(-1,))[:, :nf]
        qk[:, 1:-1] *= 2
        Q[:, :, iblk] = qk.T
    return Q / np.sqrt(nBlks), f, nBlks


def spod_run(X, t, nDFT=0, nolap=0, remove_mean=True):
    """pyLOM/SPOD/wrapper.py:24-95: returns (L, P, f) sorted by L[:,0] desc."""
    M = X.shape[0]
    Q, f, nBlks = spod_blocks(X, t, nDFT, nolap, remove_mean)
    nf = f.shape[0]
    L = np.zeros((nf, nBlks), X.dtype)
    P = np.zeros((M * nBlks, nf), X.dtype)
    for ifreq in range(nf):
        U, S, V = tsqr_svd([Q[ifreq]])
        U = U[0]
        P[:, ifreq] = np.real(U.reshape((M * nBlks), order="F"))
        L[ifreq, :] = np.abs(S * S)
    order = np.argsort(L[:, 0])[::-1]
    return L[order, :], P[:, order], f[order]


# ---------------------------------------------------------------------------
# DMD  (pyLOM/DMD/wrapper.py, the helpers of pyLOM/vmmath/maths.py)
# ---------------------------------------------------------------------------
def eigen(A):
    """(real, imag, vecs) of A — pyLOM/vmmath/maths.py:161-183 (numpy.linalg.eig)."""
    w, vecs = np.linalg.eig(A)
    return np.real(w).copy(), np.imag(w).copy(), vecs


def vandermonde(real, imag, m: int, n: int):
    """Vand[:, j] = (real + i imag)**j, j < n — pyLOM/vmmath/maths.py:203-222."""
    lam = real + 1j * imag
    V = np.zeros((m, n), dtype=np.complex128)
    for icol in range(n):
        V[:, icol] = lam ** icol
    return V


def vandermonde_time(real, imag, m: int, time):
    """Vand[:, j] = (real + i imag)**t_j — pyLOM/vmmath/maths.py:225-246."""
    lam = real + 1j * imag
    V = np.zeros((m, len(time)), dtype=np.complex128)
    for it, t in enumerate(time):
        V[:, it] = lam ** t
    return V


def dmd_order_modes(muReal, muImag, Phi, bJov):
    """pyLOM/DMD/wrapper.py:18-46: modes by decreasing |b|, then inside each
    conjugate pair the positive imaginary part first (imaginary parts of the
    pair's entries exchanged as the reference does)."""
    order = np.flip(np.abs(bJov).argsort(kind="stable"))
    muReal, muImag, Phi, bJov = muReal[order].copy(), muImag[order].copy(), Phi[:, order].copy(), bJov[order].copy()
    skip = False
    for ii in range(muImag.shape[0]):
        if skip:
            skip = False
            continue
        if muImag[ii] < 0:
            muImag[ii] = muImag[ii + 1]
            muImag[ii + 1] = -muImag[ii]
            bJov.imag[ii] = bJov.imag[ii + 1]
            bJov.imag[ii + 1] = -bJov.imag[ii]
            Phi.imag[:, ii] = Phi.imag[:, ii + 1]
            Phi.imag[:, ii + 1] = -Phi.imag[:, ii + 1]
            skip = True
            continue
        if muImag[ii] > 0:
            skip = True
    return muReal, muImag, Phi, bJov


def dmd_run(X, r, remove_mean: bool = True):
    """pyLOM/DMD/wrapper.py:49-117 on one rank: SVD of the first n-1
    snapshots, truncation, the projected operator
    Atilde = U^T Y[:, 1:] VT^T diag(1/S), its eigen-decomposition, the modes
    Phi = Y[:, 1:] VT^T diag(1/S) w / mu and the amplitudes of Jovanovic et
    al. (2014), ordered by amplitude.  Returns (muReal, muImag, Phi, bJov)."""
    X = np.asarray(X, dtype=np.float64)
    Y = subtract_mean(X, temporal_mean(X)) if remove_mean else X.copy()
    U, S, VT = np.linalg.svd(Y[:, :-1], full_matrices=False)
    N = int(r) if r >= 1 else compute_truncation_residual(S, r)
    U, S, VT = U[:, :N], S[:N], VT[:N, :]
    aux1 = U.T @ Y[:, 1:]
    aux2 = ((1.0 / S)[:, None] * VT).T
    Atilde = aux1 @ aux2
    muReal, muImag, w = eigen(Atilde)
    Phi = (Y[:, 1:] @ VT.T @ np.diag(1.0 / S) @ w) / (muReal + muImag * 1j)
    Vand = vandermonde(muReal, muImag, muReal.shape[0], Y.shape[1] - 1)
    P = (w.conj().T @ w) * np.conj(Vand @ Vand.conj().T)
    Pl = np.linalg.cholesky(P)
    G = np.diag(S) @ VT
    q = np.conj(np.diag(Vand @ G.conj().T @ w))
    bJov = np.linalg.inv(Pl.conj().T) @ (np.linalg.inv(Pl) @ q)
    return dmd_order_modes(muReal, muImag, Phi, bJov)


def dmd_frequency_damping(real, imag, dt):
    """pyLOM/DMD/wrapper.py:120-130: delta = log|mu| / dt, omega = arg(mu) / dt."""
    mod = np.sqrt(real * real + imag * imag)
    return np.log(mod) / dt, np.arctan2(imag, real) / dt


def dmd_reconstruction_jovanovic(Phi, real, imag, t, bJov):
    """pyLOM/DMD/wrapper.py:141-147: Re(Phi diag(b) Vand(t))."""
    V = vandermonde_time(real, imag, real.shape[0], t)
    return (Phi @ (np.diag(bJov) @ V)).real


def dataset_x(fields: dict, npoints: int, names=None):
    """pyLOM/dataset.py:252-292 for one temporal dimension: fields[name] =
    (ndim, value (ndim * npoints, ...)); rows of X interleaved per point."""
    names = list(names) if names else list(fields)
    nf = sum(fields[n][0] for n in names)
    v0 = np.asarray(fields[names[0]][1])
    X = np.zeros((nf * npoints,) + v0.shape[1:], dtype=np.float64)
    ifield = 0
    for name in names:
        ndim, value = fields[name]
        value = np.asarray(value)
        for idim in range(ndim):
            X[ifield:nf * npoints:nf] = value[idim:ndim * npoints:ndim]
            ifield += 1
    return X

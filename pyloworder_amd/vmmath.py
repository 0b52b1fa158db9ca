"""``pyLOM.math`` counterparts of the POD core, running on the MI355X.

Same names, argument meaning and return values as the reference's
pyLOM/vmmath (svd.py, maths.py, averaging.py, truncation.py, stats.py); each
function cites the reference function it replaces.  Inputs may be numpy arrays
(results come back as numpy) or torch tensors (results stay on the device).
One extension keyword, ``precision``, selects the arithmetic of the streaming
passes: ``"fp64"`` (default, FP64 MFMA whatever the storage) or ``"mixed"``
(float32/complex64 storage: fp16 hi/lo split MFMA, FP64 small steps).

Errors raise ``pyloworder_amd.LinAlgError`` / ``PodCoreError`` instead of the
reference's MPI_Abort (pyLOM/utils/errors.py:22-28).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from ._tensor import Arg, empty, pc_dtype, real_dtype
from .cr import cr

__all__ = [
    "qr", "svd", "tsqr", "tsqr_svd", "tsqr_svd_grouped", "GroupedGraph", "randomized_qr", "randomized_svd", "init_qr_streaming", "update_qr_streaming", "matmul", "matmulp",
    "vecmat", "transpose", "temporal_mean", "subtract_mean", "temporal_variance", "norm_variance",
    "compute_truncation_residual", "RMSE", "energy", "reference_omega", "eigen",
]

_PREC = {"fp64": _lib.PC_PREC_FP64, "mixed": _lib.PC_PREC_MIXED}


def _prec(precision):
    try:
        return _PREC[precision]
    except KeyError:
        raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}") from None


def _ctx():
    return _lib.context()


# ---------------------------------------------------------------------------
# SVD / QR  (pyLOM/vmmath/svd.py)
# ---------------------------------------------------------------------------
@cr("math.tsqr_svd")
def tsqr_svd(Ai, precision: str = "fp64", center: int = 0):
    """U, S, V of a tall-skinny matrix — pyLOM/vmmath/svd.py:227-252.

    Ui (m, n) left singular vectors, S (n,) descending, V (n, n) = V^H (rows are
    the right singular vectors).  ``center`` = 1 removes the row mean (2 also
    divides by the row variance) inside the passes, as POD.run does before
    calling tsqr_svd (pyLOM/POD/wrapper.py:33-45).
    """
    a = Arg(Ai, "Ai")
    m, n = a.t.shape
    U = empty((m, n), a.t.dtype, a.t)
    S = empty((n,), real_dtype(a.t.dtype), a.t)
    V = empty((n, n), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_tsqr_svd(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, int(center), U.data_ptr(), U.stride(0),
                                  S.data_ptr(), V.data_ptr(), V.stride(0), _prec(precision)))
    return a.give(U), a.give(S), a.give(V)


def _group_outputs(args, dt):
    """Output tensors (U, S, V) per problem.  Problems of one shape get views of
    three stacked tensors, so S and V sit at constant strides: the library then
    runs the small stages of all problems as single launches."""
    shapes = {tuple(a.t.shape) for a in args}
    if len(shapes) == 1:
        m, n = args[0].t.shape
        k = len(args)
        Ub = empty((k, m, n), dt, args[0].t)
        Sb = empty((k, n), real_dtype(dt), args[0].t)
        Vb = empty((k, n, n), dt, args[0].t)
        return [(Ub[i], Sb[i], Vb[i]) for i in range(k)]
    out = []
    for a in args:
        m, n = a.t.shape
        out.append((empty((m, n), dt, a.t), empty((n,), real_dtype(dt), a.t), empty((n, n), dt, a.t)))
    return out


@cr("math.tsqr_svd_grouped")
def tsqr_svd_grouped(As, precision: str = "fp64", center: int = 0):
    """tsqr_svd of every matrix of ``As`` in one grouped call (pc_grouped_tsqr_svd).

    No single reference function: it replaces the per-frequency loop of
    pyLOM/SPOD/wrapper.py:80-86 (one tsqr_svd per block) and any loop of
    independent PODs.  ``As`` is a list of 2-D arrays/tensors of one dtype, or a
    3-D tensor (count, m, n).  Returns lists (Us, Ss, Vs), each entry identical
    to ``tsqr_svd`` of that matrix alone.
    """
    if isinstance(As, torch.Tensor) and As.dim() == 3:
        args = [Arg(As[i], "Ai") for i in range(As.shape[0])]
    else:
        args = [Arg(A, "Ai") for A in As]
    if not args:
        return [], [], []
    dt = args[0].t.dtype
    for a in args:
        if a.t.dtype != dt:
            raise TypeError("tsqr_svd_grouped: all matrices must share one dtype")
    Us, Ss, Vs = [], [], []
    probs = (_lib.PCProblem * len(args))()
    outs = _group_outputs(args, dt)
    for i, a in enumerate(args):
        m, n = a.t.shape
        U, S, V = outs[i]
        Us.append(U)
        Ss.append(S)
        Vs.append(V)
        probs[i] = _lib.PCProblem(a.ptr(), a.ld, m, n, int(center), 0, U.data_ptr(), U.stride(0), S.data_ptr(),
                                  V.data_ptr(), V.stride(0))
    ctx = _ctx()
    ctx.check(ctx.lib.pc_grouped_tsqr_svd(ctx.h, pc_dtype(args[0].t), C.cast(probs, C.c_void_p), len(args),
                                          _prec(precision)))
    return ([a.give(U) for a, U in zip(args, Us)], [a.give(S) for a, S in zip(args, Ss)],
            [a.give(V) for a, V in zip(args, Vs)])


class GroupedGraph:
    """A grouped tsqr_svd recorded once into a HIP graph and replayed.

    No reference counterpart: for repeated solves of fixed shapes (SPOD on
    new snapshots written into the same buffers, repeated POD of a stream).
    ``As``: device tensors (kept referenced; their contents may change
    between launches).  ``launch()`` replays the recorded solves on the
    current stream and returns the output lists (Us, Ss, Vs), the same
    tensors each time, identical to ``tsqr_svd_grouped`` of the current data.
    """

    def __init__(self, As, precision: str = "fp64", center: int = 0):
        if isinstance(As, torch.Tensor) and As.dim() == 3:
            As = [As[i] for i in range(As.shape[0])]
        self.args = [Arg(A, "Ai") for A in As]
        if not self.args or any(a.was_numpy for a in self.args):
            raise TypeError("GroupedGraph records device tensors (torch, on the GPU)")
        dt = self.args[0].t.dtype
        self.Us, self.Ss, self.Vs = [], [], []
        self._probs = (_lib.PCProblem * len(self.args))()
        for a in self.args:
            if a.t.dtype != dt:
                raise TypeError("GroupedGraph: all matrices must share one dtype")
        outs = _group_outputs(self.args, dt)
        for i, a in enumerate(self.args):
            m, n = a.t.shape
            U, S, V = outs[i]
            self.Us.append(U)
            self.Ss.append(S)
            self.Vs.append(V)
            self._probs[i] = _lib.PCProblem(a.ptr(), a.ld, m, n, int(center), 0, U.data_ptr(), U.stride(0),
                                            S.data_ptr(), V.data_ptr(), V.stride(0))
        ctx = _ctx()
        self._lib = ctx.lib
        h = C.c_void_p()
        ctx.check(ctx.lib.pc_grouped_graph_create(ctx.h, pc_dtype(self.args[0].t), C.cast(self._probs, C.c_void_p),
                                                  len(self.args), _prec(precision), C.byref(h)))
        self.h = h

    def launch(self):
        ctx = _ctx()
        ctx.check(ctx.lib.pc_grouped_graph_launch(ctx.h, self.h))
        return self.Us, self.Ss, self.Vs

    def close(self):
        if getattr(self, "h", None):
            self._lib.pc_graph_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


@cr("math.tsqr")
def tsqr(Ai, precision: str = "fp64"):
    """Qi (m, n), R (n, n) — pyLOM/vmmath/svd.py:49-118.  The signs are LAPACK's
    (?geqrf, as numpy.linalg.qr in the reference's one-rank tsqr): R_kk is
    real with the sign -sign(Re alpha_k) of the Householder step."""
    a = Arg(Ai, "Ai")
    m, n = a.t.shape
    Q = empty((m, n), a.t.dtype, a.t)
    R = empty((n, n), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_tsqr(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, Q.data_ptr(), Q.stride(0), R.data_ptr(),
                              R.stride(0), _prec(precision)))
    return a.give(Q), a.give(R)


@cr("math.qr")
def qr(A, precision: str = "fp64"):
    """Reduced QR — pyLOM/vmmath/svd.py:28-36 (tall or square matrices)."""
    a = Arg(A, "A")
    m, n = a.t.shape
    if m < n:
        raise ValueError("qr: this build factors tall or square matrices (m >= n)")
    return tsqr(A, precision=precision)


@cr("math.svd")
def svd(A, method: str = "gesdd"):
    """Thin SVD (U, S, V=V^H) — pyLOM/vmmath/svd.py:38-47.  ``method`` is accepted
    for signature compatibility; one-sided Jacobi is used for every matrix."""
    a = Arg(A, "A")
    m, n = a.t.shape
    k = min(m, n)
    U = empty((m, k), a.t.dtype, a.t)
    S = empty((k,), real_dtype(a.t.dtype), a.t)
    V = empty((k, n), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_svd(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, U.data_ptr(), U.stride(0), S.data_ptr(),
                             V.data_ptr(), V.stride(0)))
    return a.give(U), a.give(S), a.give(V)


def reference_omega(n: int, r: int, seed: int, dtype) -> np.ndarray:
    """The reference's sketch matrix: np.random.seed(seed); rand(n, r) —
    pyLOM/vmmath/svd.py:131-133 (numpy legacy MT19937, frozen by numpy)."""
    seed = int(time.time()) if seed < 0 else seed
    rs = np.random.RandomState(seed)
    return rs.random_sample((n, r)).astype(dtype)


def _omega_tensor(a: Arg, r: int, omega):
    """The caller's sketch matrix as a contiguous device tensor, or None: the
    library then draws the reference's Omega on the device from the seed."""
    if omega is None:
        return None
    n = a.t.shape[1]
    o = Arg(omega, "omega", dtype=a.t.dtype).t
    if tuple(o.shape) != (n, r):
        raise ValueError(f"omega must be ({n}, {r})")
    return o.contiguous()


def _seed(seed: int) -> int:
    """seed < 0 draws a time-based seed (pyLOM/vmmath/svd.py:131, 264); numpy's
    legacy seeding takes 0 <= seed < 2**32."""
    seed = int(time.time()) if seed < 0 else int(seed)
    if seed > 0xFFFFFFFF:
        raise ValueError("Seed must be between 0 and 2**32 - 1")
    return seed


@cr("math.randomized_qr")
def randomized_qr(Ai, r: int, q: int, seed: int = -1, omega=None, precision: str = "fp64"):
    """Qi (m, r), B (r, n) — pyLOM/vmmath/svd.py:120-144.  Without ``omega`` the
    library draws the reference's Omega (numpy's seeded MT19937 stream) on
    the device."""
    a = Arg(Ai, "Ai")
    m, n = a.t.shape
    om = _omega_tensor(a, r, omega)
    Q = empty((m, r), a.t.dtype, a.t)
    B = empty((r, n), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_randomized_qr(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, r, q, _seed(seed),
                                       None if om is None else om.data_ptr(), Q.data_ptr(), Q.stride(0), B.data_ptr(),
                                       B.stride(0), _prec(precision)))
    return a.give(Q), a.give(B)


@cr("math.randomized_svd")
def randomized_svd(Ai, r: int, q: int, seed: int = -1, omega=None, precision: str = "fp64"):
    """Ui (m, r), S (r,), V (r, n) — pyLOM/vmmath/svd.py:254-273."""
    a = Arg(Ai, "Ai")
    m, n = a.t.shape
    om = _omega_tensor(a, r, omega)
    U = empty((m, r), a.t.dtype, a.t)
    S = empty((r,), real_dtype(a.t.dtype), a.t)
    V = empty((r, n), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_randomized_svd(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, r, q, _seed(seed),
                                        None if om is None else om.data_ptr(), U.data_ptr(), U.stride(0),
                                        S.data_ptr(), V.data_ptr(), V.stride(0), _prec(precision)))
    return a.give(U), a.give(S), a.give(V)


def _dev(x, dtype=None):
    """(device tensor, was_numpy): numpy inputs are copied to the GPU."""
    if isinstance(x, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(x)).cuda()
        return (t.to(dtype) if dtype is not None else t), True
    t = x if x.is_cuda else x.cuda()
    return (t.to(dtype) if dtype is not None and t.dtype != dtype else t), False


def _gemm_into(opa, A, opb, B, C, beta, precision):
    """C = op(A) op(B) + beta C on the device (pc_gemm, FP64 MFMA whatever
    ``precision`` asks of the QRs); row-major operands, op 1 = transpose.
    Shapes are taken from C and op(A)."""
    M, N = C.shape
    K = A.shape[0] if opa else A.shape[1]
    ctx = _ctx()
    ctx.check(ctx.lib.pc_gemm(ctx.h, pc_dtype(C), opa, opb, M, N, K, 1.0, A.data_ptr(), A.stride(0), B.data_ptr(),
                              B.stride(0), float(beta), C.data_ptr(), C.stride(0), _lib.PC_PREC_FP64))
    return C


def _tsqr_q(Y, precision):
    m, r = Y.shape
    Q = empty((m, r), Y.dtype, Y)
    R = empty((r, r), Y.dtype, Y)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_tsqr(ctx.h, pc_dtype(Y), Y.data_ptr(), Y.stride(0), m, r, Q.data_ptr(), Q.stride(0),
                              R.data_ptr(), R.stride(0), _prec(precision)))
    return Q


def _power_sketch(A, omega, q, precision):
    """Y = (A A^T)^q A Omega (pyLOM/vmmath/svd.py:186-191, 212-216)."""
    m, n = A.shape
    r = omega.shape[1]
    Y = _gemm_into(0, A, 0, omega, empty((m, r), A.dtype, A), 0.0, precision)
    for _ in range(q):
        Qp = _tsqr_q(Y, precision)
        O2 = _gemm_into(1, A, 0, Qp, empty((n, r), A.dtype, A), 0.0, precision)
        Y = _gemm_into(0, A, 0, O2, empty((m, r), A.dtype, A), 0.0, precision)
    return Y


@cr("math.init_qr_streaming")
def init_qr_streaming(Ai, r: int, q: int, seed=None, omega=None, precision: str = "fp64"):
    """Qi (m, r), B (r, n), Yi (m, r) — pyLOM/vmmath/svd.py:175-197 (the
    randomized QR of a first batch of snapshots, keeping the sketch Y).
    ``seed`` None draws a time-based seed as the reference does; ``omega``
    injects the sketch matrix (n, r)."""
    A, was_np = _dev(Ai)
    if A.dtype not in (torch.float32, torch.float64):
        raise TypeError("init_qr_streaming: real float32 / float64 data")
    m, n = A.shape
    if omega is None:
        omega = reference_omega(n, r, int(time.time()) if seed is None else int(seed), np.float64)
    om, _ = _dev(omega, A.dtype)
    Y = _power_sketch(A, om, q, precision)
    Q = _tsqr_q(Y, precision)
    B = _gemm_into(1, Q, 0, A, empty((r, n), A.dtype, A), 0.0, precision)
    if was_np:
        return Q.cpu().numpy(), B.cpu().numpy(), Y.cpu().numpy()
    return Q, B, Y


@cr("math.update_qr_streaming")
def update_qr_streaming(Ai, Q1, B1, Yo, r: int, q: int, omega=None, precision: str = "fp64"):
    """Q2 (m, r), B2 (r, n_old + n), Yo (m, r) — pyLOM/vmmath/svd.py:199-225:
    fold a new batch Ai (m, n) into the sketch.  Omega is drawn from numpy's
    global stream (unseeded, as svd.py:212) unless ``omega`` is given.  Yo is
    updated in place when it is a device tensor (the reference's Yo += Yn)."""
    A, was_np = _dev(Ai)
    m, n = A.shape
    dt = A.dtype
    Q1d, _ = _dev(Q1, dt)
    B1d, _ = _dev(B1, dt)
    Yd, y_np = _dev(Yo, dt)
    if omega is None:
        omega = np.random.rand(n, r)
    om, _ = _dev(omega, dt)
    Yn = _power_sketch(A, om, q, precision)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_axpy(ctx.h, pc_dtype(Yd), 1.0, Yn.data_ptr(), Yn.stride(0), Yd.data_ptr(), Yd.stride(0), m, r))
    Q2 = _tsqr_q(Yd, precision)
    Q2Q1 = _gemm_into(1, Q2, 0, Q1d, empty((r, Q1d.shape[1]), dt, A), 0.0, precision)
    nold = B1d.shape[1]
    B2 = empty((r, nold + n), dt, A)
    # B2 = [Q2^T Q1 B1, Q2^T A], both written in place into the column blocks
    _gemm_into(0, Q2Q1, 0, B1d, B2[:, :nold], 0.0, precision)
    _gemm_into(1, Q2, 0, A, B2[:, nold:], 0.0, precision)
    if y_np and isinstance(Yo, np.ndarray):
        Yo[...] = Yd.cpu().numpy()
    if was_np:
        return Q2.cpu().numpy(), B2.cpu().numpy(), (Yo if y_np else Yd)
    return Q2, B2, Yd


# ---------------------------------------------------------------------------
# products  (pyLOM/vmmath/maths.py)
# ---------------------------------------------------------------------------
def _view_op(x):
    """(tensor, op) with op 0/1/2 for plain / transposed / conj-transposed views."""
    if isinstance(x, torch.Tensor) and x.dim() == 2 and x.stride(0) == 1 and x.stride(1) >= x.shape[0] and x.shape[0] > 1:
        return x.t(), 1
    return x, 0


@cr("math.matmul")
def matmul(A, B, precision: str = "fp64"):
    """C = A B — pyLOM/vmmath/maths.py:76-90."""
    was_np = isinstance(A, np.ndarray)
    ta, opa = _view_op(A if not was_np else A)
    tb, opb = _view_op(B if not isinstance(B, np.ndarray) else B)
    a = Arg(ta, "A")
    b = Arg(tb, "B", dtype=a.t.dtype)
    am, ak = a.t.shape if opa == 0 else a.t.shape[::-1]
    bk, bn = b.t.shape if opb == 0 else b.t.shape[::-1]
    if ak != bk:
        raise ValueError(f"matmul: inner dimensions differ ({ak} vs {bk})")
    Cm = empty((am, bn), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_gemm(ctx.h, pc_dtype(a.t), opa, opb, am, bn, ak, 1.0, a.ptr(), a.ld, b.ptr(), b.ld, 0.0,
                              Cm.data_ptr(), Cm.stride(0), _prec(precision)))
    return a.give(Cm)


@cr("math.matmulp")
def matmulp(A, B, precision: str = "fp64"):
    """pyLOM/vmmath/maths.py:92-110.  On one GPU the rank sum is the product."""
    return matmul(A, B, precision=precision)


@cr("math.vecmat")
def vecmat(v, A):
    """C[i,:] = v[i] A[i,:] — pyLOM/vmmath/maths.py:112-129 (one launch, no row loop)."""
    a = Arg(A, "A")
    vv = Arg(v, "v", dtype=a.t.dtype if a.t.is_complex() else real_dtype(a.t.dtype)).t.contiguous().reshape(-1)
    m, n = a.t.shape
    if vv.shape[0] != m:
        raise ValueError("vecmat: len(v) must equal A.shape[0]")
    if a.t.is_complex() and not vv.is_complex():
        vv = vv.to(a.t.dtype)
    Cm = empty((m, n), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_vecmat(ctx.h, pc_dtype(a.t), vv.data_ptr(), a.ptr(), a.ld, m, n, Cm.data_ptr(), Cm.stride(0)))
    return a.give(Cm)


@cr("math.transpose")
def transpose(A):
    """A^T — pyLOM/vmmath/maths.py:17-29 (materialised)."""
    a = Arg(A, "A")
    m, n = a.t.shape
    B = empty((n, m), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_transpose(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, B.data_ptr(), B.stride(0)))
    return a.give(B)


# ---------------------------------------------------------------------------
# averaging  (pyLOM/vmmath/averaging.py)
# ---------------------------------------------------------------------------
@cr("math.temporal_mean")
def temporal_mean(X):
    """Row mean over snapshots, shape (m,) — averaging.py:17-29."""
    a = Arg(X, "X")
    m, n = a.t.shape
    mean = empty((m,), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_row_stats(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, mean.data_ptr(), None))
    return a.give(mean)


@cr("math.subtract_mean")
def subtract_mean(X, X_mean):
    """X - X_mean[:, None] — averaging.py:31-44 (no tiled temporary)."""
    a = Arg(X, "X")
    mu = Arg(X_mean, "v", dtype=a.t.dtype).t.contiguous().reshape(-1)
    m, n = a.t.shape
    Y = empty((m, n), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_row_center(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, mu.data_ptr(), None, Y.data_ptr(),
                                    Y.stride(0)))
    return a.give(Y)


@cr("math.temporal_variance")
def temporal_variance(X, X_mean=None):
    """Population variance over snapshots, shape (m, 1) — averaging.py:46-59."""
    a = Arg(X, "X")
    m, n = a.t.shape
    mean = empty((m,), a.t.dtype, a.t)
    var = empty((m, 1), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_row_stats(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, mean.data_ptr(), var.data_ptr()))
    return a.give(var)


@cr("math.temporal_variance")
def norm_variance(X, X_mean, X_var):
    """(X - mean) / var (VARIANCE, not std) — averaging.py:61-75."""
    a = Arg(X, "X")
    mu = Arg(X_mean, "v", dtype=a.t.dtype).t.contiguous().reshape(-1)
    var = Arg(X_var, "v", dtype=a.t.dtype).t.contiguous().reshape(-1)
    m, n = a.t.shape
    Y = empty((m, n), a.t.dtype, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_row_center(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, m, n, mu.data_ptr(), var.data_ptr(),
                                    Y.data_ptr(), Y.stride(0)))
    return a.give(Y)


# ---------------------------------------------------------------------------
# truncation / statistics  (truncation.py, stats.py)
# ---------------------------------------------------------------------------
def compute_truncation_residual(S, r):
    """Number of modes kept — pyLOM/vmmath/truncation.py:17-39 (host, n <= 512)."""
    s = S.detach().double().cpu().numpy() if isinstance(S, torch.Tensor) else np.asarray(S, dtype=np.float64)
    s = np.ascontiguousarray(s.reshape(-1))
    N = C.c_int64(0)
    lib = _lib.load()
    rc = lib.pc_truncation_residual(s.ctypes.data_as(C.POINTER(C.c_double)), s.shape[0], float(r), C.byref(N))
    if rc != 0:
        raise _lib.PodCoreError(rc, lib.pc_last_error(None).decode())
    return int(N.value)


def _diff_norms(A, B):
    a = Arg(A, "A")
    b = Arg(B, "B", dtype=a.t.dtype)
    m, n = a.t.shape if a.t.dim() == 2 else (a.t.shape[0], 1)
    out = (C.c_double * 2)()
    ctx = _ctx()
    ctx.check(ctx.lib.pc_diff_norms(ctx.h, pc_dtype(a.t), a.ptr(), a.ld, b.ptr(), b.ld, m, n, out))
    return out[0], out[1], m * n


@cr("math.RMSE")
def RMSE(A, B, relative: bool = True) -> float:
    """Root mean square error — pyLOM/vmmath/stats.py:16-34."""
    s1, s2, cnt = _diff_norms(A, B)
    return float(np.sqrt(s1 / (s2 if relative else cnt)))


@cr("math.energy")
def energy(original, rec) -> float:
    """1 - sum (o - r)^2 / sum o^2 — pyLOM/vmmath/truncation.py:41-60."""
    s1, s2, _ = _diff_norms(original, rec)
    return float(1.0 - s1 / s2)


# ---------------------------------------------------------------------------
# eigen  (pyLOM/vmmath/maths.py:161-183)
# ---------------------------------------------------------------------------
@cr("math.eigen")
def eigen(A):
    """(real, imag, vecs) of a real square matrix — pyLOM/vmmath/maths.py:161-183.

    Computed on the device (Hessenberg reduction, Francis double-shift QR,
    back substitution) with numpy.linalg.eig's conventions: conjugate pairs
    positive imaginary part first, unit-norm vectors, a pair's largest
    component real.  ``vecs`` is complex (real when every eigenvalue is real,
    as numpy returns it).  The eigenvalue order follows this solver's Schur
    form and can differ from LAPACK's.
    """
    a = Arg(A, "A", dtype=torch.float64)
    n, n2 = a.t.shape
    if n != n2:
        raise ValueError("eigen: A must be square")
    wr = empty((n,), torch.float64, a.t)
    wi = empty((n,), torch.float64, a.t)
    V = empty((n, n), torch.complex128, a.t)
    ctx = _ctx()
    ctx.check(ctx.lib.pc_eig(ctx.h, a.ptr(), a.ld, n, wr.data_ptr(), wi.data_ptr(), V.data_ptr(), V.stride(0)))
    if not bool((wi != 0).any()):
        V = V.real.contiguous()
    return a.give(wr), a.give(wi), a.give(V)

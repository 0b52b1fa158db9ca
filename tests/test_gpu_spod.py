"""GPU parity of the grouped solve and of SPOD (pyLOM/SPOD/wrapper.py) against
the CPU oracle.

Tolerances: the grouped solve must reproduce the single solves bit for bit
(same kernels, same inputs, deterministic reductions).  The Welch blocks are
float32 DFTs against the oracle's float64 FFT cast to complex64: relative error
<= 2e-6 of the largest block norm.  Per frequency, |dL_j| <= 2e-5 L_1(f) (the
1e-5 sigma bar relative to that frequency's leading mode), and the gated left
singular subspace (sigma >= 1e-3 sigma_1(f)) within 1e-3 rad of the oracle's.
P is Re(U): the complex phase of each mode is arbitrary (LAPACK's choice in the
reference), so P is compared after aligning each mode's phase to the oracle's.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from oracle import metrics

pytestmark = pytest.mark.gpu

pl = pytest.importorskip("pyloworder_amd")


def _np(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def _flow(M, N, seed, dtype=np.float64):
    """Travelling waves at three distinct frequencies over a noisy mean field."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, M)[:, None]
    t = np.arange(N) * 0.01
    X = 3.0 + 0.5 * rng.standard_normal((M, 1))
    for amp, om, k in ((1.0, 2 * np.pi * 12.5, 6.0), (0.75, 2 * np.pi * 25.0, 11.0), (0.55, 2 * np.pi * 6.25, 3.0)):
        X = X + amp * np.cos(k * 2 * np.pi * x - om * t[None, :])
    X = X + 0.05 * rng.standard_normal((M, N))
    return X.astype(dtype), t


def test_grouped_bitwise_equals_single(rng):
    shapes = [(3000, 7), (12000, 40), (800, 1), (5000, 256), (20000, 17), (4000, 300)]
    As = [rng.standard_normal(s) @ np.diag(np.logspace(0, -4, s[1])) for s in shapes]
    Ug, Sg, Vg = pl.math.tsqr_svd_grouped(As)
    for A, U, S, V in zip(As, Ug, Sg, Vg):
        U1, S1, V1 = pl.math.tsqr_svd(A)
        assert np.array_equal(S, S1) and np.array_equal(U, U1) and np.array_equal(V, V1)


def test_grouped_mixed_bitwise_equals_single(rng):
    """float32 mixed: the grouped members' one-launch Jacobi steps and
    Householder QR (n = 256 and 100 take the preconditioned block Jacobi)
    give the single solves' split-launch results bit for bit."""
    shapes = [(40000, 256), (7000, 100), (3000, 33), (9000, 256)]
    As = [(rng.standard_normal(s) @ np.diag(np.logspace(0, -3, s[1]))).astype(np.float32) for s in shapes]
    Ug, Sg, Vg = pl.math.tsqr_svd_grouped(As, precision="mixed")
    for A, U, S, V in zip(As, Ug, Sg, Vg):
        U1, S1, V1 = pl.math.tsqr_svd(A, precision="mixed")
        assert np.array_equal(S, S1) and np.array_equal(U, U1) and np.array_equal(V, V1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_grouped_complex64_uneven(precision, rng):
    shapes = [(4000, 12), (900, 33), (15000, 5)]
    As = [((rng.standard_normal(s) + 1j * rng.standard_normal(s)) @ np.diag(np.logspace(0, -2, s[1])))
          .astype(np.complex64) for s in shapes]
    Ug, Sg, Vg = pl.math.tsqr_svd_grouped(As, precision=precision)
    for A, U, S, V in zip(As, Ug, Sg, Vg):
        U1, S1, V1 = pl.math.tsqr_svd(A, precision=precision)
        assert np.array_equal(S, S1) and np.array_equal(U, U1)
        Uo, So, _ = oracle.tsqr_svd([A.astype(np.complex128)])
        e = metrics.sigma_errors(S.astype(np.float64), So)
        assert e["maxrel_gated"] <= 1e-5, e
        assert metrics.max_principal_angle(U.astype(np.complex128), Uo[0], e["n_gated"]) <= 1e-4


@pytest.mark.parametrize("dtype,precision", [(np.complex128, "fp64"), (np.complex64, "fp64"), (np.complex64, "mixed")])
def test_uniform_complex_group_bitwise_equals_single(dtype, precision, rng):
    """Equal shapes in one stacked tensor: the small stages of all members run
    as single launches, and every member still equals its lone solve bit for
    bit.  Member 3 has rank 4 (null-space completion of U)."""
    k, m, n = 7, 3000, 15
    A = (rng.standard_normal((k, m, n)) + 1j * rng.standard_normal((k, m, n))) @ np.diag(np.logspace(0, -3, n))
    A[3] = A[3][:, :4] @ (rng.standard_normal((4, n)) + 1j * rng.standard_normal((4, n)))
    At = torch.from_numpy(A.astype(dtype)).cuda()
    Ug, Sg, Vg = pl.math.tsqr_svd_grouped(At, precision=precision)
    for i in range(k):
        U1, S1, V1 = pl.math.tsqr_svd(At[i], precision=precision)
        assert torch.equal(Sg[i], S1) and torch.equal(Ug[i], U1) and torch.equal(Vg[i], V1), i
    U3 = Ug[3].to(torch.complex128)
    assert torch.allclose(U3.conj().T @ U3, torch.eye(n, dtype=torch.complex128, device="cuda"), atol=1e-5)


@pytest.mark.parametrize("M,nB,nf", [(1000, 7, 130), (37, 15, 129), (50, 300, 9), (5, 1, 3), (40, 100, 200)])
@pytest.mark.parametrize("out", [torch.float32, torch.float64])
def test_spod_modes_layout_bitwise(M, nB, nf, out, rng):
    """pc_spod_modes equals the plain gather P[j M + i, k] = Re U[order[k]][i, j]
    (pyLOM/SPOD/wrapper.py:84, 90-94), bit for bit."""
    U = torch.from_numpy((rng.standard_normal((nf, M, nB)) + 1j * rng.standard_normal((nf, M, nB)))
                         .astype(np.complex64)).cuda()
    order = torch.from_numpy(rng.permutation(nf).astype(np.int64)).cuda()
    ref = U.real.permute(2, 1, 0).index_select(2, order).reshape(nB * M, nf).to(out)
    P = torch.empty((nB * M, nf), dtype=out, device="cuda")
    ctx = pl._lib.context()
    ctx.check(ctx.lib.pc_spod_modes(ctx.h, U.data_ptr(), M, nB, nf, order.data_ptr(), P.data_ptr(),
                                    1 if out == torch.float64 else 0))
    torch.cuda.synchronize()
    assert torch.equal(P, ref)


def test_uniform_group_over_budget_falls_back(rng):
    """A uniform group whose per-problem workspace slices exceed the context's
    HBM budget runs problem by problem instead, with the same results."""
    from pyloworder_amd import _lib
    k, m, n = 7, 3000, 15
    A = torch.from_numpy((rng.standard_normal((k, m, n)) + 1j * rng.standard_normal((k, m, n)))
                         .astype(np.complex64)).cuda()
    U = torch.empty_like(A)
    S = torch.empty((k, n), dtype=torch.float32, device="cuda")
    V = torch.empty((k, n, n), dtype=torch.complex64, device="cuda")
    lib = _lib.load()
    need = lib.pc_workspace_query(0, _lib.PC_C64, m, n, 0, _lib.PC_PREC_MIXED)
    assert need > 0
    ctx = _lib.Context(torch.cuda.current_device(), 3 * need)  # room for 3 slices, not 7
    probs = (_lib.PCProblem * k)()
    for i in range(k):
        probs[i] = _lib.PCProblem(A[i].data_ptr(), n, m, n, 0, 0, U[i].data_ptr(), n, S[i].data_ptr(),
                                  V[i].data_ptr(), n)
    torch.cuda.synchronize()
    try:
        ctx.check(lib.pc_grouped_tsqr_svd(ctx.h, _lib.PC_C64, C.cast(probs, C.c_void_p), k, _lib.PC_PREC_MIXED))
        torch.cuda.synchronize()
    finally:
        ctx.close()
    for i in range(k):
        U1, S1, V1 = pl.math.tsqr_svd(A[i], precision="mixed")
        assert torch.equal(S[i], S1) and torch.equal(U[i], U1) and torch.equal(V[i], V1), i


def test_grouped_rejects_complex_centring(rng):
    A = torch.zeros((3, 100, 4), dtype=torch.complex64, device="cuda")
    with pytest.raises(pl.PodCoreError, match="problem 0: .*real data"):
        pl.math.tsqr_svd_grouped(A, center=1)


def test_uniform_group_error_names_the_problem(rng):
    A = torch.from_numpy((rng.standard_normal((5, 800, 9)) + 1j * rng.standard_normal((5, 800, 9)))
                         .astype(np.complex64)).cuda()
    A[2, 17, 3] = float("nan")
    with pytest.raises(pl.PodCoreError, match="problem 2"):
        pl.math.tsqr_svd_grouped(A, precision="mixed")


def test_grouped_errors_name_the_problem(rng):
    A = rng.standard_normal((100, 4))
    with pytest.raises(pl.PodCoreError, match="problem 1"):
        pl.math.tsqr_svd_grouped([A, rng.standard_normal((100, 300))])


@pytest.mark.parametrize("M,N,nDFT,nolap", [(700, 400, 0, 0), (300, 3000, 1024, 512), (257, 500, 99, 30),
                                            (64, 200, 2, 1), (33, 4200, 2048, 1024), (21, 4100, 4096, 0),
                                            (130, 700, 256, 100), (50, 2048, 256, 128), (41, 5000, 256, 0)])
def test_spod_blocks_vs_oracle(M, N, nDFT, nolap):
    X, t = _flow(M, N, seed=M + N)
    Qo, fo, nbo = oracle.spod_blocks(X, t, nDFT, nolap)
    for Xin in (X, X.astype(np.float32)):
        Q, f, nb = pl.SPOD.blocks(Xin, t, nDFT, nolap)
        assert nb == nbo and np.array_equal(f, fo) and Q.shape == Qo.shape
        scale = np.abs(Qo).max()
        assert np.abs(Q - Qo).max() <= 2e-6 * scale * np.sqrt(Qo.shape[2]), np.abs(Q - Qo).max() / scale


def test_spod_blocks_no_mean_removal():
    X, t = _flow(200, 300, seed=3)
    Qo, _, _ = oracle.spod_blocks(X, t, 64, 16, remove_mean=False)
    Q, _, _ = pl.SPOD.blocks(X.astype(np.float32), t, 64, 16, remove_mean=False)
    assert np.abs(Q - Qo).max() <= 2e-6 * np.abs(Qo).max() * 8


def _by_freq(L, f):
    return {float(fk): L[k] for k, fk in enumerate(f)}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spod_run_vs_oracle(precision, dtype):
    M, N = 1500, 640
    X, t = _flow(M, N, seed=11, dtype=dtype)
    Lo, Po, fo = oracle.spod_run(X.astype(np.float64), t)
    L, P, f = pl.SPOD.run(X, t, precision=precision)
    assert L.shape == Lo.shape and P.shape == Po.shape and f.shape == fo.shape
    assert L.dtype == X.dtype and P.dtype == X.dtype
    # sorted by the leading energy, descending (wrapper.py:90-94)
    assert np.all(np.diff(L[:, 0]) <= 0)
    # frequency by frequency (ties in L[:, 0] may order differently)
    lo, lg = _by_freq(Lo, fo), _by_freq(L, f)
    assert lo.keys() == lg.keys()
    for fk, row in lo.items():
        assert np.abs(lg[fk] - row).max() <= 2e-5 * row[0] + 1e-30, (fk, np.abs(lg[fk] - row).max() / row[0])
    # the dominant frequencies of the flow come out on top, in order
    assert list(f[:3]) == list(fo[:3]) == [12.5, 25.0, 6.25]


def test_spod_modes_match_oracle_up_to_phase():
    M, N = 1200, 512
    X, t = _flow(M, N, seed=5)
    Qo, fo, nb = oracle.spod_blocks(X, t)
    Q, f, _ = pl.SPOD.blocks(X, t)
    Ug, Sg, _ = pl.math.tsqr_svd_grouped(list(Q))
    for k in range(len(fo)):
        Uo, So, _ = oracle.tsqr_svd([Qo[k].astype(np.complex128)])
        U = Ug[k].astype(np.complex128)
        gated = int(np.sum(So >= 1e-3 * So[0]))
        assert metrics.max_principal_angle(U, Uo[0], gated) <= 1e-3
        # P = Re(U) column-major (wrapper.py:84), compared after phase alignment
        # of the well-separated modes
        gap = np.abs(np.diff(So)) > 1e-2 * So[0]
        for j in range(gated):
            sep = (j == 0 or gap[j - 1]) and (j == len(So) - 1 or gap[j])
            if not sep:
                continue
            c = np.vdot(U[:, j], Uo[0][:, j])
            u = U[:, j] * (c / abs(c))
            assert np.abs(np.real(u) - np.real(Uo[0][:, j])).max() <= 1e-3


def test_spod_run_P_layout_matches_grouped_U():
    M, N = 600, 300
    X, t = _flow(M, N, seed=9, dtype=np.float32)
    L, P, f = pl.SPOD.run(X, t)
    Q, f0, nb = pl.SPOD.blocks(X, t)
    Ug, Sg, _ = pl.math.tsqr_svd_grouped(list(Q))
    for k in range(len(f)):
        i = int(np.where(f0 == f[k])[0][0])
        assert np.array_equal(P[:, k], np.real(Ug[i]).reshape(M * nb, order="F"))
        assert np.array_equal(L[k], (Sg[i].astype(np.float64) ** 2).astype(np.float32))


def test_spod_rejects_bad_geometry():
    X, t = _flow(100, 50, seed=1)
    with pytest.raises(ValueError):
        pl.SPOD.run(X, t, nDFT=8192)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_grouped_graph_replays_match_direct(precision, rng):
    """A recorded group replayed three times, with the inputs changed in place
    and eager work in between: every replay equals the direct grouped call on
    the current data, bit for bit."""
    real = [torch.from_numpy(rng.standard_normal(s) @ np.diag(np.logspace(0, -3, s[1]))).float().cuda()
            for s in [(5000, 12), (3000, 40), (9000, 7)]]
    cplx = torch.from_numpy((rng.standard_normal((6, 4000, 15)) + 1j * rng.standard_normal((6, 4000, 15)))
                            .astype(np.complex64)).cuda()
    for As in (real, cplx):
        g = pl.math.GroupedGraph(As, precision=precision)
        try:
            for rep in range(3):
                if rep:
                    for A in (As if isinstance(As, list) else [As]):
                        A.mul_(1.5).add_(0.25)  # new data in the same buffers
                    torch.cuda.synchronize()
                Us, Ss, Vs = g.launch()
                _ = torch.ones(64, 64, device="cuda") @ torch.ones(64, 64, device="cuda")  # eager work between
                Ud, Sd, Vd = pl.math.tsqr_svd_grouped(As, precision=precision)
                for a, b in zip(Ss, Sd):
                    assert torch.equal(a, b), rep
                for a, b in zip(Us, Ud):
                    assert torch.equal(a, b), rep
                for a, b in zip(Vs, Vd):
                    assert torch.equal(a, b), rep
        finally:
            g.close()

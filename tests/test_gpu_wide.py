"""GPU parity of wide problems (256 < n <= 2048 real, 128 < n <= 1024
complex): stripe-pair FP64 MFMA Gram passes and the K-looped product pass,
against the CPU oracle (numpy FP64 SVD / the reference's tsqr_svd restated).

The reference's own GPU drivers run N = 250 to 2000 columns
(Examples/scalability/example_SVD_GPU_strong.py:9-16).  Tolerances as the
FP64 tests: gated singular values 1e-10 relative, every non-null one 1e-8,
gated principal angle 1e-8, orthogonality 1e-10, reconstruction RMSE 1e-12.
Mixed precision stops at n = 256 and must say so.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import metrics, workloads

pytestmark = pytest.mark.gpu
pl = pytest.importorskip("pyloworder_amd")


def _np(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def check(Y, U, S, V, gated=1e-10, allrel=1e-8, ang=1e-8, orth=1e-10, rmse=1e-12):
    Y = np.asarray(Y)
    Yo = Y.astype(np.complex128 if np.iscomplexobj(Y) else np.float64)
    Uo, So, Vo = np.linalg.svd(Yo, full_matrices=False)
    U, S, V = _np(U), _np(S), _np(V)
    e = metrics.sigma_errors(S, So)
    assert e["maxrel_gated"] <= gated, e
    assert e["maxrel_all"] <= allrel, e
    k = e["n_gated"]
    assert metrics.max_principal_angle(U.astype(Uo.dtype), Uo, k) <= ang
    assert metrics.orthogonality(U.astype(Uo.dtype)) <= orth
    assert metrics.orthogonality(np.conj(V.astype(Uo.dtype)).T) <= orth
    assert metrics.recon_rmse(Yo, U.astype(Uo.dtype), S.astype(np.float64), V.astype(Uo.dtype)) <= rmse
    return e


@pytest.mark.parametrize("m,n", [(12000, 300), (9000, 512), (20000, 1000)])
def test_wide_tsqr_svd_fp64(m, n):
    X = workloads.lowrank(m, n, seed=n, decay=0.98, top=1e3, noise=1e-6)
    U, S, V = pl.math.tsqr_svd(X)
    # U = Yc W: its orthogonality error grows like n u (5e-10 measured at n = 1000)
    e = check(X, U, S, V, orth=max(1e-10, 1e-12 * n))
    assert e["n_gated"] >= 200 or n == 300


def test_wide_pod_run_centred_null_completion():
    """remove_mean at full width: rank n - 1, the null column completed
    orthonormally (the reference's thin-SVD contract)."""
    X = workloads.lowrank(15000, 400, seed=5, decay=0.97, top=100.0, noise=1e-5)
    U, S, V = pl.POD.run(X, remove_mean=True)
    Y = X - X.mean(axis=1, keepdims=True)
    check(Y, U, S, V, allrel=1e-6)
    Uo, So, Vo = oracle.tsqr_svd([Y])
    assert metrics.sigma_errors(_np(S), So)["maxrel_gated"] <= 1e-10


def test_wide_float32_storage_fp64_math():
    X = workloads.lowrank(10000, 333, seed=3, decay=0.97, top=10.0, noise=1e-4, dtype=np.float32)
    U, S, V = pl.math.tsqr_svd(X)
    assert _np(U).dtype == np.float32
    # FP64 arithmetic, float32 outputs: S, U and V carry float32 rounding (6e-8)
    check(X, U, S, V, gated=2e-7, allrel=1e-5, ang=1e-5, orth=1e-6, rmse=1e-6)


def test_wide_complex128():
    rng = np.random.default_rng(8)
    A = (rng.standard_normal((8000, 200)) + 1j * rng.standard_normal((8000, 200))) @ np.diag(np.logspace(0, -4, 200))
    U, S, V = pl.math.tsqr_svd(A)
    check(A, U, S, V)


def test_wide_tsqr_lapack_signs():
    rng = np.random.default_rng(9)
    A = rng.standard_normal((6000, 300)) @ np.diag(np.logspace(0, -3, 300))
    Q, R = pl.math.tsqr(A)
    Qo, Ro = np.linalg.qr(A)
    assert np.abs(_np(R) - Ro).max() <= 1e-9 * np.abs(Ro).max()
    assert np.abs(_np(Q) - Qo).max() <= 1e-8


def test_wide_dense_svd_square():
    rng = np.random.default_rng(10)
    A = rng.standard_normal((600, 600))
    U, S, V = pl.math.svd(A)
    assert np.allclose(_np(S), np.linalg.svd(A, compute_uv=False), rtol=1e-10)
    assert metrics.recon_rmse(A, _np(U), _np(S), _np(V)) <= 1e-12


def test_wide_mixed_is_refused_with_a_reason():
    X = workloads.lowrank(5000, 300, seed=1, dtype=np.float32)
    with pytest.raises(pl.PodCoreError, match="mixed precision takes n <= 256"):
        pl.math.tsqr_svd(X, precision="mixed")
    with pytest.raises(pl.PodCoreError):
        pl.math.tsqr_svd(np.ones((3000, 2049)))


def test_wide_randomized_svd_fp64_vs_oracle():
    """randomized_svd (pyLOM/vmmath/svd.py:254-273) at n = 512, r = 200, q = 2
    against the oracle with the same seeded Omega: gated singular values 1e-9,
    their left subspace 1e-7 rad."""
    X = workloads.lowrank(30000, 512, seed=12, decay=0.97, top=1e3, noise=1e-6)
    U, S, V = pl.math.randomized_svd(X, 200, 2, seed=7)
    Uo, So, Vo = oracle.randomized_svd([X], 200, 2, seed=7)
    e = metrics.sigma_errors(_np(S), So)
    assert e["maxrel_gated"] <= 1e-9, e
    assert metrics.max_principal_angle(_np(U), np.concatenate(Uo), min(e["n_gated"], 150)) <= 1e-7
    Q, B = pl.math.randomized_qr(X, 200, 2, seed=7)
    Qo, Bo = oracle.randomized_qr([X], 200, 2, seed=7)
    assert np.abs(_np(Q) - Qo[0]).max() <= 1e-8 and np.abs(_np(B) - Bo).max() <= 1e-8 * np.abs(Bo).max()


def test_mixed_randomized_wide_sketch():
    """mixed randomized_svd with r = 192 > 128 (the factor's FP64 randomized
    solve takes any r <= n)."""
    X = workloads.lowrank(60000, 256, seed=13, decay=0.96, top=1e3, noise=1e-6 * np.sqrt(8e6 / 60000),
                          dtype=np.float32)
    U, S, V = pl.math.randomized_svd(X, 192, 1, seed=3, precision="mixed")
    Uo, So, Vo = oracle.randomized_svd([X.astype(np.float64)], 192, 1, seed=3)
    e = metrics.sigma_errors(_np(S), So)
    assert e["maxrel_gated"] <= 1e-5, e


def test_wide_products():
    rng = np.random.default_rng(14)
    U = np.linalg.qr(rng.standard_normal((20000, 300)))[0]
    S = np.linspace(3.0, 1.0, 300)
    V = rng.standard_normal((300, 700))
    Xr = pl.POD.reconstruct(U, S, V)
    ref = (U * S) @ V
    assert np.abs(_np(Xr) - ref).max() <= 1e-12 * np.abs(ref).max()
    Cm = pl.POD.project(U, ref)
    assert np.abs(_np(Cm) - S[:, None] * V).max() <= 1e-11 * np.abs(V).max() * 3
    A = rng.standard_normal((5000, 600))
    B = rng.standard_normal((600, 400))
    assert np.abs(_np(pl.math.matmul(A, B)) - A @ B).max() <= 1e-11 * np.abs(A @ B).max()
    AtB = pl.math.matmul(torch.from_numpy(A).cuda().T, torch.from_numpy(rng.standard_normal((5000, 300))).cuda())
    assert tuple(AtB.shape) == (600, 300)

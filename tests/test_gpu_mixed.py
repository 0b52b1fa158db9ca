"""GPU parity of the mixed-precision path (float32 storage, fp16 hi/lo split
MFMA, FP64 small steps) against the FP64 oracle on the same stored data.

Tolerance (the stated bar): singular-value relative error <= 1e-5 over the
gated modes (s >= 1e-3 s_1); the largest principal angle of the gated left
subspace and the reconstruction RMSE are bounded by FP32-level figures.
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


def oracle_svd(Y):
    U, S, V = np.linalg.svd(np.asarray(Y, dtype=np.float64), full_matrices=False)
    return U, S, V


@pytest.mark.parametrize("m,n", [(4096, 32), (20000, 64), (30011, 128), (65536, 256)])
def test_mixed_lowrank_gated_sigma(m, n):
    X = workloads.lowrank(m, n, seed=n, decay=0.93, top=1e3, noise=1e-6 * np.sqrt(8e6 / m), dtype=np.float32)
    U, S, V = pl.math.tsqr_svd(X, precision="mixed")
    Uo, So, Vo = oracle_svd(X)
    e = metrics.sigma_errors(_np(S), So)
    assert e["maxrel_gated"] <= 1e-5, e
    k = e["n_gated"]
    assert metrics.max_principal_angle(_np(U).astype(np.float64), Uo, k) <= 1e-3
    assert metrics.orthogonality(_np(U)[:, :k].astype(np.float64)) <= 1e-4
    assert metrics.recon_rmse(X, _np(U), _np(S), _np(V)) <= 1e-5


def test_mixed_pod_run_with_mean_removal():
    X = workloads.sines(20000, 200, dtype=np.float32)
    U, S, V = pl.POD.run(X, remove_mean=True, precision="mixed")
    Y = X.astype(np.float64) - X.astype(np.float64).mean(axis=1, keepdims=True)
    Uo, So, Vo = oracle_svd(Y)
    e = metrics.sigma_errors(_np(S), So)
    assert e["maxrel_gated"] <= 1e-5, e
    assert metrics.recon_rmse(Y, _np(U), _np(S), _np(V)) <= 1e-5


def test_mixed_complex64_block():
    rng = np.random.default_rng(3)
    Q = ((rng.standard_normal((20000, 15)) + 1j * rng.standard_normal((20000, 15))) @ np.diag(np.logspace(0, -2, 15))).astype(np.complex64)
    U, S, V = pl.math.tsqr_svd(Q, precision="mixed")
    So = np.linalg.svd(Q.astype(np.complex128), compute_uv=False)
    e = metrics.sigma_errors(_np(S), So)
    assert e["maxrel_gated"] <= 1e-5, e
    assert metrics.recon_rmse(Q.astype(np.complex128), _np(U), _np(S), _np(V)) <= 1e-5


def test_mixed_large_row_magnitude_spread():
    # rows whose magnitudes span 1e6 (per-row scales keep their relative accuracy)
    rng = np.random.default_rng(4)
    X = (rng.standard_normal((20000, 48)) * np.logspace(-3, 3, 20000)[:, None]).astype(np.float32)
    U, S, V = pl.math.tsqr_svd(X, precision="mixed")
    Uo, So, Vo = oracle_svd(X)
    e = metrics.sigma_errors(_np(S), So)
    assert e["maxrel_gated"] <= 1e-5, e


def test_mixed_rejects_float64():
    with pytest.raises(pl.PodCoreError):
        pl.math.tsqr_svd(np.ones((100, 4)), precision="mixed")

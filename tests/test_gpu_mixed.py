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


@pytest.mark.parametrize("n,r,q", [(256, 74, 3), (64, 12, 1), (40, 40, 0)])
def test_mixed_randomized_svd_vs_oracle(n, r, q):
    """Mixed randomized_svd (float32 storage; pyLOM/vmmath/svd.py:254-273) at the
    8M x 256 workload's sketch (r = 74, q = 3) against the oracle's FP64
    randomized_svd with the same seeded Omega: gated singular values (s >= 1e-3
    s_1 of the oracle's) to 1e-5 relative, their left subspace to 1e-3 rad."""
    m = 120_000 if n == 256 else 20_000
    X = workloads.lowrank(m, n, seed=n + r, decay=0.93, top=1e3, noise=1e-6 * np.sqrt(8e6 / m), dtype=np.float32)
    U, S, V = pl.math.randomized_svd(X, r, q, seed=31, precision="mixed")
    U, S, V = _np(U), _np(S), _np(V)
    assert U.shape == (m, r) and S.shape == (r,) and V.shape == (r, n) and U.dtype == np.float32
    Uo, So, Vo = oracle.randomized_svd([X.astype(np.float64)], r, q, seed=31)
    e = metrics.sigma_errors(S, So)
    assert e["maxrel_gated"] <= 1e-5, e
    k = min(e["n_gated"], 64)
    assert metrics.max_principal_angle(U.astype(np.float64), np.concatenate(Uo), k) <= 1e-3
    assert metrics.orthogonality(U[:, :k].astype(np.float64)) <= 1e-4
    # the same call drawing Omega on the device from the seed, and passing numpy's
    om = pl.math.reference_omega(n, r, 31, np.float32)
    U2, S2, V2 = pl.math.randomized_svd(X, r, q, seed=31, omega=om, precision="mixed")
    assert np.array_equal(_np(S2), S) and np.array_equal(_np(U2), U)


def test_mixed_randomized_qr_matches_reference_q_and_b():
    """randomized_qr (pyLOM/vmmath/svd.py:120-144), mixed: Q and B themselves,
    LAPACK's QR signs included, against the FP64 oracle on the same data."""
    m, n, r, q = 50_000, 96, 20, 2
    X = workloads.lowrank(m, n, seed=9, decay=0.85, top=10.0, noise=1e-6, dtype=np.float32)
    Q, B = pl.math.randomized_qr(X, r, q, seed=4, precision="mixed")
    Q, B = _np(Q).astype(np.float64), _np(B).astype(np.float64)
    Qo, Bo = oracle.randomized_qr([X.astype(np.float64)], r, q, seed=4)
    assert np.abs(Q - Qo[0]).max() <= 1e-4 * np.abs(Qo[0]).max()
    assert np.abs(B - Bo).max() <= 1e-5 * np.abs(Bo).max()


@pytest.mark.parametrize("n", [32, 256])
def test_mixed_pod_run_divide_variance(n):
    """POD.run(remove_mean=True, divide_variance=True) on the mixed path (exact
    FP64 row statistics, then rows divided by their variance, averaging.py:75)
    against the FP64 oracle: gated sigma 1e-5, reconstruction RMSE 1e-5."""
    rng = np.random.default_rng(n)
    X = workloads.lowrank(30000, n, seed=n, decay=0.9, top=100.0, noise=1e-4, dtype=np.float64)
    X = (X * rng.uniform(0.5, 3.0, (30000, 1)) + rng.uniform(-5, 5, (30000, 1))).astype(np.float32)
    U, S, V = pl.POD.run(X, remove_mean=True, divide_variance=True, precision="mixed")
    X64 = X.astype(np.float64)
    mu = oracle.temporal_mean(X64)
    Y = oracle.norm_variance(X64, mu, oracle.temporal_variance(X64, mu))
    Uo, So, Vo = oracle_svd(Y)
    e = metrics.sigma_errors(_np(S), So)
    assert e["maxrel_gated"] <= 1e-5, e
    assert metrics.recon_rmse(Y, _np(U), _np(S), _np(V)) <= 1e-5


def test_mixed_deep_accumulation_2M_rows():
    """Mixed tsqr_svd on 2,097,152 x 256 (about 2000 rows summed per chunk in
    FP32 registers, the headline's regime) against numpy FP64: sigma of the R
    of numpy.linalg.qr of the upcast data, gated modes to 1e-5 relative."""
    m, n = 2_097_152, 256
    g = torch.Generator(device="cuda").manual_seed(77)
    G2 = torch.linalg.qr(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))[0]
    s = 1e3 * 0.93 ** torch.arange(n, device="cuda", dtype=torch.float64)
    X = torch.empty((m, n), dtype=torch.float32, device="cuda")
    step = 1 << 19
    for r0 in range(0, m, step):
        G1 = torch.randn(step, n, generator=g, device="cuda", dtype=torch.float64) / np.sqrt(m)
        X[r0:r0 + step] = ((G1 * s) @ G2.T + 1e-6 * torch.randn(step, n, generator=g, device="cuda",
                                                                   dtype=torch.float64)).float()
    U, S, V = pl.math.tsqr_svd(X, precision="mixed", center=1)
    Xh = X.cpu().numpy().astype(np.float64)
    Xh -= Xh.mean(axis=1, keepdims=True)
    R = np.linalg.qr(Xh, mode="r")
    So = np.linalg.svd(R, compute_uv=False)
    e = metrics.sigma_errors(_np(S).astype(np.float64), So)
    assert e["n_gated"] >= 90 and e["maxrel_gated"] <= 1e-5, e

"""The MI355X path against the committed golden fixtures (tests/golden/*.npz).

FP64 path: singular values 1e-10 relative on the gated modes, gated left
subspace within 1e-8 rad; mixed path (float32 storage): 1e-5 / 1e-4.  The
randomized fixture is reproduced with the reference's own Omega stream, so the
result is the reference's numpy-path result to rounding.
"""
import os

import numpy as np
import pytest

from oracle import metrics

pytestmark = pytest.mark.gpu

pl = pytest.importorskip("pyloworder_amd")
G = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(G, name + ".npz"))


def test_example_svd():
    d = load("example_svd")
    for f in (pl.math.svd, pl.math.tsqr_svd):
        U, S, V = f(d["A"])
        assert np.allclose(S, d["S"], rtol=1e-13)
        assert np.allclose((U * S) @ V, d["A"], atol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pod_sines(precision):
    d = load("pod_sines")
    X = d["X"] if precision == "fp64" else d["X"].astype(np.float32)
    U, S, V = pl.POD.run(X, remove_mean=True, precision=precision)
    k = int(d["k"])
    tol, ang = (1e-10, 1e-8) if precision == "fp64" else (1e-5, 1e-4)
    if precision == "mixed":
        # the fixture is FP64 math on FP64 data; float32 storage moves sigma_j by
        # up to ~6e-8 sigma_1
        tol = max(tol, 6e-8 * d["S"][0] / d["S"][k - 1] * 4)
    assert np.abs(S[:k] - d["S"][:k]).max() <= tol * d["S"][:k].max()
    assert (np.abs(S[:k] - d["S"][:k]) / d["S"][:k]).max() <= tol
    assert metrics.max_principal_angle(U[:, :k].astype(np.float64), d["U"], k) <= ang


def test_randomized_reference_omega():
    d = load("randomized")
    r, q, seed = int(d["r"]), int(d["q"]), int(d["seed"])
    om = pl.math.reference_omega(d["X"].shape[1], r, seed, np.float64)
    U, S, V = pl.math.randomized_svd(d["X"], r, q, seed=seed, omega=om)
    # modes well above the noise floor (sigma >= 1e-5 sigma_1); those near it
    # are rounding-sensitive in any implementation
    k = int(np.sum(d["S"] >= 1e-5 * d["S"][0]))
    rel = np.abs(S[:k] - d["S"][:k]) / d["S"][:k]
    assert k >= 5 and rel.max() <= 1e-9, rel
    assert metrics.max_principal_angle(U[:, :k], d["U"][:, :k], k) <= 1e-7


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_randomized_device_omega_is_the_reference_stream(dt):
    """omega = None draws Omega on the device from the seed: numpy's
    RandomState(seed).random_sample((n, r)) (pyLOM/vmmath/svd.py:131-133)
    bit for bit, so every output equals the call given numpy's Omega exactly."""
    d = load("randomized")
    X = d["X"].astype(dt)
    r, q, seed = int(d["r"]), int(d["q"]), int(d["seed"])
    for sd in (seed, 0, 2**32 - 1):
        om = pl.math.reference_omega(X.shape[1], r, sd, dt)
        Q1, B1 = pl.math.randomized_qr(X, r, q, seed=sd)
        Q2, B2 = pl.math.randomized_qr(X, r, q, seed=sd, omega=om)
        assert np.array_equal(Q1, Q2) and np.array_equal(B1, B2)
    with pytest.raises(ValueError):
        pl.math.randomized_qr(X, r, q, seed=2**32)


def test_spod_fixture():
    d = load("spod")
    Q, f, nB = pl.SPOD.blocks(d["X"], d["t"], int(d["nDFT"]), int(d["nolap"]))
    assert np.array_equal(f, d["f"])
    assert np.abs(Q - d["Q"]).max() <= 2e-6 * np.abs(d["Q"]).max()
    L, P, fs = pl.SPOD.run(d["X"], d["t"], int(d["nDFT"]), int(d["nolap"]))
    assert np.array_equal(fs, d["f_sorted"])
    Lo = d["L"]
    assert np.all(np.abs(L - Lo) <= 2e-5 * Lo[:, :1] + 1e-30)


def test_truncation_and_averaging():
    d = load("truncation")
    assert [pl.math.compute_truncation_residual(d["S"], r) for r in d["r"]] == list(d["N"])
    a = load("averaging")
    assert np.allclose(pl.math.temporal_mean(a["X"]), a["mean"], rtol=0, atol=1e-15)
    var = pl.math.temporal_variance(a["X"], a["mean"])
    assert np.allclose(var.reshape(-1), a["var"].reshape(-1), rtol=1e-15)
    nv = pl.math.norm_variance(a["X"], a["mean"], var)
    assert np.allclose(nv, a["normvar"], rtol=1e-15)

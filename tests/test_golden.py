"""The oracle reproduces the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py).  A change to the oracle that moves a
fixture fails here; the GPU parity tests read the same fixtures."""
import os

import numpy as np
import pytest

import oracle
from oracle import metrics

G = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(G, name + ".npz"))


def test_example_svd_known_answer():
    d = load("example_svd")
    # the example's printed answer (Examples/example_SVD.py:17)
    assert np.allclose(d["S"], [14.2690955, 0.6268282], atol=5e-8)
    assert np.allclose(oracle.svd(d["A"])[1], d["S"], rtol=1e-14)


def test_pod_sines():
    d = load("pod_sines")
    U, S, V = oracle.pod_run(d["X"], remove_mean=True)
    assert np.allclose(S, d["S"], rtol=1e-12, atol=1e-12 * d["S"][0])
    k = int(d["k"])
    assert metrics.max_principal_angle(U[:, :k], d["U"], k) <= 1e-10


def test_randomized_with_reference_omega():
    d = load("randomized")
    Ui, S, V = oracle.randomized_svd([d["X"]], int(d["r"]), int(d["q"]), seed=int(d["seed"]))
    assert np.allclose(S, d["S"], rtol=1e-10)
    assert metrics.max_principal_angle(np.concatenate(Ui)[:, :8], d["U"][:, :8], 8) <= 1e-8


def test_spod_blocks_and_energies():
    d = load("spod")
    Q, f, nB = oracle.spod_blocks(d["X"], d["t"], int(d["nDFT"]), int(d["nolap"]))
    assert np.array_equal(f, d["f"])
    assert np.abs(Q - d["Q"]).max() <= 1e-6 * np.abs(d["Q"]).max()
    L, P, fs = oracle.spod_run(d["X"].astype(np.float64), d["t"], int(d["nDFT"]), int(d["nolap"]))
    assert np.allclose(L, d["L"], rtol=1e-9, atol=1e-12 * d["L"].max())
    assert np.array_equal(fs, d["f_sorted"])
    # the waves sit at 12.5 and 25 Hz (bins 8 and 16 of nDFT = 64, dt = 0.01);
    # the Hamming window leaks (0.23/0.54)^2 = 0.18 of the strong wave's energy
    # into bins 7 and 9, more than the 0.4^2 = 0.16 of the weak wave
    assert list(fs[:4]) == [12.5, 10.9375, 14.0625, 25.0]


def test_truncation_cases():
    d = load("truncation")
    assert [oracle.compute_truncation_residual(d["S"], r) for r in d["r"]] == list(d["N"])
    # S = [4,3,2,1]: r = -0.5 -> cumulative 0.4, 0.7 -> N = 2; r = 0.3 -> N = 3
    assert list(d["N"][:2]) == [2, 3]


def test_averaging_cases():
    d = load("averaging")
    assert np.array_equal(oracle.temporal_mean(d["X"]), d["mean"])
    assert np.array_equal(d["mean"], [3.0, 2.0, 0.0])
    assert np.allclose(oracle.temporal_variance(d["X"], d["mean"]), d["var"], rtol=0, atol=0)
    assert np.allclose(d["normvar"], (d["X"] - d["mean"][:, None]) / d["var"])

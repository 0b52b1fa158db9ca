"""Seeded synthetic inputs of the five POD workloads, numpy form (test infra).

bench.py generates the full-size inputs on the device with torch (its
generate / waves_device / uneven_problems / pool_device); these numpy twins
build the small and scaled-down cases that tests compare against the oracle.
"""
from __future__ import annotations

import numpy as np


def sines(m=20000, n=200, seed=1, dtype=np.float64):
    """Ten travelling sines + N(0, 1e-3) noise on x in [0,1), t_j = 0.05 j."""
    rng = np.random.default_rng(seed)
    x = np.arange(m) / m
    t = 0.05 * np.arange(n)
    om = rng.uniform(0.5, 5.0, 10)
    ph = rng.uniform(0.0, 2 * np.pi, 10)
    X = np.zeros((m, n))
    for k in range(10):
        X += 10 * 0.7 ** k * np.sin(2 * np.pi * (k + 1) * x[:, None] + om[k] * t[None, :] + ph[k])
    X += 1e-3 * rng.standard_normal((m, n))
    return X.astype(dtype)


def lowrank(m, n, seed=2, decay=0.93, top=1e3, noise=1e-6, dtype=np.float64, rank=None):
    """X = G1 diag(s) G2^T + noise with G1 ~ N(0,1)/sqrt(m), G2 orthonormal."""
    rng = np.random.default_rng(seed)
    k = n if rank is None else rank
    G1 = rng.standard_normal((m, k)) / np.sqrt(m)
    G2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = top * decay ** np.arange(k)
    X = (G1 * s) @ G2[:, :k].T
    if noise:
        X += noise * rng.standard_normal((m, n))
    return X.astype(dtype)


def waves(m=2000, n=512, seed=4, dtype=np.float32, noise=1e-2):
    """Six superposed travelling waves + N(0, noise) (SPOD workload shape)."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, m, endpoint=False)
    t = np.arange(n) * 0.01
    X = np.zeros((m, n))
    amp = [1.0, 0.7, 0.5, 0.35, 0.25, 0.15]
    kap = [6.0, 11.0, 17.0, 23.0, 31.0, 40.0]
    Om = [2 * np.pi * 7.8125, 2 * np.pi * 15.625, 2 * np.pi * 23.4375,
          2 * np.pi * 31.25, 2 * np.pi * 39.0625, 2 * np.pi * 46.875]
    for a, kk, w in zip(amp, kap, Om):
        X += a * np.sin(2 * np.pi * kk * x[:, None] - w * t[None, :] + rng.uniform(0, 2 * np.pi))
    X += noise * rng.standard_normal((m, n))
    return X.astype(dtype), t

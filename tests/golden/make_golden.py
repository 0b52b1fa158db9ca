"""Writes the golden fixtures of the POD core (tests/golden/*.npz).

Each fixture holds seeded inputs and the CPU oracle's outputs (oracle/, the
numpy restatement of pyLOM's algorithm); the reference itself cannot run here
(no mpi4py / h5py / MPI / LAPACKE), so its own numbers are not available.
What the fixtures pin:

  example_svd   Examples/example_SVD.py:17, A = [[1,2],[3,4],[5,6],[7,8]]:
                sigma = 14.2690955, 0.6268282 (the example's printed answer)
  pod_sines     the 20000 x 200 sine workload's formula, downscaled to
                600 x 30 (x_i = i/600, t_j = 0.05 j, 10 sines, N(0,1e-3)),
                POD.run(remove_mean=True): S, the gated U columns, V rows
  randomized    randomized_svd(r=12, q=3) with the reference's Omega
                (np.random.seed(seed); rand(n, r), pyLOM/vmmath/svd.py:131-133)
  spod          SPOD.run of two travelling waves + noise (M=120, N=192,
                nDFT=64, nolap=32): the Welch blocks Q (complex64), L, f
  truncation    compute_truncation_residual on hand-checkable S and r
  averaging     temporal_mean / temporal_variance / norm_variance on a
                small integer matrix

Run:  python tests/golden/make_golden.py   (rewrites the .npz files)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def sines(m, n, seed=1):
    """X_ij = sum_{k<10} 10 0.7^k sin(2 pi (k+1) x_i + w_k t_j + p_k) + N(0, 1e-3)."""
    rng = np.random.default_rng(seed)
    x = np.arange(m) / m
    t = 0.05 * np.arange(n)
    w = rng.uniform(0.5, 5.0, 10)
    p = rng.uniform(0.0, 2 * np.pi, 10)
    X = np.zeros((m, n))
    for k in range(10):
        X += 10 * 0.7 ** k * np.sin(2 * np.pi * (k + 1) * x[:, None] + w[k] * t[None, :] + p[k])
    return X + 1e-3 * rng.standard_normal((m, n))


def main():
    out = {}
    A = np.array([[1, 2], [3, 4], [5, 6], [7, 8]], dtype=np.double)
    _, S, _ = oracle.svd(A)
    out["example_svd"] = dict(A=A, S=S)

    X = sines(600, 30)
    U, S, V = oracle.pod_run(X, remove_mean=True)
    k = int(np.sum(S >= 1e-3 * S[0]))
    out["pod_sines"] = dict(X=X, S=S, U=U[:, :k], V=V[:k], k=np.int64(k))

    rng = np.random.default_rng(3)
    Xr = (rng.standard_normal((600, 8)) * (10.0 ** -np.arange(8))) @ rng.standard_normal((8, 24))
    Xr += 1e-6 * rng.standard_normal(Xr.shape)
    seed, r, q = 5, 12, 3
    Ui, S, V = oracle.randomized_svd([Xr], r, q, seed=seed)
    U = np.concatenate(Ui)
    out["randomized"] = dict(X=Xr, S=S, U=U, V=V, r=np.int64(r), q=np.int64(q), seed=np.int64(seed))

    M, N = 120, 192
    rng = np.random.default_rng(4)
    xs = np.linspace(0, 1, M, endpoint=False)[:, None]
    ts = np.arange(N) * 0.01
    Xs = (np.sin(2 * np.pi * 4 * xs - 2 * np.pi * 12.5 * ts) + 0.4 * np.sin(2 * np.pi * 9 * xs - 2 * np.pi * 25.0 * ts)
          + 0.05 * rng.standard_normal((M, N))).astype(np.float32)
    Q, f, nB = oracle.spod_blocks(Xs, ts, 64, 32)
    L, P, fs = oracle.spod_run(Xs.astype(np.float64), ts, 64, 32)
    out["spod"] = dict(X=Xs, t=ts, nDFT=np.int64(64), nolap=np.int64(32), Q=Q.astype(np.complex64), f=f,
                       L=L, f_sorted=fs)

    Ss = np.array([4.0, 3.0, 2.0, 1.0])
    rs = np.array([-0.5, 0.3, 2.0, 0.99, -0.99, 1e-8])
    Ns = np.array([oracle.compute_truncation_residual(Ss, r_) for r_ in rs], dtype=np.int64)
    out["truncation"] = dict(S=Ss, r=rs, N=Ns)

    Xa = np.array([[1.0, 2.0, 3.0, 6.0], [0.0, 0.0, 4.0, 4.0], [-2.0, 2.0, -2.0, 2.0]])
    mu = oracle.temporal_mean(Xa)
    var = oracle.temporal_variance(Xa, mu)
    out["averaging"] = dict(X=Xa, mean=mu, var=var, normvar=oracle.norm_variance(Xa, mu, var))

    for name, d in out.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(name, {k_: getattr(v, "shape", ()) for k_, v in d.items()})


if __name__ == "__main__":
    main()

"""CPU baseline of the tall-skinny SVD (test / measurement infrastructure).

Only bench.py's cpu_baseline leg uses this file.  Two variants of the
reference's algorithm (pyLOM/POD/wrapper.py:15-51 with
pyLOM/vmmath/svd.py:49-118, 227-252), both on the same float32 rows:

  * mpi: an MPI-rank stand-in.  P worker processes (one BLAS thread each), the
    rows split as pyLOM/utils/parall.py:24-48 splits them, in shared memory.
    Each rank removes its rows' temporal mean and takes the leaf QR of its
    block; the butterfly R-tree of svd.py:58-115 runs on the R factors (the
    oracle's restatement); svd(R); each rank forms U_i = Q1_i QW_i U_R.
  * threaded: one process, threaded BLAS, the oracle's one-rank path.

Each variant is timed as the median of ``runs`` runs after one warm-up.
"""
from __future__ import annotations

import multiprocessing as mp
import os
import statistics
import time
from multiprocessing import shared_memory

import numpy as np


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def usable_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _leaf(args):
    name_x, name_q, shape, lo, hi = args
    shx = shared_memory.SharedMemory(name=name_x)
    shq = shared_memory.SharedMemory(name=name_q)
    X = np.ndarray(shape, dtype=np.float32, buffer=shx.buf)
    Q = np.ndarray(shape, dtype=np.float32, buffer=shq.buf)
    Y = X[lo:hi] - X[lo:hi].mean(axis=1, keepdims=True)
    q, r = np.linalg.qr(Y)
    Q[lo:hi] = q
    del X, Q
    shx.close()
    shq.close()
    return r


def _form_u(args):
    name_q, shape, lo, hi, W = args
    shq = shared_memory.SharedMemory(name=name_q)
    Q = np.ndarray(shape, dtype=np.float32, buffer=shq.buf)
    Q[lo:hi] = Q[lo:hi] @ W
    del Q
    shq.close()
    return hi - lo


_LIMIT = None


def _init_worker():
    # a forked worker inherits the parent's BLAS thread pool: limit it to one
    # thread, as one MPI rank per core runs
    global _LIMIT
    from threadpoolctl import threadpool_limits
    _LIMIT = threadpool_limits(1)


def _tree(Rs):
    """Butterfly R-tree of pyLOM/vmmath/svd.py:58-115 on the ranks' R factors:
    the final R and, per rank, its n x n block QW_i of the tree's Q."""
    from .pod_oracle import tsqr
    n = Rs[0].shape[1]
    # tsqr of the stacked R's simulated rank by rank: the leaf QR of an R is
    # R itself up to signs, so the tree runs on them exactly as on the ranks
    Qs, R = tsqr([r.astype(np.float64) for r in Rs])
    return R, [q[:n].astype(np.float32) for q in Qs]


def mpi_standin(X: np.ndarray, nproc: int, runs: int = 5):
    """Median seconds of POD.run(remove_mean=True) on float32 X with nproc ranks."""
    m, n = X.shape
    from .pod_oracle import worksplit
    bounds = [worksplit(0, m, r, nproc) for r in range(nproc)]
    shx = shared_memory.SharedMemory(create=True, size=X.nbytes)
    shq = shared_memory.SharedMemory(create=True, size=X.nbytes)
    try:
        np.ndarray(X.shape, dtype=np.float32, buffer=shx.buf)[:] = X
        ctx = mp.get_context("fork")
        times = []
        with ctx.Pool(nproc, initializer=_init_worker) as pool:
            for it in range(runs + 1):
                t0 = time.perf_counter()
                Rs = pool.map(_leaf, [(shx.name, shq.name, X.shape, lo, hi) for lo, hi in bounds])
                R, QW = _tree(Rs)
                Ur, S, V = np.linalg.svd(R.astype(np.float32), full_matrices=False)
                pool.map(_form_u, [(shq.name, X.shape, lo, hi, QW[i] @ Ur) for i, (lo, hi) in enumerate(bounds)])
                if it:
                    times.append(time.perf_counter() - t0)
        return statistics.median(times), times
    finally:
        shx.close()
        shx.unlink()
        shq.close()
        shq.unlink()


def threaded(X: np.ndarray, runs: int = 5):
    """Median seconds of the oracle's one-rank POD.run on float32 X."""
    from .pod_oracle import pod_run
    times = []
    for it in range(runs + 1):
        t0 = time.perf_counter()
        pod_run(X, remove_mean=True)
        if it:
            times.append(time.perf_counter() - t0)
    return statistics.median(times), times

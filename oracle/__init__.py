"""CPU oracle for the POD core — TEST INFRASTRUCTURE ONLY.

This package is a numpy restatement of the reference's POD hot path
(pyLOM/vmmath/svd.py, averaging.py, maths.py, truncation.py, stats.py and
pyLOM/POD/wrapper.py, pyLOM/SPOD/wrapper.py).  Every function cites the
reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product package ``pyloworder_amd`` never imports it and has no
CPU fallback.

Pinning (how each function is checked, see tests/test_oracle.py):
  * svd / tsqr_svd / qr / tsqr: the known-answer 4x2 case of
    Examples/example_SVD.py:17 (sigma = 14.2690955, 0.6268282), numpy LAPACK
    SVD/QR (closed form up to signs), brute force of the simulated-rank
    butterfly tree against the one-block QR for P = 1..9 ranks (including
    non-powers of two), and the invariants A = QR, Q^T Q = I, R upper
    triangular.
  * randomized_qr / randomized_svd: Omega is numpy's legacy MT19937 stream
    (svd.py:131-133) which is frozen by numpy; exact-rank inputs are recovered
    exactly and the subspace matches the deterministic power iteration.
  * temporal_mean / subtract_mean / temporal_variance / norm_variance /
    matmul / matmulp / vecmat / reconstruct / RMSE / energy: closed forms
    checked on hand-computed integer cases.
  * compute_truncation_residual: hand-computed cases of the three r regimes.
  * init_qr_streaming / update_qr_streaming: the first equals randomized_qr
    plus its last sketch; the update reproduces a low-rank history exactly
    (Q2 B2 = [A1, A2] when rank(A1) + rank(A2) <= r, no power steps); a test
    shows the sketch depends on the QR sign convention (``positive`` selects
    diag(R) >= 0 instead of LAPACK's signs, which the reference and the
    device path both use).
  * spod_run: a single-frequency travelling wave whose spectrum peaks at the
    known bin, Parseval-type energy identities, and a hand-built 2-block case.
"""
from .pod_oracle import *  # noqa: F401,F403
from . import metrics  # noqa: F401

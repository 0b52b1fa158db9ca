"""Pins of the CPU oracle (oracle/) — every function is checked against a
known answer, a closed form, an invariant or brute force, chosen so that a
plausible slip (dropped term, wrong sign or index, transposed operand) fails.
"""
import numpy as np
import pytest
import scipy.fft
import scipy.linalg

import oracle
from oracle import metrics, workloads


# ---------------------------------------------------------------------------
# svd / tsqr / tsqr_svd
# ---------------------------------------------------------------------------
def test_known_answer_example_svd():
    # Examples/example_SVD.py:17: A = [[1,2],[3,4],[5,6],[7,8]]
    A = np.array([[1, 2], [3, 4], [5, 6], [7, 8]], dtype=np.double)
    _, S, _ = oracle.svd(A)
    assert np.allclose(S, [14.2690955, 0.6268282], atol=5e-8)
    for P in (1, 2):
        Ui, S2, V = oracle.tsqr_svd(oracle.split_rows(A, P))
        U = np.concatenate(Ui)
        assert np.allclose(S2, S, rtol=1e-14)
        assert np.allclose((U * S2) @ V, A, atol=1e-13)


@pytest.mark.parametrize("P", range(1, 10))
def test_tsqr_tree_equals_one_block_qr(P):
    rng = np.random.default_rng(P)
    A = rng.standard_normal((40 * P + 7, 6))
    Qi, R = oracle.tsqr(oracle.split_rows(A, P))
    Q = np.concatenate(Qi)
    _, R1 = np.linalg.qr(A)
    # R unique up to row signs; A = QR; Q orthonormal; R upper triangular
    assert np.allclose(np.abs(R), np.abs(R1), atol=1e-12)
    assert np.allclose(Q @ R, A, atol=1e-12)
    assert np.allclose(Q.T @ Q, np.eye(6), atol=1e-13)
    assert np.abs(np.tril(R, -1)).max() == 0.0
    # R^T R = A^T A (svd.py:85 "R is correct on processor 0")
    assert np.allclose(R.T @ R, A.T @ A, rtol=1e-12, atol=1e-10)


def test_tsqr_requires_tall_blocks():
    A = np.ones((10, 4))
    with pytest.raises(ValueError):
        oracle.tsqr(oracle.split_rows(A, 4))  # 3,3,2,2 rows < n


def test_worksplit_covers_rows_exactly():
    for m in (1, 7, 100, 1001):
        for P in (1, 2, 3, 5, 8):
            spans = [oracle.worksplit(0, m, r, P) for r in range(P)]
            if P < m:
                assert spans[0][0] == 0 and spans[-1][1] == m
                assert all(spans[i][1] == spans[i + 1][0] for i in range(P - 1))
                sizes = [b - a for a, b in spans]
                assert max(sizes) - min(sizes) <= 1


@pytest.mark.parametrize("P", [1, 3, 4])
def test_tsqr_svd_equals_lapack_svd(P):
    X = workloads.lowrank(3000, 40, seed=11, decay=0.8, noise=1e-5)
    Ui, S, V = oracle.tsqr_svd(oracle.split_rows(X, P))
    U = np.concatenate(Ui)
    Ur, Sr, Vr = np.linalg.svd(X, full_matrices=False)
    assert np.allclose(S, Sr, rtol=1e-12)
    assert metrics.max_principal_angle(U, Ur, 30) < 1e-10
    assert metrics.recon_rmse(X, U, S, V) < 1e-13


def test_tsqr_svd_complex_block():
    rng = np.random.default_rng(3)
    Q = (rng.standard_normal((500, 15)) + 1j * rng.standard_normal((500, 15))).astype(np.complex64)
    Ui, S, V = oracle.tsqr_svd([Q])
    Sr = np.linalg.svd(Q.astype(np.complex128), compute_uv=False)
    assert np.allclose(S, Sr, rtol=1e-5)
    assert metrics.recon_rmse(Q, Ui[0], S, V) < 1e-5


# ---------------------------------------------------------------------------
# randomized
# ---------------------------------------------------------------------------
def test_reference_omega_stream_is_numpy_legacy_mt19937():
    # np.random.seed(0); rand() -> 0.5488135039273248 (published MT19937 value)
    om = oracle.reference_omega(2, 2, 0, np.float64)
    assert np.allclose(om, [[0.5488135, 0.71518937], [0.60276338, 0.54488318]], atol=1e-8)
    om = oracle.reference_omega(3, 2, 25, np.float64)
    assert np.allclose(om, [[0.87012414, 0.58227693], [0.27883894, 0.18591123], [0.41110013, 0.11737555]], atol=1e-8)


def test_randomized_recovers_exact_low_rank():
    rng = np.random.default_rng(5)
    X = rng.standard_normal((2000, 6)) @ rng.standard_normal((6, 50))
    Ui, S, V = oracle.randomized_svd(oracle.split_rows(X, 2), 8, 2, seed=25)
    Sr = np.linalg.svd(X, compute_uv=False)
    assert np.allclose(S[:6], Sr[:6], rtol=1e-10)
    assert np.all(S[6:] < 1e-10 * S[0])
    U = np.concatenate(Ui)
    assert metrics.recon_rmse(X, U, S, V) < 1e-12


def test_randomized_power_iteration_subspace():
    # q power iterations: Q spans (A A^T)^q A Omega; compare with a direct computation
    X = workloads.lowrank(1500, 30, seed=9, decay=0.7, noise=0.0)
    om = oracle.reference_omega(30, 5, 7, np.float64)
    Qi, B = oracle.randomized_qr([X], 5, 2, omega=om)
    Y = X @ om
    for _ in range(2):
        Y = X @ (X.T @ Y)
    Qd, _ = np.linalg.qr(Y)
    assert metrics.max_principal_angle(Qi[0], Qd, 5) < 1e-8
    assert np.allclose(B, Qi[0].T @ X, atol=1e-10)


# ---------------------------------------------------------------------------
# averaging, maths, truncation, stats — hand-computed
# ---------------------------------------------------------------------------
def test_averaging_hand_cases():
    X = np.array([[1.0, 2.0, 3.0, 6.0], [2.0, 2.0, 2.0, 2.0], [0.0, -4.0, 4.0, 0.0]])
    mu = oracle.temporal_mean(X)
    assert np.array_equal(mu, [3.0, 2.0, 0.0])
    Y = oracle.subtract_mean(X, mu)
    assert np.array_equal(Y, [[-2, -1, 0, 3], [0, 0, 0, 0], [0, -4, 4, 0]])
    var = oracle.temporal_variance(X, mu)
    assert var.shape == (3, 1)
    assert np.allclose(var[:, 0], [14 / 4, 0.0, 32 / 4])
    Z = oracle.norm_variance(X[[0, 2]], mu[[0, 2]], var[[0, 2]])
    assert np.allclose(Z[0], np.array([-2, -1, 0, 3]) / 3.5)   # divides by the variance, not std


def test_matmulp_is_the_rank_sum():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((100, 5))
    B = rng.standard_normal((100, 3))
    parts = oracle.split_rows(np.hstack([A, B]), 4)
    C = oracle.matmulp([p[:, :5].T for p in parts], [p[:, 5:] for p in parts])
    assert np.allclose(C, A.T @ B, atol=1e-13)


def test_vecmat_and_reconstruct():
    v = np.array([2.0, -1.0])
    A = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    assert np.array_equal(oracle.vecmat(v, A), [[2, 4, 6], [-4, -5, -6]])
    U = np.eye(3)[:, :2]
    assert np.array_equal(oracle.pod_reconstruct(U, v, A), [[2, 4, 6], [-4, -5, -6], [0, 0, 0]])


@pytest.mark.parametrize("S,r,N", [
    ([4, 3, 2, 1], -0.5, 2), ([4, 3, 2, 1], 0.3, 3), ([4, 3, 2, 1], 0.5, 2), ([4, 3, 2, 1], -0.95, 4),
])
def test_truncation_hand_cases(S, r, N):
    assert oracle.compute_truncation_residual(np.array(S, float), r) == N
    U = np.eye(4)
    Ur, Sr, Vr = oracle.pod_truncate(U, np.array(S, float), U, r)
    assert Ur.shape == (4, N) and Sr.shape == (N,) and Vr.shape == (N, 4)


def test_rmse_energy_hand():
    A = np.array([[3.0, 4.0]])
    B = np.array([[3.0, 0.0]])
    assert oracle.RMSE(A, B) == pytest.approx(4 / 5)
    assert oracle.RMSE(A, B, relative=False) == pytest.approx(np.sqrt(16 / 2))
    assert oracle.energy(A, B) == pytest.approx(1 - 16 / 25)


def test_pod_run_variants():
    X = workloads.sines(2000, 60)
    U, S, V = oracle.pod_run(X, remove_mean=True, nranks=3)
    Y = X - X.mean(axis=1, keepdims=True)
    assert np.allclose(S, np.linalg.svd(Y, compute_uv=False), rtol=1e-10, atol=1e-10 * S[0])
    U2, S2, V2 = oracle.pod_run(X, remove_mean=False)
    assert np.allclose(S2, np.linalg.svd(X, compute_uv=False), rtol=1e-12)


# ---------------------------------------------------------------------------
# SPOD
# ---------------------------------------------------------------------------
def test_hammwin_closed_form():
    w = oracle.hammwin(5)
    assert np.allclose(w, [0.08, 0.54, 1.0, 0.54, 0.08])


def test_spod_single_wave_peaks_at_its_bin():
    M, N, nDFT = 64, 512, 64
    dt = 0.01
    t = np.arange(N) * dt
    f0 = 10 / (nDFT * dt)                     # bin 10 exactly
    x = np.linspace(0, 1, M, endpoint=False)
    X = np.sin(2 * np.pi * 3 * x[:, None] - 2 * np.pi * f0 * t[None, :])
    L, P, f = oracle.spod_run(X, t, nDFT=nDFT, nolap=32)
    assert np.isclose(f[0], f0)
    # Hamming leakage puts ((0.46/2)/0.54)^2 ~ 0.18 of the peak in the two neighbour bins
    assert L[0, 0] > 4 * L[1, 0]
    assert L.shape == (nDFT // 2 + 1, (N - 32) // (nDFT - 32))
    assert P.shape == (M * L.shape[1], L.shape[0])


def test_spod_blocks_match_direct_welch():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((20, 40))
    t = np.arange(40) * 0.1
    Q, f, nb = oracle.spod_blocks(X, t, nDFT=8, nolap=4)
    assert nb == 9
    Y = X - X.mean(axis=1, keepdims=True)
    w = oracle.hammwin(8)
    blk = Y[:, 4:12] * w
    q = (1 / np.mean(w) / 8) * np.fft.fft(blk, axis=1)[:, :5]
    q[:, 1:-1] *= 2
    assert np.allclose(Q[:, :, 1].T, q / np.sqrt(9), atol=1e-6)
    assert Q.shape == (5, 20, 9)
    # energy per frequency equals the sum of squared singular values (Frobenius)
    L, P, fo = oracle.spod_run(X, t, nDFT=8, nolap=4)
    tot = np.sort([np.sum(np.abs(Q[i]) ** 2) for i in range(5)])
    assert np.allclose(np.sort(L.sum(axis=1)), tot, rtol=1e-4)


# ---------------------------------------------------------------------------
# metrics (used by every parity test)
# ---------------------------------------------------------------------------
def test_metrics_detect_planted_errors():
    X = workloads.lowrank(500, 10, seed=1, decay=0.5, noise=0.0)
    U, S, V = np.linalg.svd(X, full_matrices=False)
    assert metrics.sigma_errors(S, S)["maxrel_all"] == 0.0
    S2 = S.copy(); S2[3] *= 1 + 1e-6
    assert metrics.sigma_errors(S2, S)["maxrel_gated"] == pytest.approx(1e-6, rel=1e-3)
    U2 = U.copy(); U2[:, [2, 5]] = U2[:, [5, 2]]       # swapped modes: large angle on k=3
    assert metrics.max_principal_angle(U2, U, 3) > 0.5
    assert metrics.max_principal_angle(-U, U, 10) < 1e-12   # signs do not matter
    assert metrics.recon_rmse(X, U, S, V) < 1e-14
    assert metrics.recon_rmse(X, U, S, V.T) > 1e-2          # transposed operand is caught


# ---------------------------------------------------------------------------
# streaming randomized QR (pyLOM/vmmath/svd.py:175-225)
# ---------------------------------------------------------------------------
def test_init_qr_streaming_is_randomized_qr_plus_sketch():
    rng = np.random.default_rng(21)
    A = rng.standard_normal((300, 20)) @ np.diag(np.logspace(0, -3, 20))
    Q, B, Y = oracle.init_qr_streaming([A], 6, 2, seed=4)
    Q0, B0 = oracle.randomized_qr([A], 6, 2, seed=4)
    assert np.array_equal(Q[0], Q0[0]) and np.array_equal(B, B0)
    # the returned sketch is the last power iterate, and Q spans it
    assert np.allclose(Q[0] @ (Q[0].T @ Y[0]), Y[0], atol=1e-10 * np.abs(Y[0]).max())


def test_update_qr_streaming_recovers_low_rank_history():
    """rank(A1) + rank(A2) <= r, no power steps: the updated basis reproduces
    every snapshot seen so far, Q2 B2 = [A1, A2], whatever the sign
    convention of the QR."""
    rng = np.random.default_rng(22)
    A1 = rng.standard_normal((400, 5)) @ rng.standard_normal((5, 30))
    A2 = rng.standard_normal((400, 4)) @ rng.standard_normal((4, 25))
    for positive in (False, True):
        Q, B, Y = oracle.init_qr_streaming([A1], 12, 0, seed=3, positive=positive)
        Yo = [Y[0].copy()]
        Q2, B2, Yo2 = oracle.update_qr_streaming([A2], Q, B, Yo, 12, 0, omega=np.random.default_rng(5).random((25, 12)),
                                                 positive=positive)
        X = np.hstack((A1, A2))
        assert np.abs(Q2[0] @ B2 - X).max() <= 1e-12 * np.abs(X).max()
        assert Yo2 is Yo and B2.shape == (12, 55)  # Yo is updated in place
        assert np.allclose(Q2[0].T @ Q2[0], np.eye(12), atol=1e-12)


def test_update_qr_streaming_sign_convention_changes_the_sketch():
    """Yo accumulates (A A^T)^q A Omega whose columns carry the QR's signs: the
    two conventions give different sketches (the reason the device tsqr
    reproduces LAPACK's signs, and the streaming path is pinned elementwise)."""
    rng = np.random.default_rng(23)
    A1 = rng.standard_normal((200, 10))
    A2 = rng.standard_normal((200, 8))
    om1 = rng.random((10, 4)); om2 = rng.random((8, 4))
    outs = []
    for positive in (False, True):
        Q, B, Y = oracle.init_qr_streaming([A1], 4, 1, omega=om1, positive=positive)
        _, _, Yo = oracle.update_qr_streaming([A2], Q, B, [Y[0].copy()], 4, 1, omega=om2, positive=positive)
        outs.append(Yo[0])
    assert not np.allclose(outs[0], outs[1])


def _dynamics(m=400, n=60, seed=3):
    """Snapshots of exact linear dynamics: two oscillating pairs and one
    decaying real mode, X[:, t] = sum_k Re(a_k phi_k mu_k^t) (rank 5)."""
    rng = np.random.default_rng(seed)
    mus = [0.98 * np.exp(0.3j), 0.9 * np.exp(1.1j), 0.7]
    amps = [3.0, 1.5, 0.8]
    X = np.zeros((m, n))
    t = np.arange(n)
    for mu, a in zip(mus, amps):
        phi = rng.standard_normal(m) + (1j * rng.standard_normal(m) if np.iscomplex(mu) else 0)
        X += np.real(a * np.outer(phi, mu ** t))
    return X, mus


def test_dmd_recovers_known_eigenvalues_and_reconstructs():
    """Exact rank-5 dynamics: DMD's eigenvalues are the generating ones
    (conjugate pairs included), and Re(Phi diag(b) Vand) reproduces every
    snapshot used (Jovanovic amplitudes minimise that residual, here 0)."""
    X, mus = _dynamics()
    muR, muI, Phi, b = oracle.dmd_run(X, 5, remove_mean=False)
    got = np.sort_complex(muR + 1j * muI)
    want = np.sort_complex(np.array([mus[0], np.conj(mus[0]), mus[1], np.conj(mus[1]), mus[2]]))
    assert np.abs(got - want).max() < 1e-10
    Xr = oracle.dmd_reconstruction_jovanovic(Phi, muR, muI, np.arange(X.shape[1] - 1), b)
    assert np.abs(Xr - X[:, :-1]).max() < 1e-9 * np.abs(X).max()
    # ordering: |b| descending, each conjugate pair positive imaginary part first
    assert np.all(np.diff(np.abs(b)) <= 1e-9 * np.abs(b).max())
    for i in range(len(muI) - 1):
        if muI[i] != 0 and abs(muR[i] - muR[i + 1]) < 1e-12:
            assert muI[i] > 0 and abs(muI[i] + muI[i + 1]) < 1e-12


def test_dmd_order_modes_hand_case():
    """pyLOM/DMD/wrapper.py:18-46 on a hand case: |b| = 1, 5, 5, 3 with the
    pair listed negative-imaginary first."""
    muR = np.array([0.1, 0.5, 0.5, 0.9])
    muI = np.array([0.0, -0.2, 0.2, 0.0])
    b = np.array([1.0, 3 - 4j, 3 + 4j, 3.0])
    Phi = np.array([[1.0, 1 - 1j, 1 + 1j, 2.0]])
    r, i, P, bb = oracle.dmd_order_modes(muR, muI, Phi, b)
    assert np.allclose(r, [0.5, 0.5, 0.9, 0.1]) and np.allclose(i, [0.2, -0.2, 0.0, 0.0])
    assert np.allclose(bb, [3 + 4j, 3 - 4j, 3.0, 1.0]) and np.allclose(P, [[1 + 1j, 1 - 1j, 2.0, 1.0]])


def test_dmd_frequency_damping_closed_form():
    delta, omega, dt = np.array([-0.1, 0.0, 0.05]), np.array([2.0, -1.0, 0.5]), 0.01
    mu = np.exp((delta + 1j * omega) * dt)
    d, w = oracle.dmd_frequency_damping(mu.real, mu.imag, dt)
    assert np.allclose(d, delta, atol=1e-12) and np.allclose(w, omega, atol=1e-12)


def test_dataset_x_hand_case():
    """pyLOM/dataset.py:252-292: two points, a scalar field P and a 2-component
    field U (u, v per point), X rows = p0:(P, u, v), p1:(P, u, v)."""
    P = np.array([[1.0, 2.0], [3.0, 4.0]])                       # points 0, 1
    U = np.array([[10.0, 11.0], [20.0, 21.0], [30.0, 31.0], [40.0, 41.0]])  # (u0, v0, u1, v1)
    X = oracle.dataset_x({"P": (1, P), "U": (2, U)}, 2)
    assert np.array_equal(X, [[1, 2], [10, 11], [20, 21], [3, 4], [30, 31], [40, 41]])
    X2 = oracle.dataset_x({"P": (1, P), "U": (2, U)}, 2, names=["U"])
    assert np.array_equal(X2, U)

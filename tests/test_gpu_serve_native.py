"""Native request loop (pc_serve_*) on the GPU.

Tolerance: bitwise equality (0) with the direct calls.  Each slot has its own
stream and library context, every kernel sums in an order fixed by the problem
shape alone and none uses floating-point atomics, so a request's result does
not depend on what else the server runs.  The direct calls themselves are
checked against the oracle in test_gpu_core / test_gpu_mixed.
"""
import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

pl = pytest.importorskip("pyloworder_amd")
from pyloworder_amd import POD as P  # noqa: E402
from pyloworder_amd import _lib  # noqa: E402
from pyloworder_amd.serve import NativeServer  # noqa: E402


def _pool(m, n, seed, decay, dtype=torch.float32):
    g = torch.Generator(device="cuda").manual_seed(seed)
    G2 = torch.linalg.qr(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))[0]
    s = 1e3 * decay ** torch.arange(n, device="cuda", dtype=torch.float64)
    G1 = torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64) / np.sqrt(m)
    X = (G1 @ (s[:, None] * G2.T)).to(dtype)
    return X + 1e-6 * torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64).to(dtype)


SHAPES = [(200_000, 64), (50_000, 200), (300_000, 17), (120_000, 128), (80_000, 8), (250_000, 256), (60_000, 300)]


def _direct_pod(X, precision):
    prec = precision if (X.dtype == torch.float32 and X.shape[1] <= 256) else "fp64"
    U, S, V = P.run(X, remove_mean=True, precision=prec)
    torch.cuda.synchronize()
    return U, S, V


@pytest.fixture(scope="module")
def pools():
    return [_pool(300_000, 300, 5, 0.95), _pool(300_000, 300, 6, 0.92)]


def test_native_pod_project_reconstruct_equal_direct(pools):
    with NativeServer(slots=3, budget=40e9, m_max=300_000, n_max=300) as srv:
        pods = []
        for i, (m, n) in enumerate(SHAPES):
            X = pools[i % 2][:m, :n]
            pods.append((X, srv.submit_pod(X)))
        assert srv.drain(timeout=120)
        follow = []
        for i, (X, t) in enumerate(pods):
            rec = srv.wait(t)
            assert rec["status"] == 0, rec
            U, S, VT = srv.basis(rec["basis"])
            Ud, Sd, Vd = _direct_pod(X, "mixed")
            assert torch.equal(U, Ud) and torch.equal(S, Sd) and torch.equal(VT, Vd), (X.shape, i)
            m, n = X.shape
            Y = pools[(i + 1) % 2][:m, : max(4, n // 2)]
            k = min(n, 10)
            follow.append((srv.submit_project(rec["basis"], Y, modes=k), "p", (U, S, VT, Y, k)))
            follow.append((srv.submit_reconstruct(rec["basis"], modes=min(n, 6)), "r", (U, S, VT, None, min(n, 6))))
        for t, kind, (U, S, VT, Y, k) in follow:
            got = srv.output(t, U.dtype)
            if kind == "p":
                want = P.project(U[:, :k], Y, precision="fp64")
            else:
                want = P.reconstruct(U[:, :k], S[:k], VT[:k, :], precision="fp64")
            torch.cuda.synchronize()
            assert torch.equal(got, want), (kind, U.shape, k)
        st = srv.stats()
        assert st["completed"] == 3 * len(SHAPES) and st["rejected"] == 0 and st["queued"] == 0
        assert st["resident_bytes"] > 0 and st["inflight_bytes"] == 0


def test_native_concurrent_submitters_equal_alone(pools):
    """Requests submitted from three threads at once give the lone results."""
    shapes = SHAPES[:6] * 2
    with NativeServer(slots=4, budget=40e9, m_max=300_000, n_max=300) as srv:
        tickets = [None] * len(shapes)

        def feed(idx):
            for i in idx:
                m, n = shapes[i]
                tickets[i] = srv.submit_pod(pools[i % 2][:m, :n])

        th = [threading.Thread(target=feed, args=(list(range(j, len(shapes), 3)),)) for j in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert srv.drain(timeout=120)
        for i, (m, n) in enumerate(shapes):
            rec = srv.wait(tickets[i])
            assert rec["status"] == 0
            U, S, VT = srv.basis(rec["basis"])
            Ud, Sd, Vd = _direct_pod(pools[i % 2][:m, :n], "mixed")
            assert torch.equal(U, Ud) and torch.equal(S, Sd) and torch.equal(VT, Vd)
            assert rec["t_admit"] >= rec["t_submit"] and rec["t_done"] >= rec["t_admit"]


def test_native_fp64_server_and_float64_inputs(pools):
    X64 = pools[0][:40_000, :48].double()
    X32 = pools[1][:40_000, :48]
    with NativeServer(slots=2, budget=10e9, precision="fp64", m_max=40_000, n_max=64) as srv:
        t64, t32 = srv.submit_pod(X64), srv.submit_pod(X32)
        for X, t in ((X64, t64), (X32, t32)):
            rec = srv.wait(t, timeout=60)
            assert rec is not None and rec["status"] == 0
            U, S, VT = srv.basis(rec["basis"])
            Ud, Sd, Vd = P.run(X, remove_mean=True, precision="fp64")
            assert torch.equal(U, Ud) and torch.equal(S, Sd) and torch.equal(VT, Vd)


def _workspaces(slots, m_max, n_max):
    """Bytes a server of this shape reserves for its slots' workspaces."""
    with NativeServer(slots=slots, budget=20e9, m_max=m_max, n_max=n_max) as probe:
        return int(20e9) - probe.stats()["output_budget"]


def test_native_admission_eviction_and_errors(pools):
    m, n = 100_000, 64
    one = m * n * 4 + n * n * 4 + n * 4          # bytes of one basis
    ws = _workspaces(2, 300_000, 300)
    with pytest.raises(_lib.PodCoreError):       # the workspaces alone exceed the budget
        NativeServer(slots=2, budget=ws // 2, m_max=300_000, n_max=300)
    with NativeServer(slots=2, budget=ws + int(2.5 * one), skip_ahead=2, m_max=300_000, n_max=300) as srv:
        ts = [srv.submit_pod(pools[i % 2][:m, :n]) for i in range(4)]
        assert srv.drain(timeout=120)
        recs = [srv.wait(t) for t in ts]
        assert all(r["status"] == 0 for r in recs)
        st = srv.stats()
        assert st["output_budget"] == int(2.5 * one)
        assert st["evictions"] >= 2 and st["bases"] <= 2 and st["resident_bytes"] <= 2.5 * one
        # the oldest basis was evicted: a projection onto it fails, onto the newest runs
        bad = srv.submit_project(recs[0]["basis"], pools[0][:m, :8])
        assert srv.wait(bad)["status"] == _lib.STATUS_CODE["PC_ERR_INVALID"]
        good = srv.submit_project(recs[-1]["basis"], pools[0][:m, :8])
        assert srv.wait(good)["status"] == 0
        srv.release(good)
        # outputs larger than the whole budget are rejected, not queued
        big = srv.submit_pod(pools[0][:300_000, :300])
        assert srv.wait(big, timeout=60)["status"] == _lib.STATUS_CODE["PC_ERR_NOMEM"]
        # shapes beyond the server's limits are refused at submission
        with pytest.raises(_lib.PodCoreError):
            srv.submit_pod(torch.zeros((400_000, 8), device="cuda"))
        with pytest.raises(_lib.PodCoreError):
            srv.submit_pod(torch.zeros((4, 8), device="cuda"))
        assert srv.stats()["rejected"] == 2


def test_native_pinned_basis_survives_eviction(pools):
    m, n = 100_000, 64
    one = m * n * 4 + n * n * 4 + n * 4
    with NativeServer(slots=1, budget=_workspaces(1, 300_000, 300) + int(1.5 * one), m_max=300_000,
                      n_max=300) as srv:
        first = srv.wait(srv.submit_pod(pools[0][:m, :n]))
        lib, h = srv.lib, srv.h
        assert lib.pc_serve_basis(h, first["basis"], 1, None) == 0       # pin
        second = srv.submit_pod(pools[1][:m, :n])                         # needs the space
        assert srv.wait(second, timeout=2.0) is None                      # waits behind the pin
        assert lib.pc_serve_basis(h, first["basis"], -1, None) == 0      # unpin: admitted now
        rec = srv.wait(second, timeout=60)
        assert rec is not None and rec["status"] == 0
        assert srv.stats()["evictions"] == 1


def _slow_input(m, n, seed):
    """X produced on a side stream by a chain of products that takes some
    milliseconds, returned while that stream may still be running."""
    side = torch.cuda.Stream()
    g = torch.Generator(device="cuda").manual_seed(seed)
    with torch.cuda.stream(side):
        A = torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64)
        B = torch.linalg.qr(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))[0]
        for _ in range(40):
            A = A @ B
        X = A.float()
    return side, X


def test_native_request_waits_for_its_input_stream():
    """A request submitted while its input is still being produced on another
    stream reads the finished input (the slot waits on the submitter's stream)."""
    with NativeServer(slots=2, budget=20e9, m_max=400_000, n_max=64) as srv:
        side, X = _slow_input(400_000, 64, 11)
        with torch.cuda.stream(side):
            t = srv.submit_pod(X)
        rec = srv.wait(t, timeout=60)
        assert rec is not None and rec["status"] == 0, rec
        U, S, VT = srv.basis(rec["basis"])
        side.synchronize()
        Ud, Sd, Vd = _direct_pod(X, "mixed")
        assert torch.equal(U, Ud) and torch.equal(S, Sd) and torch.equal(VT, Vd)

"""Serving loop on the GPU: every result is the same alone and in the mix.

Tolerance: bitwise equality (0).  Each slot has its own stream and library
context (own workspace), every kernel sums in an order fixed by the problem
shape alone and none uses floating-point atomics, so concurrency cannot
change a result.  Results are also checked against the oracle.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import metrics

pytestmark = pytest.mark.gpu

pl = pytest.importorskip("pyloworder_amd")
from pyloworder_amd import serve  # noqa: E402
from pyloworder_amd.serve import POD, PROJECT, RECONSTRUCT, Request, Server  # noqa: E402


def _pool(m, n, seed, decay):
    g = torch.Generator(device="cuda").manual_seed(seed)
    G2 = torch.linalg.qr(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))[0]
    s = 1e3 * decay ** torch.arange(n, device="cuda", dtype=torch.float64)
    G1 = torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64) / np.sqrt(m)
    X = (G1 @ (s[:, None] * G2.T)).float()
    return X + 1e-6 * torch.randn(m, n, generator=g, device="cuda", dtype=torch.float32)


SHAPES = [(200_000, 64), (50_000, 200), (300_000, 17), (120_000, 128), (80_000, 8), (250_000, 256)]


def _requests(pools):
    reqs = [Request(POD, X=pools[i % 2][:m, :n]) for i, (m, n) in enumerate(SHAPES)]
    return reqs


def _follow_ups(pools, pod_reqs, start=0):
    out = []
    for i, r in enumerate(pod_reqs, start):
        m, n = r.X.shape
        out.append(Request(PROJECT, X=pools[(i + 1) % 2][:m, : max(4, n // 2)], basis=r.basis, modes=min(n, 10)))
        out.append(Request(RECONSTRUCT, basis=r.basis, modes=min(n, 6)))
    return out


def _run(server, reqs):
    for r in reqs:
        server.submit(r)
    assert server.drain(timeout=120)
    for r in reqs:
        assert r.error is None, r.error


def _same(a, b):
    if isinstance(a, tuple):
        return all(_same(x, y) for x, y in zip(a, b))
    return torch.equal(a, b)


def test_alone_equals_mix():
    pools = [_pool(300_000, 256, 1, 0.93), _pool(300_000, 256, 2, 0.96)]
    # alone: one slot, each request submitted only after the previous finished
    alone_pod = _requests(pools)
    with Server(serve.DeviceExecutor("mixed"), slots=1, budget=60e9) as s:
        for r in alone_pod:
            _run(s, [r])
        alone_fu = _follow_ups(pools, alone_pod)
        for r in alone_fu:
            _run(s, [r])
    # mix: four slots, everything submitted at once; the follow-ups of the
    # first half are queued while the second half's solves are still running
    mix_pod = _requests(pools)
    with Server(serve.DeviceExecutor("mixed"), slots=4, budget=60e9) as s:
        for r in mix_pod[:3]:
            s.submit(r)
        assert s.drain(timeout=120)
        for r in mix_pod[3:]:
            s.submit(r)
        fu_a = _follow_ups(pools, mix_pod[:3])
        for r in fu_a:
            s.submit(r)
        assert s.drain(timeout=120)
        fu_b = _follow_ups(pools, mix_pod[3:], start=3)
        _run(s, fu_b)
        overlap = max(len([q for q in s.completed if q.admitted < r.done and q.done > r.admitted and q is not r])
                      for r in s.completed)
    mix_fu = fu_a + fu_b
    assert overlap >= 1  # requests really ran concurrently
    for a, b in zip(alone_pod, mix_pod):
        assert _same(a.result, b.result), a.X.shape
    for a, b in zip(alone_fu, mix_fu):
        assert a.kind == b.kind and _same(a.result, b.result), (a.kind, a.modes)


def test_served_results_match_oracle():
    X = _pool(60_000, 40, 3, 0.9)
    with Server(serve.DeviceExecutor("mixed"), slots=2, budget=20e9) as s:
        r = s.submit(Request(POD, X=X))
        assert s.drain(timeout=60)
        p = s.submit(Request(PROJECT, X=X, basis=r.basis, modes=5))
        q = s.submit(Request(RECONSTRUCT, basis=r.basis))
        assert s.drain(timeout=60)
    U, S, V = (t.double().cpu().numpy() for t in r.result)
    Xn = X.double().cpu().numpy()
    Uo, So, Vo = oracle.pod_run(Xn, remove_mean=True)
    e = metrics.sigma_errors(S, So)
    assert e["maxrel_gated"] <= 1e-5
    assert metrics.max_principal_angle(U, Uo, e["n_gated"]) <= 1e-4
    C = p.result.double().cpu().numpy()
    assert np.allclose(C, U[:, :5].T @ Xn, rtol=0, atol=1e-5 * np.abs(Xn).max() * np.sqrt(Xn.shape[0]))
    Xr = q.result.double().cpu().numpy()
    Y = Xn - Xn.mean(axis=1, keepdims=True)
    assert np.abs(Xr - Y).max() <= 1e-4 * np.abs(Y).max()


def test_missing_basis_is_rejected_not_fatal():
    with Server(serve.DeviceExecutor("mixed"), slots=1, budget=1e9) as s:
        r = s.submit(Request(RECONSTRUCT, basis=99))
        assert s.drain(timeout=10)
    assert isinstance(r.error, KeyError)


def test_request_waits_for_its_input_stream():
    """A request submitted while its input is still being produced on another
    stream runs on the finished input (its slot stream waits on the event
    recorded at submission)."""
    side = torch.cuda.Stream()
    g = torch.Generator(device="cuda").manual_seed(12)
    with torch.cuda.stream(side):
        A = torch.randn(400_000, 64, generator=g, device="cuda", dtype=torch.float64)
        B = torch.linalg.qr(torch.randn(64, 64, generator=g, device="cuda", dtype=torch.float64))[0]
        for _ in range(40):
            A = A @ B
        X = A.float()
    with Server(serve.DeviceExecutor("mixed"), slots=2, budget=20e9) as s:
        with torch.cuda.stream(side):
            r = s.submit(Request(POD, X=X))
        assert s.drain(timeout=60)
    assert r.error is None, r.error
    side.synchronize()
    U, S, V = pl.POD.run(X, remove_mean=True, precision="mixed")
    torch.cuda.synchronize()
    assert _same(r.result, (U, S, V))

"""Host logic of the serving loop (pyloworder_amd/serve.py) with a fake
executor: continuous admission, HBM budget, skip-ahead, LRU eviction of
resident bases, rejections.  No GPU."""
import threading
import time

import pytest

from pyloworder_amd import serve
from pyloworder_amd.serve import POD, PROJECT, RECONSTRUCT, Request, Server


class FakeExecutor:
    """Requests carry X = (rows, cols, seconds); bytes = rows * cols."""

    def __init__(self):
        self.lock = threading.Lock()
        self.running = 0
        self.max_running = 0
        self.order = []

    def bytes_needed(self, req, basis):
        if req.kind == POD:
            return req.X[0] * req.X[1]
        return 10

    def run(self, req, basis):
        with self.lock:
            self.running += 1
            self.max_running = max(self.max_running, self.running)
            self.order.append(req.rid)
        time.sleep(req.X[2] if req.X is not None else 0.01)
        with self.lock:
            self.running -= 1
        if req.kind == POD:
            return ("U", req.X, "V")
        return ("C", req.basis)

    def make_stream(self):
        return None

    def stream_context(self, stream):
        return serve._NullCtx()

    def synchronize(self, stream):
        pass

    def resident_bytes(self, result):
        return result[1][0] * result[1][1]


def test_overlap_limited_by_slots():
    ex = FakeExecutor()
    with Server(ex, slots=3, budget=1e9) as s:
        for _ in range(9):
            s.submit(Request(POD, X=(10, 10, 0.05)))
        assert s.drain(timeout=10)
    assert ex.max_running == 3
    assert len(s.completed) == 9 and all(r.error is None for r in s.completed)


def test_late_arrival_joins_running_batch():
    """A request that arrives while others run is admitted as soon as a slot
    frees, before the earlier requests have all finished."""
    ex = FakeExecutor()
    with Server(ex, slots=2, budget=1e9) as s:
        long = s.submit(Request(POD, X=(10, 10, 0.6)))
        short = s.submit(Request(POD, X=(10, 10, 0.05)))
        time.sleep(0.1)  # short done, long still running
        late = s.submit(Request(POD, X=(10, 10, 0.05)))
        assert s.drain(timeout=10)
    assert late.admitted < long.done
    assert short.done < long.done


def test_budget_serialises_large_requests():
    ex = FakeExecutor()
    with Server(ex, slots=4, budget=150, keep_bases=False) as s:
        for _ in range(4):
            s.submit(Request(POD, X=(10, 10, 0.03)))  # 100 bytes each: one at a time
        assert s.drain(timeout=10)
    assert ex.max_running == 1


def test_skip_ahead_admits_small_behind_large():
    ex = FakeExecutor()
    with Server(ex, slots=4, budget=200, keep_bases=False, skip_ahead=4) as s:
        a = s.submit(Request(POD, X=(10, 15, 0.3)))    # 150 bytes, runs
        big = s.submit(Request(POD, X=(10, 10, 0.01)))  # 100: must wait for a
        small = s.submit(Request(POD, X=(5, 5, 0.01)))  # 25: fits beside a
        assert s.drain(timeout=10)
    assert small.admitted < a.done <= big.admitted


def test_lru_eviction_keeps_pinned_and_recent():
    ex = FakeExecutor()
    with Server(ex, slots=1, budget=250) as s:
        r0 = s.submit(Request(POD, X=(10, 10, 0.01)))
        s.drain(timeout=10)
        r1 = s.submit(Request(POD, X=(10, 10, 0.01)))
        s.drain(timeout=10)
        # use basis 0 so basis 1 becomes the least recently used
        s.submit(Request(PROJECT, X=(1, 1, 0.01), basis=r0.basis))
        s.drain(timeout=10)
        r2 = s.submit(Request(POD, X=(10, 10, 0.01)))  # 2 x 100 resident + 100 in flight > 250
        s.drain(timeout=10)
        ids = s.resident_ids()
    assert r1.basis not in ids and r0.basis in ids and r2.basis in ids
    assert s.evictions == 1


def test_rejections():
    ex = FakeExecutor()
    with Server(ex, slots=2, budget=100) as s:
        too_big = s.submit(Request(POD, X=(20, 20, 0.01)))
        missing = s.submit(Request(RECONSTRUCT, basis=12345))
        ok = s.submit(Request(POD, X=(5, 5, 0.01)))
        assert s.drain(timeout=10)
    assert isinstance(too_big.error, MemoryError)
    assert isinstance(missing.error, KeyError)
    assert ok.error is None and len(s.rejected) == 2


def test_errors_reported_on_request_and_loop_continues():
    class Failing(FakeExecutor):
        def run(self, req, basis):
            if req.X[2] < 0:
                raise RuntimeError("boom")
            return super().run(req, basis)
    ex = Failing()
    with Server(ex, slots=2, budget=1e9) as s:
        bad = s.submit(Request(POD, X=(1, 1, -1)))
        good = s.submit(Request(POD, X=(1, 1, 0.01)))
        assert s.drain(timeout=10)
    assert isinstance(bad.error, RuntimeError) and good.error is None


def test_summary_percentiles():
    ex = FakeExecutor()
    with Server(ex, slots=2, budget=1e9) as s:
        for _ in range(6):
            s.submit(Request(POD, X=(1, 1, 0.01)))
        s.drain(timeout=10)
    out = serve.summary(s, 1.0)
    assert out["requests"] == 6 and out["pod"]["count"] == 6
    assert 0 < out["p50_ms"] <= out["p99_ms"]
    assert serve.percentile([1, 2, 3, 4], 50) in (2, 3) and serve.percentile([], 50) is None


def test_reserve_must_fit_budget():
    with pytest.raises(ValueError):
        Server(FakeExecutor(), slots=2, budget=100, reserve_per_slot=60)

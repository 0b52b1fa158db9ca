"""Serving loop for POD requests on one MI355X: continuous batching under an
HBM admission budget, with resident bases.

No reference counterpart: pyLOM runs one POD per script.  The requests reuse
the library's entry points (POD.run = tsqr_svd with fused centring,
POD.project, POD.reconstruct), so every result is the one those calls give.

  * A request is admitted as soon as a slot is idle and its device bytes fit
    the budget: requests that arrive while others run join at once (continuous
    batching), not after the current batch drains.  Admission is FIFO with a
    bounded skip-ahead so that a large request at the head does not idle the
    device, and it evicts least-recently-used idle bases to make room.
  * Each slot is a worker thread with its own HIP stream and library context
    (own workspace), so admitted requests overlap on the device.  The library
    calls release the GIL (ctypes) and every kernel is deterministic, so a
    result does not depend on what else is running.
  * POD results stay resident as bases (U, S, V) for later project /
    reconstruct requests, counted against the budget until evicted.

The scheduling logic is independent of the device: ``Server`` takes an
executor with ``bytes_needed`` / ``run``; ``DeviceExecutor`` is the MI355X
one.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import Any, Callable, Dict, List, Optional

POD, PROJECT, RECONSTRUCT = "pod", "project", "reconstruct"


@dataclasses.dataclass
class Request:
    """One unit of work.  ``X``: input rows x columns (pod, project);
    ``basis``: id of a resident basis (project, reconstruct); ``modes``: number
    of leading modes used by project / reconstruct (0 = all)."""
    kind: str
    X: Any = None
    basis: Optional[int] = None
    modes: int = 0
    arrival: float = 0.0
    rid: int = -1
    # filled by the server
    ready: Any = None  # marker of the input's producer at submission
    admitted: float = 0.0
    done: float = 0.0
    result: Any = None
    error: Optional[BaseException] = None
    nbytes: int = 0
    slot: int = -1

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclasses.dataclass
class Basis:
    U: Any
    S: Any
    V: Any
    nbytes: int
    last_used: float = 0.0
    pins: int = 0


class DeviceExecutor:
    """Runs requests with the library on the calling thread's current stream."""

    def __init__(self, precision: str = "mixed", remove_mean: bool = True):
        from . import _lib
        self.precision = precision
        self.remove_mean = remove_mean
        self._lib = _lib
        self._prec = _lib.PC_PREC_MIXED if precision == "mixed" else _lib.PC_PREC_FP64

    @staticmethod
    def _esz(X) -> int:
        return X.element_size()

    # the mixed path takes n <= 256 (float32); wider POD requests run FP64
    MIXED_MAX_N = 256

    def precision_for(self, X) -> str:
        """Arithmetic of a POD solve of X: the executor's, except that mixed
        requests wider than the mixed kernels take FP64 (same results
        contract, FP64 MFMA)."""
        if self.precision == "mixed" and X.shape[1] > self.MIXED_MAX_N:
            return "fp64"
        return self.precision

    def workspace_bound(self, m_max: int, n_max: int, dtype=None) -> int:
        """Workspace one slot's context may hold: that of the largest solve of
        either precision a request may take."""
        from ._lib import PC_F32, PC_PREC_FP64
        lib = self._lib.load()
        dt = PC_F32 if dtype is None else dtype
        need = [lib.pc_workspace_query(0, dt, m_max, min(n_max, self.MIXED_MAX_N), 0, self._prec)]
        if n_max > self.MIXED_MAX_N or self._prec == PC_PREC_FP64:
            need.append(lib.pc_workspace_query(0, dt, m_max, n_max, 0, PC_PREC_FP64))
        return int(max(max(need), 0))

    def bytes_needed(self, req: Request, basis: Optional[Basis]) -> int:
        """Device bytes of the request's outputs (the slot's workspace is
        reserved separately); a POD solve's outputs stay resident afterwards."""
        if req.kind == POD:
            m, n = req.X.shape
            e = self._esz(req.X)
            return int(m * n * e + n * n * e + n * 8)
        U = basis.U
        m, N = U.shape
        k = req.modes or N
        e = U.element_size()
        if req.kind == PROJECT:
            n = req.X.shape[1]
            return int(k * n * e + 2 * k * n * 8 + (1 << 20))
        n = basis.V.shape[1]
        return int(m * n * e)

    def _dtype(self, X):
        from ._tensor import pc_dtype
        return pc_dtype(X)

    def run(self, req: Request, basis: Optional[Basis]):
        from . import POD as P
        if req.kind == POD:
            return P.run(req.X, remove_mean=self.remove_mean, precision=self.precision_for(req.X))
        k = req.modes or basis.U.shape[1]
        # projections and reconstructions run FP64 MFMA on the stored data
        if req.kind == PROJECT:
            return P.project(basis.U[:, :k], req.X, precision="fp64")
        return P.reconstruct(basis.U[:, :k], basis.S[:k], basis.V[:k, :], precision="fp64")

    def make_stream(self):
        import torch
        return torch.cuda.Stream()

    def input_ready(self, req: Request):
        """Event marking the work queued so far on the submitter's current
        stream (the producer of req.X); None when there is no input."""
        import torch
        if req.X is None:
            return None
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(req.X.device))
        return ev

    def wait_input(self, stream, marker):
        if marker is not None and stream is not None:
            stream.wait_event(marker)

    def stream_context(self, stream):
        import torch
        return torch.cuda.stream(stream)

    def synchronize(self, stream):
        stream.synchronize()

    def resident_bytes(self, result) -> int:
        U, S, V = result
        return int(U.numel() * U.element_size() + S.numel() * S.element_size() + V.numel() * V.element_size())


class _NullCtx:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class Server:
    """Continuous-batching request loop.

    ``slots``: requests in flight at once (one worker thread and stream each);
    ``budget``: device bytes for in-flight requests, resident bases and
    ``reserve_per_slot`` bytes of workspace per slot;
    ``skip_ahead``: how many queued requests behind a head that does not fit
    may be admitted before it; ``keep_bases``: keep POD results resident;
    ``keep_results``: keep each result on its request after ``on_done`` (a
    callback run on completion) has seen it — off for long runs, so that the
    device memory of results and evicted bases is released.
    """

    def __init__(self, executor=None, slots: int = 4, budget: float = 200e9, skip_ahead: int = 8,
                 keep_bases: bool = True, reserve_per_slot: int = 0, keep_results: bool = True,
                 on_done: Optional[Callable[[Request], None]] = None,
                 clock: Callable[[], float] = time.perf_counter):
        self.ex = executor if executor is not None else DeviceExecutor()
        self.slots = slots
        # each slot's context keeps the workspace of the largest request it ran
        self.budget = int(budget) - slots * int(reserve_per_slot)
        if self.budget <= 0:
            raise ValueError("budget does not cover the per-slot workspace reserve")
        self.skip_ahead = skip_ahead
        self.keep_bases = keep_bases
        self.keep_results = keep_results
        self.on_done = on_done
        self.clock = clock
        self.pending: collections.deque = collections.deque()
        self.bases: Dict[int, Basis] = {}
        self.inflight_bytes = 0
        self.used_bases = 0
        self.completed: List[Request] = []
        self.rejected: List[Request] = []
        self._next_basis = 0
        self._next_rid = 0
        self._cv = threading.Condition()
        self._idle = list(range(slots))
        self._work = [None] * slots
        self._stop = False
        self._n_inflight = 0
        self.evictions = 0
        self._threads = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(slots)]
        for t in self._threads:
            t.start()

    # -- public ---------------------------------------------------------------
    def submit(self, req: Request) -> Request:
        # the request's slot stream waits for what the submitter's stream has
        # queued so far (req.X may still be in flight there)
        mark = getattr(self.ex, "input_ready", None)
        req.ready = mark(req) if mark is not None else None
        with self._cv:
            if req.arrival == 0.0:
                req.arrival = self.clock()
            req.rid = self._next_rid
            self._next_rid += 1
            if req.kind != POD:
                b = self.bases.get(req.basis)
                if b is None:
                    req.error = KeyError(f"basis {req.basis} is not resident")
                    req.done = self.clock()
                    self.rejected.append(req)
                    return req
                b.pins += 1  # a basis with queued or running users is never evicted
            self.pending.append(req)
            self._admit_locked()
            self._cv.notify_all()
        return req

    def add_basis(self, U, S, V) -> int:
        """Register an existing (U, S, V) as a resident basis; returns its id."""
        with self._cv:
            return self._add_basis_locked((U, S, V))

    def resident_ids(self) -> List[int]:
        with self._cv:
            return list(self.bases)

    def drain(self, timeout: Optional[float] = None) -> bool:
        """Wait until nothing is queued or in flight."""
        end = None if timeout is None else self.clock() + timeout
        with self._cv:
            while self.pending or self._n_inflight:
                left = None if end is None else end - self.clock()
                if left is not None and left <= 0:
                    return False
                self._cv.wait(timeout=left if left is not None else 0.5)
                self._admit_locked()
        return True

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join(timeout=30)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
        return False

    # -- scheduling -----------------------------------------------------------
    def _add_basis_locked(self, res) -> int:
        bid = self._next_basis
        self._next_basis += 1
        nb = self.ex.resident_bytes(res)
        self.bases[bid] = Basis(res[0], res[1], res[2], nb, self.clock())
        self.used_bases += nb
        return bid

    def _evict_for_locked(self, need: int) -> bool:
        """Evict idle bases, least recently used first, until need fits."""
        while self.inflight_bytes + self.used_bases + need > self.budget:
            idle = [(b.last_used, bid) for bid, b in self.bases.items() if b.pins == 0]
            if not idle:
                return False
            _, bid = min(idle)
            b = self.bases.pop(bid)
            self.used_bases -= b.nbytes
            self.evictions += 1
        return True

    def _admit_locked(self):
        i = 0
        while self._idle and i < len(self.pending) and i <= self.skip_ahead:
            req = self.pending[i]
            basis = self.bases[req.basis] if req.kind != POD else None
            need = int(self.ex.bytes_needed(req, basis))
            if need > self.budget:
                del self.pending[i]
                if basis is not None:
                    basis.pins -= 1
                req.error = MemoryError(f"request needs {need} bytes, budget is {self.budget}")
                req.done = self.clock()
                self.rejected.append(req)
                continue
            if not self._evict_for_locked(need):
                i += 1
                continue
            del self.pending[i]
            req.nbytes = need
            req.admitted = self.clock()
            self.inflight_bytes += need
            slot = self._idle.pop()
            req.slot = slot
            self._work[slot] = (req, basis)
            self._n_inflight += 1
            self._cv.notify_all()

    def _worker(self, slot: int):
        stream = self.ex.make_stream()
        while True:
            with self._cv:
                while self._work[slot] is None and not self._stop:
                    self._cv.wait()
                if self._stop and self._work[slot] is None:
                    return
                req, basis = self._work[slot]
            try:
                if req.ready is not None:
                    self.ex.wait_input(stream, req.ready)
                with self.ex.stream_context(stream) if stream is not None else _NullCtx():
                    res = self.ex.run(req, basis)
                self.ex.synchronize(stream)
                req.result = res
            except BaseException as e:  # reported on the request, the loop keeps serving
                req.error = e
            with self._cv:
                req.done = self.clock()
                self.inflight_bytes -= req.nbytes
                if basis is not None:
                    basis.pins -= 1
                    basis.last_used = req.done
                if req.kind == POD and req.error is None and self.keep_bases:
                    req.basis = self._add_basis_locked(req.result)
                if self.on_done is not None:
                    self.on_done(req)
                if not self.keep_results:
                    req.result = None
                self._work[slot] = None
                self._idle.append(slot)
                self._n_inflight -= 1
                self.completed.append(req)
                self._admit_locked()
                self._cv.notify_all()


def percentile(xs, q):
    if not xs:
        return None
    s = sorted(xs)
    return s[min(len(s) - 1, int(round(q / 100.0 * (len(s) - 1))))]


def summary(server: Server, wall: float) -> dict:
    """p50 / p99 latency per kind, solves/s and requests/s over ``wall`` seconds."""
    out = {"requests": len(server.completed), "rejected": len(server.rejected), "evictions": server.evictions,
           "wall_s": round(wall, 3)}
    ok = [r for r in server.completed if r.error is None]
    out["errors"] = len(server.completed) - len(ok)
    lat = [r.latency * 1e3 for r in ok]
    out["p50_ms"] = percentile(lat, 50)
    out["p99_ms"] = percentile(lat, 99)
    for kind in (POD, PROJECT, RECONSTRUCT):
        lk = [r.latency * 1e3 for r in ok if r.kind == kind]
        out[kind] = {"count": len(lk), "p50_ms": percentile(lk, 50), "p99_ms": percentile(lk, 99)}
    out["solves_per_s"] = round(sum(1 for r in ok if r.kind == POD) / wall, 2) if wall > 0 else None
    out["requests_per_s"] = round(len(ok) / wall, 2) if wall > 0 else None
    return out


# ---------------------------------------------------------------------------
# native request loop (pc_serve_*, include/podcore.h)
# ---------------------------------------------------------------------------
_HIP = None


def _dev_copy(dst: int, src: int, nbytes: int) -> None:
    """Device-to-device copy through the HIP runtime (synchronous)."""
    import ctypes as C
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.restype = C.c_int
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rc = _HIP.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")


class NativeServer:
    """Python handle of the native request loop: the admission, eviction,
    slots and streams all run in the library (C++ threads), so the serving
    loop holds no Python lock while requests run.

    ``submit_*`` return a ticket; ``wait`` returns the request's record (status,
    basis id, timings) once done, or None on timeout; ``output`` copies a
    projection / reconstruction output into a new tensor and releases it;
    ``basis`` copies a resident basis (U, S, VT).  Inputs stay referenced here
    until their request is done.
    """

    def __init__(self, slots: int = 4, budget: float = 200e9, skip_ahead: int = 8, precision: str = "mixed",
                 remove_mean: bool = True, m_max: int = 4_000_000, n_max: int = 512, device: Optional[int] = None):
        import ctypes as C
        import torch
        from . import _lib
        self._C, self._torch, self._lib_mod = C, torch, _lib
        self.lib = _lib.load()
        dev = torch.cuda.current_device() if device is None else device
        prec = _lib.PC_PREC_MIXED if precision == "mixed" else _lib.PC_PREC_FP64
        h = C.c_void_p()
        rc = self.lib.pc_serve_create(dev, slots, int(budget), skip_ahead, prec, 1 if remove_mean else 0,
                                      int(m_max), int(n_max), C.byref(h))
        if rc != 0:
            raise _lib.PodCoreError(rc, "pc_serve_create failed")
        self.h = h
        self.device = dev
        self._inputs: Dict[int, Any] = {}

    # -- requests -------------------------------------------------------------
    def _submit(self, kind, X=None, basis=-1, modes=0) -> int:
        from ._tensor import pc_dtype
        C, _lib = self._C, self._lib_mod
        q = _lib.PCRequest()
        q.kind = kind
        q.basis = int(basis)
        q.modes = int(modes)
        if X is not None:
            if not (X.is_cuda and X.dim() == 2 and X.stride(1) == 1):
                raise ValueError("X must be a 2-D CUDA tensor with unit column stride")
            q.dtype = pc_dtype(X)
            q.X = X.data_ptr()
            q.ldx = X.stride(0)
            q.m, q.n = X.shape
            # X may still be in flight on torch's stream: the request waits for it
            q.stream = self._torch.cuda.current_stream(X.device).cuda_stream
        t = C.c_int64()
        rc = self.lib.pc_serve_submit(self.h, C.byref(q), C.byref(t))
        if rc != 0:
            raise _lib.PodCoreError(rc, "pc_serve_submit: " + self._err())
        if X is not None:
            self._inputs[t.value] = X
        return t.value

    def submit_pod(self, X) -> int:
        return self._submit(self._lib_mod.PC_REQ_POD, X)

    def submit_project(self, basis: int, X, modes: int = 0) -> int:
        return self._submit(self._lib_mod.PC_REQ_PROJECT, X, basis, modes)

    def submit_reconstruct(self, basis: int, modes: int = 0) -> int:
        return self._submit(self._lib_mod.PC_REQ_RECONSTRUCT, None, basis, modes)

    def wait(self, ticket: int, timeout: Optional[float] = None) -> Optional[dict]:
        C, _lib = self._C, self._lib_mod
        r = _lib.PCResult()
        rc = self.lib.pc_serve_wait(self.h, ticket, -1.0 if timeout is None else float(timeout), C.byref(r))
        if rc == _lib.PC_PENDING:
            return None
        if rc != 0:
            raise _lib.PodCoreError(rc, "pc_serve_wait: " + self._err())
        self._inputs.pop(ticket, None)
        out = {"status": r.status, "basis": r.basis, "out": r.out, "shape": (r.out_rows, r.out_cols),
               "ld": r.out_ld, "t_submit": r.t_submit, "t_admit": r.t_admit, "t_done": r.t_done,
               "latency": r.t_done - r.t_submit}
        if r.status != 0:
            out["error"] = self._err()
        return out

    def _err(self) -> str:
        return self.lib.pc_serve_last_error(self.h).decode()

    def output(self, ticket: int, dtype, release: bool = True):
        """Copy of a finished projection / reconstruction output."""
        rec = self.wait(ticket)
        if rec["status"] != 0:
            raise self._lib_mod.PodCoreError(rec["status"], rec.get("error", "request failed"))
        rows, cols = rec["shape"]
        t = self._torch.empty((rows, cols), dtype=dtype, device=f"cuda:{self.device}")
        _dev_copy(t.data_ptr(), rec["out"], t.numel() * t.element_size())
        if release:
            self.release(ticket)
        return t

    def release(self, ticket: int) -> None:
        rc = self.lib.pc_serve_release(self.h, ticket)
        if rc != 0:
            raise self._lib_mod.PodCoreError(rc, "pc_serve_release: " + self._err())

    def basis(self, bid: int):
        """Copies (U, S, VT) of a resident basis (pinned while copied)."""
        C, _lib, torch = self._C, self._lib_mod, self._torch
        v = _lib.PCBasisView()
        rc = self.lib.pc_serve_basis(self.h, bid, 1, C.byref(v))
        if rc != 0:
            raise _lib.PodCoreError(rc, "pc_serve_basis: " + self._err())
        try:
            dt = torch.float32 if v.dtype == _lib.PC_F32 else torch.float64
            dev = f"cuda:{self.device}"
            U = torch.empty((v.m, v.n), dtype=dt, device=dev)
            S = torch.empty((v.n,), dtype=dt, device=dev)
            VT = torch.empty((v.n, v.n), dtype=dt, device=dev)
            for t, p in ((U, v.U), (S, v.S), (VT, v.VT)):
                _dev_copy(t.data_ptr(), p, t.numel() * t.element_size())
        finally:
            self.lib.pc_serve_basis(self.h, bid, -1, None)
        return U, S, VT

    def resident(self) -> List[int]:
        """Ids of the resident bases."""
        C = self._C
        cnt = C.c_int64(0)
        self.lib.pc_serve_bases(self.h, None, 0, C.byref(cnt))
        ids = (C.c_int64 * max(1, cnt.value))()
        self.lib.pc_serve_bases(self.h, ids, cnt.value, C.byref(cnt))
        return list(ids[: cnt.value])

    def basis_shape(self, bid: int):
        """(m, n) of a resident basis, or None once it is evicted."""
        v = self._lib_mod.PCBasisView()
        if self.lib.pc_serve_basis(self.h, bid, 0, self._C.byref(v)) != 0:
            return None
        return v.m, v.n

    def drain(self, timeout: Optional[float] = None) -> bool:
        return self.lib.pc_serve_drain(self.h, -1.0 if timeout is None else float(timeout)) == 0

    def stats(self) -> dict:
        C = self._C
        v = (C.c_int64 * 9)()
        self.lib.pc_serve_stats(self.h, v, 9)
        keys = ("completed", "rejected", "evictions", "inflight_bytes", "resident_bytes", "queued", "running",
                "bases", "output_budget")
        return dict(zip(keys, list(v)))

    def close(self) -> None:
        if self.h:
            self.lib.pc_serve_destroy(self.h)
            self.h = None
            self._inputs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
        return False

"""Timing channels with the reference's names (pyLOM/utils/cr.py:234-281).

``@cr("math.tsqr_svd")`` records host wall time per call under that channel.
With ``PYLOM_CR_SYNC=1`` the device stream is synchronised at both ends so the
time covers the GPU work (the reference's channels never synchronise, so
asynchronous work lands in whichever channel synchronises later).
"""
from __future__ import annotations

import functools
import os
import time
from collections import defaultdict

_ch = defaultdict(lambda: [0, 0.0, float("inf"), 0.0])  # n, sum, min, max
_open = {}
_SYNC = os.environ.get("PYLOM_CR_SYNC", "0") == "1"


def _sync():
    if _SYNC:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()


def cr_start(name, _rank=0):
    _sync()
    _open[name] = time.perf_counter()


def cr_stop(name, _rank=0):
    _sync()
    t0 = _open.pop(name, None)
    if t0 is None:
        return
    dt = time.perf_counter() - t0
    c = _ch[name]
    c[0] += 1
    c[1] += dt
    c[2] = min(c[2], dt)
    c[3] = max(c[3], dt)


def cr(name):
    def deco(fn):
        @functools.wraps(fn)
        def wrap(*a, **k):
            cr_start(name)
            try:
                return fn(*a, **k)
            finally:
                cr_stop(name)
        return wrap
    return deco


def cr_info():
    """Channel table: name, n, sum, min, max, avg (seconds)."""
    rows = []
    for k, (n, s, mn, mx) in sorted(_ch.items()):
        rows.append((k, n, s, mn, mx, s / max(n, 1)))
        print(f"{k:32s} n={n:6d} sum={s:.6f} min={mn:.6f} max={mx:.6f} avg={s / max(n, 1):.6f}")
    return rows


def cr_reset():
    _ch.clear()
    _open.clear()

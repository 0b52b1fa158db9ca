"""C-ABI checks that need no GPU: the library loads, every symbol declared in
include/podcore.h is exported, and the host-only entry points behave."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from pyloworder_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    txt = (ROOT / "include" / "podcore.h").read_text()
    return sorted(set(re.findall(r"^(?:int|int64_t|const char\*)\s+(pc_[a-z_]+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # and the binding table covers exactly the header
    assert set(names) == set(_lib.SIGNATURES)


def test_version_string():
    lib = _lib.load()
    assert b"gfx950" in lib.pc_version()


def _trunc(S, r):
    lib = _lib.load()
    s = np.ascontiguousarray(np.asarray(S, dtype=np.float64))
    N = C.c_int64(-1)
    rc = lib.pc_truncation_residual(s.ctypes.data_as(C.POINTER(C.c_double)), s.shape[0], r, C.byref(N))
    assert rc == 0
    return N.value


@pytest.mark.parametrize("S,r,N", [
    ([4, 3, 2, 1], -0.5, 2),     # cumulative 0.4, 0.7 -> i+1 = 2 (truncation.py:34-38)
    ([4, 3, 2, 1], 0.3, 3),      # ||S[i:]||/||S|| = 1, .73, .41, .18 -> 3
    ([4, 3, 2, 1], -1.0, 4),     # never exceeds 1 -> all
    ([4, 3, 2, 1], 1e-8, 4),
    ([1.0], 0.5, 1),
    ([10, 1e-3, 1e-9], 1e-6, 2),
])
def test_truncation_residual_host(S, r, N):
    from oracle import compute_truncation_residual
    assert compute_truncation_residual(np.array(S, dtype=float), r) == N
    assert _trunc(S, r) == N


def test_truncation_matches_oracle_random(rng):
    from oracle import compute_truncation_residual
    for _ in range(200):
        n = int(rng.integers(1, 60))
        S = np.sort(rng.random(n) ** 3)[::-1] * 10
        r = float(rng.choice([-1, 1]) * rng.random())
        assert _trunc(S, r) == compute_truncation_residual(S, r)


def test_bad_arguments_fail_without_gpu():
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.pc_ctx_create(0, 0, 0, C.byref(h))
    if rc == 0:  # a GPU is visible: then argument checks must still reject NULLs
        assert lib.pc_tsqr_svd(h, 1, None, 4, 4, 4, 0, None, 4, None, None, 4, 0) == -1
        lib.pc_ctx_destroy(h)
    else:
        assert rc in (-4, -1)
        assert len(lib.pc_last_error(None)) > 0
    # NULL context is rejected, never dereferenced
    assert lib.pc_tsqr_svd(None, 1, None, 4, 4, 4, 0, None, 4, None, None, 4, 0) == -1
    assert lib.pc_truncation_residual(None, 3, 0.5, None) == -1

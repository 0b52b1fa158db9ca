"""pc_tsqr_svd_noalloc (the allocation-free tall-skinny SVD on the library's
static device scratch) against pc_tsqr_svd, called through the C ABI.

Both run the same kernels with the same reduction orders, so the outputs
must be identical, bit for bit: checked on two calls in a row, the second
with a smaller input (the scratch keeps the first call's contents), for the
FP64 and the mixed path.  The call must not grow any buffer: a context that
never allocated a workspace still holds none afterwards.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

pl = pytest.importorskip("pyloworder_amd")
from pyloworder_amd import _lib  # noqa: E402
from oracle import workloads  # noqa: E402


def _outs(m, n, dt):
    rdt = torch.float32 if dt in (torch.float32, torch.complex64) else torch.float64
    return (torch.empty((m, n), dtype=dt, device="cuda"), torch.empty((n,), dtype=rdt, device="cuda"),
            torch.empty((n, n), dtype=dt, device="cuda"))


def _call(lib, h, fn, X, center, prec, outs):
    U, S, V = outs
    m, n = X.shape
    rc = getattr(lib, fn)(h, _lib.PC_F32 if X.dtype == torch.float32 else _lib.PC_F64, X.data_ptr(), X.stride(0), m,
                          n, center, U.data_ptr(), U.stride(0), S.data_ptr(), V.data_ptr(), V.stride(0), prec)
    assert rc == 0, lib.pc_last_error(h)
    torch.cuda.synchronize()
    return [t.clone() for t in outs]


@pytest.mark.parametrize("dt,prec", [(np.float64, _lib.PC_PREC_FP64), (np.float32, _lib.PC_PREC_MIXED)])
def test_noalloc_equals_tsqr_svd_twice(dt, prec):
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.pc_ctx_create(0, 0, 0, C.byref(h)) == 0
    try:
        assert lib.pc_workspace_bytes(h) == 0
        # the second call is smaller than the first (fewer rows and columns)
        for (m, n, center) in ((25_000, 200, 1), (9_001, 48, 0)):
            X = torch.from_numpy(workloads.lowrank(m, n, seed=n, decay=0.9, top=100.0, noise=1e-5,
                                                   dtype=dt)).cuda()
            na = _call(lib, h, "pc_tsqr_svd_noalloc", X, center, prec, _outs(m, n, X.dtype))
            assert lib.pc_workspace_bytes(h) == 0  # nothing grown: the static scratch was used
            ref = pl.math.tsqr_svd(X, precision="mixed" if prec == _lib.PC_PREC_MIXED else "fp64", center=center)
            for a, b in zip(na, ref):
                assert torch.equal(a, b)
    finally:
        lib.pc_ctx_destroy(h)


def test_noalloc_rejects_beyond_the_scratch():
    lib = _lib.load()
    ctx = _lib.context()
    m, n = 100_000, 256  # needs more than the 128 MiB scratch
    assert lib.pc_workspace_query(0, _lib.PC_F32, m, n, 0, _lib.PC_PREC_MIXED) > 128 << 20
    X = torch.zeros((m, n), dtype=torch.float32, device="cuda")
    U, S, V = _outs(m, n, torch.float32)
    rc = lib.pc_tsqr_svd_noalloc(ctx.h, _lib.PC_F32, X.data_ptr(), n, m, n, 0, U.data_ptr(), n, S.data_ptr(),
                                 V.data_ptr(), n, _lib.PC_PREC_MIXED)
    assert rc == -3 and b"scratch" in lib.pc_last_error(ctx.h)

"""Launch tuner of the product pass (pc_autotune): it changes the chunk count
of U = Y W only, so the tuned solve is bitwise equal (tolerance 0) to the
untuned one; the chosen count is recorded and reused."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

pl = pytest.importorskip("pyloworder_amd")
from pyloworder_amd import _lib  # noqa: E402
from pyloworder_amd._tensor import pc_dtype, real_dtype  # noqa: E402


def _solve(ctx, X, prec):
    m, n = X.shape
    U = torch.empty((m, n), dtype=X.dtype, device="cuda")
    S = torch.empty((n,), dtype=real_dtype(X.dtype), device="cuda")
    V = torch.empty((n, n), dtype=X.dtype, device="cuda")
    ctx.check(ctx.lib.pc_tsqr_svd(ctx.h, pc_dtype(X), X.data_ptr(), X.stride(0), m, n, 1, U.data_ptr(), n,
                                  S.data_ptr(), V.data_ptr(), n, prec))
    torch.cuda.synchronize()
    return U, S, V


@pytest.mark.parametrize("m,n,dtype,prec", [(600_000, 256, torch.float32, _lib.PC_PREC_MIXED),
                                            (300_000, 96, torch.float32, _lib.PC_PREC_MIXED),
                                            (200_000, 64, torch.float64, _lib.PC_PREC_FP64)])
def test_tuned_product_pass_is_bitwise_equal(m, n, dtype, prec):
    g = torch.Generator(device="cuda").manual_seed(n)
    X = torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64).to(dtype)
    lib = _lib.load()
    assert lib.pc_autotune_clear() == 0
    plain = _lib.Context()
    U0, S0, V0 = _solve(plain, X, prec)
    ch = C.c_int64(-1)
    assert lib.pc_autotune_query(pc_dtype(X), m, n, prec, C.byref(ch)) == 0   # nothing tuned yet
    tuned = _lib.Context()
    assert lib.pc_autotune(tuned.h, 1) == 0
    U1, S1, V1 = _solve(tuned, X, prec)                                      # tunes, then solves
    assert lib.pc_autotune_query(pc_dtype(X), m, n, prec, C.byref(ch)) == 1
    assert ch.value >= 1
    U2, S2, V2 = _solve(plain, X, prec)                                      # table used on another context
    for a, b in ((U0, U1), (S0, S1), (V0, V1), (U0, U2), (S0, S2)):
        assert torch.equal(a, b)
    assert lib.pc_autotune(tuned.h, 2) != 0
    lib.pc_autotune_clear()

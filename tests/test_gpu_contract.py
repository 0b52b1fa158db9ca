"""Boundary contract of include/podcore.h, called through ctypes.

Every entry point that takes ``precision`` either computes in the arithmetic
asked for or returns PC_ERR_UNSUPPORTED; omega == NULL draws the reference's
sketch matrix on the device from ``seed`` (numpy's legacy seeding range
0 <= seed < 2**32, as np.random.seed in pyLOM/vmmath/svd.py:131-133).
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

pl = pytest.importorskip("pyloworder_amd")
from pyloworder_amd import _lib  # noqa: E402

OK, INVALID, UNSUPPORTED = 0, -1, -6


@pytest.fixture
def ctx():
    return _lib.context()


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_fp64_only_products_refuse_mixed(ctx):
    rng = np.random.default_rng(5)
    U = _t(rng.standard_normal((4096, 8)).astype(np.float32))
    X = _t(rng.standard_normal((4096, 16)).astype(np.float32))
    S = _t(np.ones(8, np.float32))
    V = _t(rng.standard_normal((8, 16)).astype(np.float32))
    Cm = torch.empty((8, 16), dtype=torch.float32, device="cuda")
    Xr = torch.empty((4096, 16), dtype=torch.float32, device="cuda")
    L = ctx.lib
    for prec, want in ((_lib.PC_PREC_FP64, OK), (_lib.PC_PREC_MIXED, UNSUPPORTED), (7, INVALID)):
        assert L.pc_project(ctx.h, _lib.PC_F32, U.data_ptr(), 8, X.data_ptr(), 16, 4096, 8, 16, Cm.data_ptr(), 16,
                            prec) == want
        assert L.pc_reconstruct(ctx.h, _lib.PC_F32, U.data_ptr(), 8, S.data_ptr(), V.data_ptr(), 16, 4096, 8, 16,
                                Xr.data_ptr(), 16, prec) == want
        assert L.pc_gemm(ctx.h, _lib.PC_F32, 1, 0, 8, 16, 4096, 1.0, U.data_ptr(), 8, X.data_ptr(), 16, 0.0,
                         Cm.data_ptr(), 16, prec) == want
        if want != OK:
            assert len(L.pc_last_error(ctx.h)) > 0
    torch.cuda.synchronize()


def test_reconstruct_zero_modes_gives_zero(ctx):
    X = torch.full((3000, 7), 5.0, dtype=torch.float64, device="cuda")
    U = torch.ones((3000, 1), dtype=torch.float64, device="cuda")
    assert ctx.lib.pc_reconstruct(ctx.h, _lib.PC_F64, U.data_ptr(), 1, None, U.data_ptr(), 7, 3000, 0, 7,
                                  X.data_ptr(), 7, _lib.PC_PREC_FP64) == OK
    torch.cuda.synchronize()
    assert float(X.abs().max()) == 0.0


def test_randomized_seed_range_and_null_omega(ctx):
    rng = np.random.default_rng(6)
    A = _t(rng.standard_normal((2000, 20)))
    Q = torch.empty((2000, 5), dtype=torch.float64, device="cuda")
    B = torch.empty((5, 20), dtype=torch.float64, device="cuda")
    L = ctx.lib
    args = (ctx.h, _lib.PC_F64, A.data_ptr(), 20, 2000, 20, 5, 1)
    assert L.pc_randomized_qr(*args, 2**32, None, Q.data_ptr(), 5, B.data_ptr(), 20, _lib.PC_PREC_FP64) == INVALID
    assert L.pc_randomized_qr(*args, 2**32 - 1, None, Q.data_ptr(), 5, B.data_ptr(), 20, _lib.PC_PREC_FP64) == OK
    torch.cuda.synchronize()
    om = _t(np.random.RandomState(2**32 - 1).random_sample((20, 5)))
    Q2 = torch.empty_like(Q)
    B2 = torch.empty_like(B)
    assert L.pc_randomized_qr(*args, 0, om.data_ptr(), Q2.data_ptr(), 5, B2.data_ptr(), 20, _lib.PC_PREC_FP64) == OK
    torch.cuda.synchronize()
    assert torch.equal(Q, Q2) and torch.equal(B, B2)

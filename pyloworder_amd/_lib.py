"""ctypes binding of libpodcore.so (the C ABI declared in include/podcore.h).

The library is built in-tree by ``pyloworder_amd.csrc.build``.  There is no
CPU fallback: if the library is missing or no GPU is visible, every call
raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpodcore.so"

PC_F32, PC_F64, PC_C64, PC_C128 = 0, 1, 2, 3
PC_PREC_FP64, PC_PREC_MIXED = 0, 1
STATUS = {
    0: "PC_OK", -1: "PC_ERR_INVALID", -2: "PC_ERR_SHAPE", -3: "PC_ERR_NOMEM",
    -4: "PC_ERR_HIP", -5: "PC_ERR_LINALG", -6: "PC_ERR_UNSUPPORTED",
}
STATUS_CODE = {v: k for k, v in STATUS.items()}

i64 = C.c_int64
vp = C.c_void_p
ip = C.c_int

# name -> (restype, argtypes); every symbol declared in include/podcore.h
SIGNATURES = {
    "pc_ctx_create": (ip, [ip, i64, C.c_uint32, C.POINTER(vp)]),
    "pc_ctx_destroy": (ip, [vp]),
    "pc_set_stream": (ip, [vp, vp]),
    "pc_last_error": (C.c_char_p, [vp]),
    "pc_workspace_bytes": (i64, [vp]),
    "pc_workspace_query": (i64, [ip, ip, i64, i64, i64, ip]),
    "pc_synchronize": (ip, [vp]),
    "pc_profile": (ip, [vp, ip]),
    "pc_profile_read": (ip, [vp, C.POINTER(C.c_double), C.POINTER(i64), ip]),
    "pc_version": (C.c_char_p, []),
    "pc_row_stats": (ip, [vp, ip, vp, i64, i64, i64, vp, vp]),
    "pc_row_center": (ip, [vp, ip, vp, i64, i64, i64, vp, vp, vp, i64]),
    "pc_tsqr_svd": (ip, [vp, ip, vp, i64, i64, i64, ip, vp, i64, vp, vp, i64, ip]),
    "pc_tsqr_svd_noalloc": (ip, [vp, ip, vp, i64, i64, i64, ip, vp, i64, vp, vp, i64, ip]),
    "pc_tsqr": (ip, [vp, ip, vp, i64, i64, i64, vp, i64, vp, i64, ip]),
    "pc_svd": (ip, [vp, ip, vp, i64, i64, i64, vp, i64, vp, vp, i64]),
    "pc_randomized_qr": (ip, [vp, ip, vp, i64, i64, i64, i64, ip, C.c_uint64, vp, vp, i64, vp, i64, ip]),
    "pc_randomized_svd": (ip, [vp, ip, vp, i64, i64, i64, i64, ip, C.c_uint64, vp, vp, i64, vp, vp, i64, ip]),
    "pc_gemm": (ip, [vp, ip, ip, ip, i64, i64, i64, C.c_double, vp, i64, vp, i64, C.c_double, vp, i64, ip]),
    "pc_vecmat": (ip, [vp, ip, vp, vp, i64, i64, i64, vp, i64]),
    "pc_axpy": (ip, [vp, ip, C.c_double, vp, i64, vp, i64, i64, i64]),
    "pc_transpose": (ip, [vp, ip, vp, i64, i64, i64, vp, i64]),
    "pc_reconstruct": (ip, [vp, ip, vp, i64, vp, vp, i64, i64, i64, i64, vp, i64, ip]),
    "pc_project": (ip, [vp, ip, vp, i64, vp, i64, i64, i64, i64, vp, i64, ip]),
    "pc_diff_norms": (ip, [vp, ip, vp, i64, vp, i64, i64, i64, C.POINTER(C.c_double)]),
    "pc_truncation_residual": (ip, [C.POINTER(C.c_double), i64, C.c_double, C.POINTER(i64)]),
    "pc_grouped_tsqr_svd": (ip, [vp, ip, vp, i64, ip]),
    "pc_spod_blocks": (ip, [vp, vp, i64, i64, i64, i64, i64, ip, vp]),
    "pc_spod_modes": (ip, [vp, vp, i64, i64, i64, vp, vp, ip]),
    "pc_grouped_graph_create": (ip, [vp, ip, vp, i64, ip, C.POINTER(vp)]),
    "pc_grouped_graph_launch": (ip, [vp, vp]),
    "pc_graph_destroy": (ip, [vp]),
    "pc_eig": (ip, [vp, vp, i64, i64, vp, vp, vp, i64]),
    "pc_dataset_x": (ip, [vp, ip, i64, C.POINTER(vp), C.POINTER(i64), C.POINTER(i64), i64, i64, vp, i64]),
    "pc_dmd_operator": (ip, [vp, vp, i64, vp, i64, i64, i64, i64, vp, vp, i64, vp, i64]),
    "pc_dmd_modes": (ip, [vp, vp, i64, i64, i64, i64, vp, vp, i64, vp, vp, vp, vp, vp, vp, i64, vp]),
    "pc_dmd_reconstruct": (ip, [vp, vp, i64, i64, i64, vp, vp, vp, vp, i64, vp, i64]),
    "pc_autotune": (ip, [vp, ip]),
    "pc_autotune_query": (ip, [ip, i64, i64, ip, C.POINTER(i64)]),
    "pc_autotune_clear": (ip, []),
    "pc_serve_create": (ip, [ip, ip, i64, ip, ip, ip, i64, i64, C.POINTER(vp)]),
    "pc_serve_destroy": (ip, [vp]),
    "pc_serve_submit": (ip, [vp, vp, C.POINTER(i64)]),
    "pc_serve_wait": (ip, [vp, i64, C.c_double, vp]),
    "pc_serve_release": (ip, [vp, i64]),
    "pc_serve_drain": (ip, [vp, C.c_double]),
    "pc_serve_basis": (ip, [vp, i64, ip, vp]),
    "pc_serve_stats": (ip, [vp, C.POINTER(i64), ip]),
    "pc_serve_bases": (ip, [vp, C.POINTER(i64), i64, C.POINTER(i64)]),
    "pc_serve_last_error": (C.c_char_p, [vp]),
}


class PCProblem(C.Structure):
    """Mirror of pc_problem (include/podcore.h)."""
    _fields_ = [("A", vp), ("lda", i64), ("m", i64), ("n", i64), ("center", C.c_int32), ("pad_", C.c_int32),
                ("U", vp), ("ldu", i64), ("S", vp), ("VT", vp), ("ldvt", i64)]


PC_PENDING = 1
PC_REQ_POD, PC_REQ_PROJECT, PC_REQ_RECONSTRUCT = 0, 1, 2


class PCRequest(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dtype", C.c_int32), ("X", vp), ("ldx", i64), ("m", i64), ("n", i64),
                ("basis", i64), ("modes", i64), ("stream", vp)]


class PCResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("pad_", C.c_int32), ("basis", i64), ("out", vp), ("out_rows", i64),
                ("out_cols", i64), ("out_ld", i64), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double)]


class PCBasisView(C.Structure):
    _fields_ = [("U", vp), ("S", vp), ("VT", vp), ("m", i64), ("n", i64), ("dtype", C.c_int32), ("pad_", C.c_int32)]


class PodCoreError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class LinAlgError(PodCoreError):
    """Raised for PC_ERR_LINALG (the reference raises numpy/cupy LinAlgError)."""


_lib = None
_lock = threading.Lock()


def load(path: str | os.PathLike | None = None):
    """Load the shared library (no GPU needed) and bind every symbol."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise ImportError(f"{p} is not built; run pyloworder_amd.csrc.build.build() (hipcc, gfx950)")
        lib = C.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)  # AttributeError if a declared symbol is missing
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


class Context:
    """One pc_ctx per (thread, device); the stream is set before each call."""

    def __init__(self, device: int = 0, budget: int = 0):
        self.lib = load()
        h = vp()
        rc = self.lib.pc_ctx_create(device, budget, 0, C.byref(h))
        if rc != 0:
            raise PodCoreError(rc, self.lib.pc_last_error(None).decode())
        self.h = h
        self.device = device

    def check(self, rc):
        if rc != 0:
            msg = self.lib.pc_last_error(self.h).decode()
            if rc == -5:
                raise LinAlgError(rc, msg)
            raise PodCoreError(rc, msg)

    def set_stream(self, stream_handle: int):
        self.check(self.lib.pc_set_stream(self.h, vp(stream_handle)))

    def close(self):
        if self.h:
            self.lib.pc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


_tls = threading.local()


def context(device: int | None = None) -> Context:
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pyloworder_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    dev = torch.cuda.current_device() if device is None else device
    cache = getattr(_tls, "ctx", None)
    if cache is None:
        cache = _tls.ctx = {}
    ctx = cache.get(dev)
    if ctx is None:
        ctx = cache[dev] = Context(dev)
    ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    return ctx

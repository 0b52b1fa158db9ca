"""Benchmark of the POD core on one MI355X (prints ONE JSON line).

Workload ("tall-skinny SVD 8M x 256"): X = G1 diag(s) G2^T + N(0, 1e-6),
G1 ~ N(0,1)/sqrt(m) (8,000,000 x 256), G2 orthonormal, s_k = 1e3 0.93^k,
generated on the device from a fixed seed (not timed).  One step = one
POD.run(X, remove_mean=True) = row centring + tall-skinny SVD, with the input
already resident in HBM.

  * mixed: float32 storage (8.2 GB), fp16 hi/lo split MFMA, FP64 small steps —
    the headline `value` (algorithmic HBM GB/s = 3 m n 4 B / time per solve);
  * fp64:  float64 storage (16.4 GB), FP64 MFMA — reported beside it.

Singular-value errors of the mixed result are taken against the FP64 GPU
path run on the same stored (float32, upcast) data; that path is itself
checked against the CPU oracle by tests/test_gpu_core.py.  The CPU baseline
is the oracle's numpy restatement of the reference algorithm (butterfly TSQR +
svd(R) + Q U_R) on a bounded row sample of the same matrix.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

# Grouped and serving workloads spread independent solves over one stream per
# hardware queue (the library reads this variable); HIP opens 4 queues per
# process unless told otherwise.  Set before anything initialises HIP.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

M_DEFAULT, N_DEFAULT = 8_000_000, 256
HBM_PEAK_GBS = 8000.0
MFMA_F16_PEAK_TFS = 2500.0  # dense fp16 MFMA, MI355X


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--m", type=int, default=M_DEFAULT)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run one plain device copy of the float32 input (known bytes) to calibrate "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--cpu-rows", type=int, default=500_000)
    ap.add_argument("--mode", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--no-extras", action="store_true",
                    help="pod workload: skip the randomized, grouped and serving summaries")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default product-pass launch geometry")
    ap.add_argument("--python-serve", action="store_true", help="serve through the Python loop, not pc_serve_*")
    ap.add_argument("--rand-r", type=int, default=74, help="randomized sketch width (64 modes + 10)")
    ap.add_argument("--rand-q", type=int, default=3, help="randomized power iterations")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: Poisson arrival rate, requests/s")
    ap.add_argument("--duration", type=float, default=20.0, help="serve: arrival window, s")
    ap.add_argument("--slots", type=int, default=4, help="serve: requests in flight")
    ap.add_argument("--workload", default="pod", choices=["pod", "grouped", "spod", "serve", "copy"],
                    help="pod: tall-skinny SVD 8M x 256 (default); grouped: SPOD 200000 x 2048 "
                         "(129 complex64 200000 x 15 blocks) + 64 uneven real POD problems; serve: "
                         "Poisson stream of POD / project / reconstruct requests; copy: one plain device "
                         "copy of an 8M x 256 float32 matrix (counter calibration)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def generate(m, n, seed=2024):
    """Low-rank-plus-noise matrix in float64 on the device, in row chunks."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    G2 = torch.linalg.qr(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))[0]
    s = 1e3 * 0.93 ** torch.arange(n, device="cuda", dtype=torch.float64)
    B = (s[:, None] * G2.T).contiguous()   # X = G1 diag(s) G2^T
    X = torch.empty((m, n), device="cuda", dtype=torch.float64)
    step = 1 << 20
    inv = 1.0 / np.sqrt(m)
    for r0 in range(0, m, step):
        r1 = min(m, r0 + step)
        G1 = torch.randn(r1 - r0, n, generator=g, device="cuda", dtype=torch.float64) * inv
        X[r0:r1] = G1 @ B + 1e-6 * torch.randn(r1 - r0, n, generator=g, device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    return X


class Solver:
    """Direct C-ABI calls (no Python-side allocation inside the timed loop)."""

    def __init__(self, X, precision, autotune=False):
        import torch
        from pyloworder_amd import _lib
        from pyloworder_amd._tensor import pc_dtype, real_dtype
        self.torch = torch
        self.ctx = _lib.context()
        # product-pass launch geometry tuned on the first solve (warm-up); the
        # tuner changes speed only, never the result
        self.ctx.check(self.ctx.lib.pc_autotune(self.ctx.h, 1 if autotune else 0))
        self.X = X
        m, n = X.shape
        self.m, self.n = m, n
        self.U = torch.empty((m, n), dtype=X.dtype, device="cuda")
        self.S = torch.empty((n,), dtype=real_dtype(X.dtype), device="cuda")
        self.V = torch.empty((n, n), dtype=X.dtype, device="cuda")
        self.dt = pc_dtype(X)
        self.prec = _lib.PC_PREC_MIXED if precision == "mixed" else _lib.PC_PREC_FP64

    def __call__(self):
        c = self.ctx
        c.check(c.lib.pc_tsqr_svd(c.h, self.dt, self.X.data_ptr(), self.X.stride(0), self.m, self.n, 1,
                                  self.U.data_ptr(), self.U.stride(0), self.S.data_ptr(), self.V.data_ptr(),
                                  self.V.stride(0), self.prec))

    def tuned_chunks(self):
        import ctypes as C
        v = C.c_int64(0)
        return v.value if self.ctx.lib.pc_autotune_query(self.dt, self.m, self.n, self.prec, C.byref(v)) == 1 else None

    def profile(self, on):
        self.ctx.check(self.ctx.lib.pc_profile(self.ctx.h, 1 if on else 0))

    def profile_read(self):
        import ctypes as C
        ms = (C.c_double * 5)()
        cnt = (C.c_int64 * 5)()
        self.ctx.check(self.ctx.lib.pc_profile_read(self.ctx.h, ms, cnt, 5))
        return list(ms), list(cnt)


class RandSolver:
    """Direct C-ABI randomized_svd (pyLOM/vmmath/svd.py:254-273), outputs
    preallocated; Omega drawn on the device from the seed."""

    def __init__(self, X, r, q, seed=2024, precision="mixed"):
        import torch
        from pyloworder_amd import _lib
        from pyloworder_amd._tensor import pc_dtype, real_dtype
        self.ctx = _lib.context()
        self.X = X
        m, n = X.shape
        self.m, self.n, self.r, self.q, self.seed = m, n, r, q, seed
        self.U = torch.empty((m, r), dtype=X.dtype, device="cuda")
        self.S = torch.empty((r,), dtype=real_dtype(X.dtype), device="cuda")
        self.V = torch.empty((r, n), dtype=X.dtype, device="cuda")
        self.dt = pc_dtype(X)
        self.prec = _lib.PC_PREC_MIXED if precision == "mixed" else _lib.PC_PREC_FP64

    def __call__(self):
        c = self.ctx
        c.check(c.lib.pc_randomized_svd(c.h, self.dt, self.X.data_ptr(), self.X.stride(0), self.m, self.n, self.r,
                                        self.q, self.seed, None, self.U.data_ptr(), self.U.stride(0),
                                        self.S.data_ptr(), self.V.data_ptr(), self.V.stride(0), self.prec))


def recon_rmse_device(X, U, S, V, center):
    """Relative reconstruction RMSE (pyLOM/vmmath/stats.py:16-34) of the
    (row-centred) data against U diag(S) V, computed on the device in FP64
    row chunks."""
    import torch
    num = den = 0.0
    step = 1 << 20
    for r0 in range(0, X.shape[0], step):
        Y = X[r0:r0 + step].double()
        if center:
            Y -= Y.mean(dim=1, keepdim=True)
        R = (U[r0:r0 + step].double() * S.double()) @ V.double()
        num += float(((Y - R) ** 2).sum())
        den += float((Y ** 2).sum())
    return float(np.sqrt(num / den))


def randomized_floor(m, n, r, q, b):
    """Algorithmic bytes of the reference's randomized_svd (2q + 2 passes over
    A, q + 1 sketches written and read, U written) and its HBM floor."""
    nbytes = (2 * q + 2) * m * n * b + 2 * (q + 1) * m * r * b + m * r * b
    return nbytes, nbytes / (HBM_PEAK_GBS * 1e9)


def time_solves(solver, steps, warmup):
    import torch
    for _ in range(warmup):
        solver()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(st)
        solver()
        b.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = sorted(a.elapsed_time(b) for a, b in evs)
    return wall / steps * 1e3, lat


def stage_profile(solver, reps=3):
    solver.profile(True)
    for _ in range(reps):
        solver()
    ms, cnt = solver.profile_read()
    solver.profile(False)
    return {k: (ms[i] / max(cnt[i], 1), cnt[i]) for i, k in enumerate(["gram", "zgram", "gemm"])}


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "mixed_passes_8Mx256_traffic.md")


def measured_traffic(kernel, m, n):
    """HBM bytes per launch of a mixed pass, from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy calibration);
    None when the summary does not cover this kernel at this size."""
    if (m, n) != (M_DEFAULT, N_DEFAULT) or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == f"`{kernel}`":
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def pass_flops(name, m, n):
    """fp16 MFMA flops a mixed pass issues (n = 32 NBLK padded columns): P1
    G = Ys^T Ys, three split products on the NU = NBLK (NBLK + 1) / 2 upper
    32 x 32 blocks: 3 NU 2048 m; P2 Z = Ys W1 (hi and lo of Ys against fp16
    W1, 2 products) over the nonzero k range of the upper triangular W1
    (32 (J + 1) rows for column block J): 2 NU 2048 m, plus the same
    three-product Gram of Z; P3 C = Ys W, three products: 6 m n^2.
    Implementation work, not algorithmic work (rocprofv3's
    SQ_INSTS_VALU_MFMA_MOPS_F16 x 512 reads the same per dispatch)."""
    nblk = -(-n // 32)
    nu = nblk * (nblk + 1) // 2
    gram = 3 * nu * 32 * 32 * 2 * m
    return {"gram": gram, "zgram": 2 * nu * 32 * 32 * 2 * m + gram, "gemm": 6 * m * (32 * nblk) ** 2}[name]


def algorithmic_work(name, m, n, b):
    """(bytes, flops) a pass needs by the algorithm's own count: one read of
    the m x n input (and, for the product pass, one write of U), and the
    products of the tall-skinny SVD, 4 m n^2 per solve: m n^2 for the
    symmetric Gram of P1 and of P2 each, 2 m n^2 for U = Y W in P3."""
    return {"gram": (m * n * b, m * n * n), "zgram": (m * n * b, m * n * n),
            "gemm": (2 * m * n * b, 2 * m * n * n)}[name]


def roofline_for(stage_ms, m, n, b):
    """Dominant streaming kernel (longest average launch) against the bound
    its algorithmic work sets: the larger of algorithmic bytes / 8 TB/s and
    algorithmic flops / the fp16 dense peak.  The fp16 split issues about
    twice the algorithmic flops; that rate is reported beside it
    (``split_mfma``), not used for the bound."""
    kern = {"gram": "k_gram_mx2", "zgram": "k_zgram_mx2", "gemm": "k_gemm_mx2"}
    name = max(stage_ms, key=lambda k: stage_ms[k][0])
    t = stage_ms[name][0] * 1e-3
    nbytes, flops = algorithmic_work(name, m, n, b)
    gbs = nbytes / t / 1e9
    tfs = flops / t / 1e12
    issued = pass_flops(name, m, n)
    common = {"traffic": measured_traffic(kern[name], m, n),
              "traffic_source": "profiles/mixed_passes_8Mx256_traffic.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE "
                                "per dispatch, copy-calibrated); null at other sizes",
              "kernel": kern[name], "algorithmic_bytes": nbytes, "algorithmic_flops": flops,
              "avg_ms": round(t * 1e3, 3), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
              "mfma_frac": round(tfs / MFMA_F16_PEAK_TFS, 4),
              "split_mfma": {"issued_flops": issued, "tflops": round(issued / t / 1e12, 1),
                             "frac_of_fp16_peak": round(issued / t / 1e12 / MFMA_F16_PEAK_TFS, 4)}}
    if flops / (MFMA_F16_PEAK_TFS * 1e12) > nbytes / (HBM_PEAK_GBS * 1e9):
        return name, {"bound": "mfma", "achieved": round(tfs, 1), "peak": MFMA_F16_PEAK_TFS, "unit": "TFLOP/s",
                      "frac": round(tfs / MFMA_F16_PEAK_TFS, 4), **common}
    return name, {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(gbs / HBM_PEAK_GBS, 4), **common}


def cpu_baseline(m_sample, n, seed=7):
    """The reference's algorithm on the host cores (the oracle's numpy
    restatement): an MPI-rank stand-in (one process per core, one BLAS thread
    each, leaf QR + butterfly R-tree + svd(R) + Q U_R, shared-memory row
    blocks) as the reported value, and one threaded-BLAS process beside it;
    medians of 5 and 3 runs after a warm-up, on bounded row samples."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import cpu_baseline as cb
    from oracle import workloads
    nproc = min(16, cb.usable_cores())
    X = workloads.lowrank(m_sample, n, seed=seed, decay=0.93, top=1e3, noise=1e-6, dtype=np.float32)
    t_mpi, runs_mpi = cb.mpi_standin(X, nproc, runs=5)
    m_thr = max(n, m_sample // 5)
    t_thr, runs_thr = cb.threaded(X[:m_thr], runs=3)
    gbs = 3 * m_sample * n * 4 / t_mpi / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": nproc, "kind": "port",
            "sample": f"POD.run(remove_mean=True) of the reference algorithm (oracle restatement) on {m_sample} x {n} "
                      f"float32 rows of the same synthetic low-rank matrix, {nproc} processes as MPI ranks "
                      f"(OPENBLAS_NUM_THREADS=1 each); median of 5 runs {t_mpi:.3f} s",
            "cpu_model": cb.cpu_model(), "runs_s": [round(t, 3) for t in runs_mpi],
            "threaded_blas": {"gbs": round(3 * m_thr * n * 4 / t_thr / 1e9, 3), "rows": m_thr,
                              "median_s": round(t_thr, 3), "runs_s": [round(t, 3) for t in runs_thr],
                              "threads": os.environ.get("OMP_NUM_THREADS", str(os.cpu_count()))}}


def waves_device(m, n, seed=4, noise=1e-2):
    """Six superposed travelling waves + N(0, noise), float32, on the device."""
    import torch
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.arange(m, device="cuda", dtype=torch.float64)[:, None] / m
    t = torch.arange(n, device="cuda", dtype=torch.float64)[None, :] * 0.01
    X = torch.zeros((m, n), device="cuda", dtype=torch.float32)
    amp = [1.0, 0.7, 0.5, 0.35, 0.25, 0.15]
    kap = [6.0, 11.0, 17.0, 23.0, 31.0, 40.0]
    fr = [7.8125, 15.625, 23.4375, 31.25, 39.0625, 46.875]
    for a_, k_, f_ in zip(amp, kap, fr):
        X += (a_ * torch.sin(2 * np.pi * k_ * x - 2 * np.pi * f_ * t + rng.uniform(0, 2 * np.pi))).float()
    X += noise * torch.randn((m, n), generator=g, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    return X, (np.arange(n) * 0.01)


def uneven_problems(count, seed=77):
    """count real float32 problems: m log-uniform in [5e4, 2e6], n uniform in
    [16, 256], low-rank + noise with decay U(0.9, 0.97)."""
    import torch
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = []
    for _ in range(count):
        m = int(round(np.exp(rng.uniform(np.log(5e4), np.log(2e6)))))
        n = int(rng.integers(16, 257))
        d = rng.uniform(0.9, 0.97)
        G2 = torch.linalg.qr(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))[0]
        s = 1e3 * d ** torch.arange(n, device="cuda", dtype=torch.float64)
        G1 = torch.randn(m, n, generator=g, device="cuda", dtype=torch.float64) / np.sqrt(m)
        X = (G1 @ (s[:, None] * G2.T)).float()
        X += 1e-6 * torch.randn(m, n, generator=g, device="cuda", dtype=torch.float32)
        out.append(X)
        del G1
    torch.cuda.synchronize()
    return out


def run_grouped(a, precisions=("mixed", "fp64"), emit=True, separate=True):
    """SPOD (nDFT 256, 50% overlap) of 200000 x 2048 float32 + 64 uneven PODs.
    Prints the JSON line when ``emit``, and returns it."""
    import ctypes as C
    import torch
    import pyloworder_amd as pl
    from pyloworder_amd import _lib
    ctx = _lib.context()
    M, N, nDFT, nolap = 200_000, 2048, 256, 128
    log(f"[bench] grouped: generating SPOD input {M} x {N} and 64 uneven problems")
    X, t = waves_device(M, N)
    probs_X = uneven_problems(64)
    nf, nB = nDFT // 2 + 1, (N - nolap) // (nDFT - nolap)
    spod_bytes = nf * 3 * M * nB * 8
    real_bytes = sum(3 * P.shape[0] * P.shape[1] * 4 for P in probs_X)
    res = {}
    sync = torch.cuda.synchronize
    for prec in precisions:
        # full SPOD.run (blocks + grouped per-frequency SVDs + sort/P layout)
        # results are released inside the loops so that each call reuses the
        # caching allocator's blocks instead of allocating beside a live result
        for _ in range(a.warmup):
            pl.SPOD.run(X, t, nDFT, nolap, precision=prec)
        sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            L, P, f = pl.SPOD.run(X, t, nDFT, nolap, precision=prec)
            del L, P, f
        sync()
        t_spod = (time.perf_counter() - t0) / a.steps
        # the SVD stage alone: one grouped call over the 129 blocks
        Q, _, _ = pl.SPOD.blocks(X, t, nDFT, nolap)
        for _ in range(max(2, a.warmup)):  # the first calls size the workspaces
            pl.math.tsqr_svd_grouped(Q, precision=prec)
        sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            Us, Ss, Vs = pl.math.tsqr_svd_grouped(Q, precision=prec)
            del Us, Ss, Vs
        sync()
        t_blocksvd = (time.perf_counter() - t0) / a.steps
        # the same stage recorded once into a HIP graph and replayed
        g = pl.math.GroupedGraph(Q, precision=prec)
        g.launch()
        sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            g.launch()
        sync()
        t_graph = (time.perf_counter() - t0) / a.steps
        g.close()
        del g, Q
        # 64 real problems (row centring fused) in one grouped call, outputs preallocated
        outs = [(torch.empty_like(P_), torch.empty(P_.shape[1], device="cuda"),
                 torch.empty((P_.shape[1], P_.shape[1]), device="cuda")) for P_ in probs_X]
        arr = (_lib.PCProblem * len(probs_X))()
        for i, (P_, (U_, S_, V_)) in enumerate(zip(probs_X, outs)):
            arr[i] = _lib.PCProblem(P_.data_ptr(), P_.stride(0), P_.shape[0], P_.shape[1], 1, 0, U_.data_ptr(),
                                    U_.stride(0), S_.data_ptr(), V_.data_ptr(), V_.stride(0))
        pc = _lib.PC_PREC_MIXED if prec == "mixed" else _lib.PC_PREC_FP64

        def grouped():
            ctx.check(ctx.lib.pc_grouped_tsqr_svd(ctx.h, _lib.PC_F32, C.cast(arr, C.c_void_p), len(probs_X), pc))
        for _ in range(a.warmup):
            grouped()
        sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            grouped()
        sync()
        t_real = (time.perf_counter() - t0) / a.steps
        # the same 64 problems as separate calls, for comparison
        t_single = float("nan")
        if separate:
            t0 = time.perf_counter()
            for P_, (U_, S_, V_) in zip(probs_X, outs):
                ctx.check(ctx.lib.pc_tsqr_svd(ctx.h, _lib.PC_F32, P_.data_ptr(), P_.stride(0), P_.shape[0],
                                              P_.shape[1], 1, U_.data_ptr(), U_.stride(0), S_.data_ptr(),
                                              V_.data_ptr(), V_.stride(0), pc))
            sync()
            t_single = time.perf_counter() - t0
        # reconstruction RMSE of every real problem (centred rows, as solved)
        rm = []
        for P_, (U_, S_, V_) in zip(probs_X, outs):
            Yc = P_.double() - P_.double().mean(dim=1, keepdim=True)
            R_ = (U_.double() * S_.double()) @ V_.double()
            rm.append(float(torch.linalg.norm(Yc - R_) / torch.linalg.norm(Yc)))
            del Yc, R_
        del outs
        tot = t_spod + t_real
        res[prec] = {"spod_run_ms": round(t_spod * 1e3, 2), "spod_svd_stage_ms": round(t_blocksvd * 1e3, 2),
                     "spod_svd_stage_graph_ms": round(t_graph * 1e3, 2),
                     "spod_svd_stage_gbs": round(spod_bytes / t_blocksvd / 1e9, 1),
                     "real64_grouped_ms": round(t_real * 1e3, 2),
                     "real64_separate_calls_ms": None if t_single != t_single else round(t_single * 1e3, 2),
                     "real64_max_recon_rmse": max(rm),
                     "real64_gbs": round(real_bytes / t_real / 1e9, 1),
                     "solves_per_s": round((nf + len(probs_X)) / tot, 1),
                     "gbs": round((spod_bytes + real_bytes) / tot / 1e9, 1)}
        log(f"[bench] grouped {prec}: {res[prec]}")
    out = {"metric": "grouped SPOD + uneven POD solves/s (cfg: 129 x 200000x15 c64 + 64 real)",
           "value": res["mixed"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(1e3 * (nf + 64) / res["mixed"]["solves_per_s"], 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c64/f32 storage, fp16x2 MFMA / FP64 small", "data": "synthetic (device-generated)",
           "config": {"workload": "SPOD.run nDFT=256 nolap=128 on 200000 x 2048 float32 + 64 real POD problems "
                                  "m log-U[5e4, 2e6], n U[16, 256]", "nf": nf, "nBlks": nB,
                      "algorithmic_GB": round((spod_bytes + real_bytes) / 1e9, 2)},
           "mixed": res.get("mixed"), "fp64": res.get("fp64")}
    if emit:
        print(json.dumps(out), flush=True)
    del X, probs_X
    torch.cuda.empty_cache()
    return out


def pool_device(m, n, decay, seed):
    """Low-rank + noise float32 pool (as the 8M x 256 workload, decay given)."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    G2 = torch.linalg.qr(torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64))[0]
    s = 1e3 * decay ** torch.arange(n, device="cuda", dtype=torch.float64)
    B = (s[:, None] * G2.T).contiguous()
    X = torch.empty((m, n), device="cuda", dtype=torch.float32)
    step = 1 << 20
    for r0 in range(0, m, step):
        r1 = min(m, r0 + step)
        G1 = torch.randn(r1 - r0, n, generator=g, device="cuda", dtype=torch.float64) / np.sqrt(m)
        X[r0:r1] = (G1 @ B).float() + 1e-6 * torch.randn(r1 - r0, n, generator=g, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    return X


def run_serve(a, emit=True):
    """Poisson arrivals of POD solves (50 %), projections (25 %) and
    reconstructions (25 %) on resident bases; m log-uniform in [1e4, 4e6],
    n log-uniform in [8, 512]; float32, mixed path for n <= 256 and FP64 MFMA
    beyond (and for project / reconstruct); 200 GB admission budget."""
    import threading
    import torch
    from pyloworder_amd import serve
    from pyloworder_amd.serve import POD, PROJECT, RECONSTRUCT, Request, Server
    M_MAX, N_MAX = 4_000_000, 512
    log(f"[bench] serve: generating 3 input pools {M_MAX} x {N_MAX} float32")
    pools = [pool_device(M_MAX, N_MAX, d, 100 + i) for i, d in enumerate((0.9, 0.935, 0.97))]
    rng = np.random.default_rng(12)
    ex = serve.DeviceExecutor("mixed")
    reserve = ex.workspace_bound(M_MAX, N_MAX)
    algo = {"bytes": 0}
    lock = threading.Lock()

    def on_done(r):
        if r.error is not None:
            return
        if r.kind == POD:
            m, n = r.X.shape
            b = 3 * m * n * 4
        elif r.kind == PROJECT:
            m, n = r.X.shape
            b = (m * r.modes + m * n + r.modes * n) * 4
        else:
            U = r.result
            m, n = U.shape
            b = (m * r.modes + r.modes * n + m * n) * 4
        with lock:
            algo["bytes"] += b
    srv = Server(ex, slots=a.slots, budget=200e9, reserve_per_slot=reserve, keep_results=False, on_done=on_done)

    def logu(lo, hi):
        return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))

    def make(kind):
        if kind != POD:
            ids = srv.resident_ids()
            if ids:
                bid = ids[int(rng.integers(len(ids)))]
                b = srv.bases.get(bid)
                if b is not None:
                    m, N = b.U.shape
                    k = int(min(N, rng.integers(4, 65)))
                    if kind == PROJECT:
                        n2 = int(round(logu(8, N_MAX)))
                        return Request(PROJECT, X=pools[int(rng.integers(3))][:m, :n2], basis=bid, modes=k)
                    return Request(RECONSTRUCT, basis=bid, modes=k)
        m = int(round(logu(1e4, M_MAX)))
        n = int(round(logu(8, N_MAX)))
        n = min(n, m)
        return Request(POD, X=pools[int(rng.integers(3))][:m, :n])

    # warm-up: allocate every slot's workspace and the caching allocator's pools
    for _ in range(2 * a.slots):
        srv.submit(Request(POD, X=pools[0][:M_MAX, :N_MAX]))
    srv.drain()
    srv.completed.clear()
    algo["bytes"] = 0
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / a.rate)
        if t > a.duration:
            break
        arrivals.append(t)
    kinds = rng.choice([POD, PROJECT, RECONSTRUCT], size=len(arrivals), p=[0.5, 0.25, 0.25])
    log(f"[bench] serve: {len(arrivals)} requests over {a.duration} s at {a.rate}/s, {a.slots} slots")
    t0 = time.perf_counter()
    late = 0
    for ta, kind in zip(arrivals, kinds):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        elif now - ta > 0.01:
            late += 1
        r = make(str(kind))
        r.arrival = t0 + ta
        srv.submit(r)
    srv.drain()
    wall = max(r.done for r in srv.completed) - t0
    srv.close()
    summ = serve.summary(srv, wall)
    okr = [r for r in srv.completed if r.error is None]
    summ["queue_p50_ms"] = serve.percentile([(r.admitted - r.arrival) * 1e3 for r in okr], 50)
    summ["run_p50_ms"] = serve.percentile([(r.done - r.admitted) * 1e3 for r in okr], 50)
    gbs = algo["bytes"] / wall / 1e9
    out = {"metric": "serving: solves/s and p50/p99 latency (POD / project / reconstruct stream)",
           "value": summ["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": len(arrivals), "warmup": 2 * a.slots,
           "ms_per_step": round(wall * 1e3 / max(1, summ["requests"]), 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None,
           "dtype": "f32 storage; POD n <= 256: fp16x2 MFMA / FP64 small, n > 256: FP64 MFMA; project / reconstruct: "
                    "FP64 MFMA",
           "data": "synthetic (device-generated)",
           "config": {"workload": "Poisson arrivals, 50% POD.run / 25% project / 25% reconstruct on resident bases, "
                                  "m log-U[1e4, 4e6], n log-U[8, 512], decay in {0.9, 0.935, 0.97}, float32",
                      "rate_per_s": a.rate, "duration_s": a.duration, "slots": a.slots, "budget_GB": 200,
                      "workspace_reserve_per_slot_GB": round(reserve / 1e9, 3)},
           "p50_ms": summ["p50_ms"], "p99_ms": summ["p99_ms"], "serve": summ,
           "algorithmic_gbs": round(gbs, 1), "late_submissions": late}
    if emit:
        print(json.dumps(out), flush=True)
    del pools
    torch.cuda.empty_cache()
    return out


def run_serve_native(a, emit=True):
    """The workload of run_serve through the native request loop (pc_serve_*):
    admission, eviction, slots and streams run in the library's C++ threads."""
    import torch
    from pyloworder_amd.serve import NativeServer, percentile
    M_MAX, N_MAX = 4_000_000, 512
    log(f"[bench] serve (native): generating 3 input pools {M_MAX} x {N_MAX} float32")
    pools = [pool_device(M_MAX, N_MAX, d, 100 + i) for i, d in enumerate((0.9, 0.935, 0.97))]
    rng = np.random.default_rng(12)
    srv = NativeServer(slots=a.slots, budget=200e9, precision="mixed", m_max=M_MAX, n_max=N_MAX)

    def logu(lo, hi):
        return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))

    reqs = []         # (ticket, kind, algorithmic bytes)
    outs = []         # projection / reconstruction tickets whose output is not released yet
    recs = {}         # ticket -> record, kept when an output is released

    def release_done():
        # outputs count against the budget until released: copy nothing, just
        # record and release what has finished
        still = []
        for t in outs:
            rec = srv.wait(t, 0.0)
            if rec is None:
                still.append(t)
                continue
            recs[t] = rec
            if rec["status"] == 0:
                srv.release(t)
        outs[:] = still

    def submit(kind):
        ids = srv.resident() if kind != "pod" else []
        shape = None
        if ids:
            bid = ids[int(rng.integers(len(ids)))]
            shape = srv.basis_shape(bid)
        if shape is not None:
            m, N = shape
            k = int(min(N, rng.integers(4, 65)))
            if kind == "project":
                n2 = int(round(logu(8, N_MAX)))
                t = srv.submit_project(bid, pools[int(rng.integers(3))][:m, :n2], modes=k)
                outs.append(t)
                return t, kind, (m * k + m * n2 + k * n2) * 4
            t = srv.submit_reconstruct(bid, modes=k)
            outs.append(t)
            return t, "reconstruct", (m * k + k * N + m * N) * 4
        m = int(round(logu(1e4, M_MAX)))
        n = min(int(round(logu(8, N_MAX))), m)
        t = srv.submit_pod(pools[int(rng.integers(3))][:m, :n])
        return t, "pod", 3 * m * n * 4

    # warm-up: the stream-ordered allocator's pools
    warm = [srv.submit_pod(pools[0][:M_MAX, :N_MAX]) for _ in range(2 * a.slots)]
    srv.drain()
    for t in warm:
        srv.wait(t)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / a.rate)
        if t > a.duration:
            break
        arrivals.append(t)
    kinds = rng.choice(["pod", "project", "reconstruct"], size=len(arrivals), p=[0.5, 0.25, 0.25])
    log(f"[bench] serve (native): {len(arrivals)} requests over {a.duration} s at {a.rate}/s, {a.slots} slots")
    t0 = time.perf_counter()
    late = 0
    for ta, kind in zip(arrivals, kinds):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        elif now - ta > 0.01:
            late += 1
        release_done()
        reqs.append(submit(str(kind)))
    t_wait = time.perf_counter()
    while outs and time.perf_counter() - t_wait < 120:
        release_done()
        time.sleep(0.002)
    if outs or not srv.drain(timeout=120):
        raise RuntimeError(f"serve (native): requests still open after 120 s: {srv.stats()}")
    t_end = time.perf_counter() - t0
    lat = {"pod": [], "project": [], "reconstruct": []}
    ok = rejected = errors = 0
    queue, run = [], []
    why = {}
    algo = 0
    first_submit = last_done = None
    for t, kind, b in reqs:
        rec = recs[t] if t in recs else srv.wait(t)
        if rec["status"] != 0:
            errors += 1
            if rec["t_admit"] == 0.0:
                rejected += 1
            key = f"{kind}: {rec.get('error', rec['status'])}"
            why[key] = why.get(key, 0) + 1
            continue
        ok += 1
        algo += b
        lat[kind].append(rec["latency"] * 1e3)
        queue.append((rec["t_admit"] - rec["t_submit"]) * 1e3)
        run.append((rec["t_done"] - rec["t_admit"]) * 1e3)
        first_submit = rec["t_submit"] if first_submit is None else min(first_submit, rec["t_submit"])
        last_done = rec["t_done"] if last_done is None else max(last_done, rec["t_done"])
    wall = (last_done - first_submit) if ok else t_end
    st = srv.stats()
    srv.close()
    all_lat = lat["pod"] + lat["project"] + lat["reconstruct"]
    summ = {"requests": len(reqs), "ok": ok, "rejected": rejected, "errors": errors, "evictions": st["evictions"],
            "wall_s": round(wall, 3), "p50_ms": percentile(all_lat, 50), "p99_ms": percentile(all_lat, 99),
            "solves_per_s": round(len(lat["pod"]) / wall, 2), "requests_per_s": round(ok / wall, 2),
            "queue_p50_ms": percentile(queue, 50), "run_p50_ms": percentile(run, 50), "failures": why}
    for k in lat:
        summ[k] = {"count": len(lat[k]), "p50_ms": percentile(lat[k], 50), "p99_ms": percentile(lat[k], 99)}
    out = {"metric": "serving: solves/s and p50/p99 latency (POD / project / reconstruct stream)",
           "value": summ["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": len(arrivals), "warmup": 2 * a.slots,
           "ms_per_step": round(wall * 1e3 / max(1, ok), 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None,
           "dtype": "f32 storage; POD n <= 256: fp16x2 MFMA / FP64 small, n > 256: FP64 MFMA; project / reconstruct: "
                    "FP64 MFMA",
           "data": "synthetic (device-generated)",
           "config": {"workload": "Poisson arrivals, 50% POD.run / 25% project / 25% reconstruct on resident bases, "
                                  "m log-U[1e4, 4e6], n log-U[8, 512], decay in {0.9, 0.935, 0.97}, float32",
                      "rate_per_s": a.rate, "duration_s": a.duration, "slots": a.slots, "budget_GB": 200,
                      "runtime": "native pc_serve_*"},
           "p50_ms": summ["p50_ms"], "p99_ms": summ["p99_ms"], "serve": summ,
           "algorithmic_gbs": round(algo / wall / 1e9, 1), "late_submissions": late}
    if emit:
        print(json.dumps(out), flush=True)
    del pools
    torch.cuda.empty_cache()
    return out


def run_spod(a):
    """SPOD.run alone (nDFT 256, 50% overlap) on 200000 x 2048 float32: the
    Welch blocks, one batched grouped solve over the 129 frequencies and the
    sorted modes.  For per-kernel traces of that path."""
    import torch
    import pyloworder_amd as pl
    M, N, nDFT, nolap = 200_000, 2048, 256, 128
    X, t = waves_device(M, N)
    res = {}
    for prec in ("mixed", "fp64"):
        for _ in range(a.warmup):
            pl.SPOD.run(X, t, nDFT, nolap, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            L, P, f = pl.SPOD.run(X, t, nDFT, nolap, precision=prec)
            del L, P, f
        torch.cuda.synchronize()
        res[prec] = round((time.perf_counter() - t0) / a.steps * 1e3, 3)
        log(f"[bench] spod {prec}: {res[prec]} ms per SPOD.run")
    print(json.dumps({"workload": "spod", "ms_per_run": res, "steps": a.steps, "warmup": a.warmup,
                      "config": {"M": M, "N": N, "nDFT": nDFT, "nolap": nolap}}), flush=True)


def run_copy(a):
    """One plain copy of an m x n float32 matrix: known bytes (m n 4 read, m n 4
    written) to calibrate the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    X = torch.empty((a.m, a.n), dtype=torch.float32, device="cuda")
    X.fill_(1.0)
    Y = torch.empty_like(X)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Y.copy_(X)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    b = X.numel() * 4
    print(json.dumps({"workload": "copy", "bytes_read": b, "bytes_written": b, "ms": round(t * 1e3, 3),
                      "gbs": round(2 * b / t / 1e9, 1)}), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (see DESIGN.md for sharding)"}))
        sys.exit(2)
    if a.workload == "grouped":
        return run_grouped(a)
    if a.workload == "serve":
        return run_serve(a) if a.python_serve else run_serve_native(a)
    if a.workload == "spod":
        return run_spod(a)
    if a.workload == "copy":
        return run_copy(a)
    import torch
    import pyloworder_amd  # noqa: F401  (fails loudly if libpodcore.so is missing)
    m, n = a.m, a.n
    log(f"[bench] generating {m} x {n} on the device")
    X64 = generate(m, n)
    out = {}
    res64 = None
    if a.mode in ("both", "fp64") and not a.no_fp64:
        log("[bench] fp64 path")
        s64 = Solver(X64, "fp64", autotune=not a.no_autotune)
        ms64, lat64 = time_solves(s64, a.steps, a.warmup)
        res64 = {"ms_per_solve": round(ms64, 3), "solves_per_s": round(1e3 / ms64, 3),
                 "hbm_gbs": round(3 * m * n * 8 / (ms64 * 1e-3) / 1e9, 1),
                 "p50_ms": round(lat64[len(lat64) // 2], 3), "p99_ms": round(lat64[min(len(lat64) - 1, int(0.99 * len(lat64)))], 3),
                 "stages_ms": {k: round(v[0], 3) for k, v in stage_profile(s64).items()},
                 "product_pass_chunks": s64.tuned_chunks()}
        log(f"[bench] fp64: {res64}")
        del s64
    X32 = X64.float()
    if a.copy_calib:
        Yc = torch.empty_like(X32)
        Yc.copy_(X32)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: {X32.numel() * 4} bytes read, {X32.numel() * 4} bytes written")
        del Yc
    del X64
    torch.cuda.empty_cache()
    log("[bench] mixed path")
    smx = Solver(X32, "mixed", autotune=not a.no_autotune)
    msmx, latmx = time_solves(smx, a.steps, a.warmup)
    tuned = smx.tuned_chunks()
    log(f"[bench] mixed: {msmx:.3f} ms per solve, product pass chunks {tuned}")
    stages = stage_profile(smx)
    dom, roof = roofline_for(stages, m, n, 4)
    # accuracy of the mixed result against FP64 math on the same stored data
    log("[bench] reference solve (fp64 math on the float32 data)")
    sref = Solver(X32, "fp64")
    sref()
    torch.cuda.synchronize()
    Sm = smx.S.double().cpu().numpy()
    Sr = sref.S.double().cpu().numpy()
    gated = Sr >= 1e-3 * Sr[0]
    nz = Sr > 1e-10 * Sr[0]
    rel = np.abs(Sm - Sr) / np.where(nz, Sr, 1.0)
    k = int(gated.sum())
    Um = smx.U[:, :k].double()
    Ur = sref.U[:, :k].double()
    P = Um - Ur @ (Ur.T @ Um)
    ang = float(torch.linalg.matrix_norm(P, ord=2).clamp(max=1.0).arcsin())
    del Um, Ur, P
    acc = {"sigma_relerr_gated": float(rel[gated].max()), "n_gated": k,
           "sigma_relerr_all": float(rel[nz].max()), "sigma_normwise": float(np.abs(Sm - Sr).max() / Sr[0]),
           "max_principal_angle_gated": ang,
           "recon_rmse": recon_rmse_device(X32, smx.U, smx.S, smx.V, True),
           "recon_rmse_fp64_math": recon_rmse_device(X32, sref.U, sref.S, sref.V, True)}
    log(f"[bench] mixed accuracy {acc}")
    rand = None
    if not a.no_extras:
        # randomized_svd r = 74 (64 modes + 10), q = 3 on the same float32 data
        log(f"[bench] randomized mixed r={a.rand_r} q={a.rand_q}")
        rs = RandSolver(X32, a.rand_r, a.rand_q)
        msr, latr = time_solves(rs, a.steps, a.warmup)
        kk = min(64, a.rand_r)
        Sr_ = rs.S.double().cpu().numpy()
        relr = np.abs(Sr_[:kk] - Sr[:kk]) / Sr[:kk]
        Ur = sref.U[:, :kk].double()
        Pm = rs.U[:, :kk].double() - Ur @ (Ur.T @ rs.U[:, :kk].double())
        angr = float(torch.linalg.matrix_norm(Pm, ord=2).clamp(max=1.0).arcsin())
        del Ur, Pm
        nbytes, floor_s = randomized_floor(m, n, a.rand_r, a.rand_q, 4)
        rand = {"r": a.rand_r, "q": a.rand_q, "kept_modes": kk, "ms_per_solve": round(msr, 3),
                "solves_per_s": round(1e3 / msr, 3), "p50_ms": round(latr[len(latr) // 2], 3),
                "p99_ms": round(latr[min(len(latr) - 1, int(0.99 * len(latr)))], 3),
                "algorithmic_bytes": nbytes, "hbm_floor_ms": round(floor_s * 1e3, 3),
                "floor_frac": round(floor_s * 1e3 / msr, 4), "algorithmic_gbs": round(nbytes / (msr * 1e-3) / 1e9, 1),
                "method": "two mixed Gram passes give A = Q_Y M; the reference's randomized SVD (same Omega, "
                          "q power steps) of the 256 x 256 FP64 factor M; one mixed product pass U = A W",
                "sigma_relerr_kept_vs_full_svd": float(relr.max()),
                "max_principal_angle_kept_vs_full_svd": angr,
                "recon_rmse": recon_rmse_device(X32, rs.U, rs.S, rs.V, False)}
        log(f"[bench] randomized: {rand}")
        del rs
    del sref
    value = 3 * m * n * 4 / (msmx * 1e-3) / 1e9
    grouped = served = None
    if not a.no_extras:
        del smx, X32
        torch.cuda.empty_cache()
        ga = argparse.Namespace(**vars(a))
        ga.steps, ga.warmup = min(a.steps, 3), 1
        log("[bench] grouped summary (mixed)")
        g = run_grouped(ga, precisions=("mixed",), emit=False, separate=False)
        grouped = {"config": g["config"], "solves_per_s": g["value"], **g["mixed"]}
        sa = argparse.Namespace(**vars(a))
        sa.duration = 10.0
        log("[bench] serving summary")
        sv = run_serve(sa, emit=False) if a.python_serve else run_serve_native(sa, emit=False)
        sm = sv["serve"]
        served = {"config": sv["config"], "solves_per_s": sv["value"], "p50_ms": sv["p50_ms"],
                  "p99_ms": sv["p99_ms"], "requests_per_s": sm.get("requests_per_s"),
                  "requests": sm.get("requests"), "rejected": sm.get("rejected"), "errors": sm.get("errors"),
                  "by_kind": {k: sm.get(k) for k in ("pod", "project", "reconstruct") if k in sm},
                  "algorithmic_gbs": sv["algorithmic_gbs"]}
    cpu = None
    if not a.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(a.cpu_rows, n)
    out = {
        "metric": "TSQR-SVD HBM GB/s (mixed precision POD.run, 8M x 256)",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(msmx, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage, fp16x2 MFMA / FP64 small",
        "data": "synthetic (device-generated low-rank + noise, seed 2024)",
        "config": {"workload": "POD.run remove_mean=True on X 8M x 256 float32 (mixed) and float64 (fp64)",
                   "m": m, "n": n, "parallelism": "single GPU"},
        "solves_per_s": round(1e3 / msmx, 3),
        "hbm_roofline_frac_solve": round(value / HBM_PEAK_GBS, 4),
        "p50_ms": round(latmx[len(latmx) // 2], 3),
        "p99_ms": round(latmx[min(len(latmx) - 1, int(0.99 * len(latmx)))], 3),
        "stages_ms": {k_: round(v[0], 3) for k_, v in stages.items()},
        "autotune": None if a.no_autotune else {"product_pass_chunks": tuned},
        "dominant_kernel": dom,
        "roofline": roof,
        "accuracy_mixed_vs_fp64": acc,
        "fp64": res64,
        "mixed_vs_fp64_speedup": round(res64["ms_per_solve"] / msmx, 3) if res64 else None,
        "randomized": rand,
        "grouped": grouped,
        "serving": served,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

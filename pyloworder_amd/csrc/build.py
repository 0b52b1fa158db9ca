"""Build libpodcore.so (HIP, gfx950) in-tree with hipcc.

Each translation unit is compiled to an object in parallel, then linked.
Usage: python -m pyloworder_amd.csrc.build  (or build() from Python).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
LIB = PKG / "libpodcore.so"
SOURCES = ["ctx.cpp", "serve.cpp", "k_stream.hip", "k_mfma_f64.hip", "k_mfma_mx.hip", "k_mfma_mx2.hip", "k_spod.hip", "k_rng.hip", "k_wide.hip", "solve.hip"]
ARCH = os.environ.get("PODCORE_ARCH", "gfx950")


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _flags(debug: bool):
    f = ["-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wno-unused-result",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-ffp-contract=fast"]
    f += ["-O0", "-g"] if debug else ["-O3"]
    return f


def build(verbose: bool = False, debug: bool = False, jobs: int | None = None) -> Path:
    hipcc = _hipcc()
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    flags = _flags(debug)

    def comp(src):
        out = objdir / (Path(src).stem + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [hipcc, *flags, *lang, "-c", str(HERE / src), "-o", str(out)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        if r.returncode != 0:
            raise RuntimeError(f"compile of {src} failed:\n{msg}")
        if verbose and msg.strip():
            print(msg, file=sys.stderr)
        return out

    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, len(SOURCES))) as ex:
        objs = list(ex.map(comp, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, debug="--debug" in sys.argv))

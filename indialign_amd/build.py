"""Build libindialign.so (the HIP library behind the C ABI) for gfx950.

Usage:  python -m indialign_amd.build      (or indialign_amd.build.build())

Every HIP translation unit is compiled with hipcc for gfx950 in parallel,
then linked into indialign_amd/libindialign.so.  Device code is compiled
without FP contraction so FP64 expressions round exactly as written.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libindialign.so")
OBJ_DIR = os.path.join(HERE, "_obj")
SOURCES = ["k_bank.hip", "k_chain.hip", "k_seed.hip", "k_learn.hip", "mlp.hip", "engine.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    errs = [l for l in r.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if errs:
        print(f"[build] {src}: {len(errs)} warnings\n" + "\n".join(errs[:20]), file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"[build] wrote {OUT}")
    return OUT


if __name__ == "__main__":
    build()

"""Build libkizuri_amd.so (HIP for gfx950 + host C++) in-tree with hipcc.

Compile flags that matter for results:
  -ffp-contract=off      no a*b+c fusion: the reference's FP32 build has no FMA
  (no -ffast-math)       IEEE semantics, correctly rounded division and sqrt
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkizuri_amd.so")
SOURCES = ["narrowphase.hip", "compact.hip", "solver.hip", "context.hip", "schedule.cpp", "host_geometry.cpp", "world.cpp",
           "mlp.hip", "batch.hip", "broadphase.hip", "bucket.hip"]

FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++20", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result", "-Wno-unused-value",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = ["hipcc", "-c"] + FLAGS
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += [os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "-shared", "--offload-arch=gfx950", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)

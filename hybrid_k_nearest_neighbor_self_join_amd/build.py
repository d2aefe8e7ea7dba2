"""Builds libgpuknnselfjoin.so (HIP, gfx950) and the `knnjoin_main` CLI in-tree.

    python -m hybrid_k_nearest_neighbor_self_join_amd.build [--dtype float|double|both]

DTYPE of the reference's KNNJoinPy symbol is fixed at build time, as in the reference
(params.h:5, Makefile:61-71): `libgpuknnselfjoin.so` exports KNNJoinPy(float*, ...), and
`double/libgpuknnselfjoin.so` (api.hip compiled with -DKNN_PY_DTYPE_DOUBLE) exports
KNNJoinPy(double*, ...); every other entry point is the same in both.  The default builds both.

Device code is compiled with -ffp-contract=off (no FMA contraction: distances and cell
coordinates are single IEEE operations in a fixed order) and correctly rounded FP32
divide/sqrt, and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgpuknnselfjoin.so")
LIB_DOUBLE = os.path.join(PKG, "double", "libgpuknnselfjoin.so")
CLI = os.path.join(PKG, "knnjoin_main")

SOURCES = ["prims.hip", "estimator.hip", "join.hip", "range.hip", "learned.hip", "batched.hip",
           "outlier.hip", "brute.hip", "hybrid.hip", "api.hip"]
# select.hip is compiled once per part (float / double x three ranges of the dimensionality),
# so its many kernel instantiations build in parallel, and once per bucket count (the learned
# schedule's and the other passes', KNN_SELECT_BUCKETS_LEARNED / _OTHER in knn_launch.h)
SELECT_PARTS = 6
SELECT_BUCKETS = (40, 56)
ARCH = os.environ.get("KNN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", INCLUDE,
    "-I", CSRC,
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(job) -> str:
    src, obj, extra = job
    cmd = [hipcc(), *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _link(objs, out: str, verbose: bool):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(tmp, out)


def build(verbose: bool = False, dtype: str = "both") -> str:
    if dtype not in ("float", "double", "both"):
        raise ValueError("dtype must be float, double or both")
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    common = [s for s in srcs if s != "api.hip"]
    jobs = [(s, os.path.join(BUILD, os.path.splitext(s)[0] + ".o"), []) for s in common]
    jobs += [("select.hip", os.path.join(BUILD, f"select_{p}_b{nb}.o"),
              [f"-DKNN_SELECT_PART={p}", f"-DKNN_SELECT_BUCKETS={nb}"])
             for nb in SELECT_BUCKETS for p in range(SELECT_PARTS)]
    common += [f"select_{p}_b{nb}" for nb in SELECT_BUCKETS for p in range(SELECT_PARTS)]
    if dtype in ("float", "both"):
        jobs.append(("api.hip", os.path.join(BUILD, "api.o"), []))
    if dtype in ("double", "both"):
        jobs.append(("api.hip", os.path.join(BUILD, "api_double.o"), ["-DKNN_PY_DTYPE_DOUBLE"]))
    with cf.ThreadPoolExecutor(max_workers=min(len(jobs), max(2, os.cpu_count() or 2), 16)) as ex:
        objs = list(ex.map(_compile, jobs))
    base = objs[:len(common)]
    if dtype in ("float", "both"):
        _link(base + [os.path.join(BUILD, "api.o")], LIB, verbose)
    if dtype in ("double", "both"):
        _link(base + [os.path.join(BUILD, "api_double.o")], LIB_DOUBLE, verbose)
    cli_src = os.path.join(CSRC, "cli_main.cpp")
    if os.path.exists(cli_src) and dtype in ("float", "both"):
        cmd = [hipcc(), "-O2", "-std=c++17", "-I", INCLUDE, cli_src, "-o", CLI,
               f"-L{PKG}", "-lgpuknnselfjoin", f"-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI link failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", LIB if dtype != "double" else LIB_DOUBLE)
    return LIB if dtype != "double" else LIB_DOUBLE


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--dtype", choices=["float", "double", "both"], default="both",
                    help="DTYPE of KNNJoinPy (params.h:5): float -> libgpuknnselfjoin.so, "
                         "double -> double/libgpuknnselfjoin.so")
    build(verbose=True, dtype=ap.parse_args().dtype)
    sys.exit(0)

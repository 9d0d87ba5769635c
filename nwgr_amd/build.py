"""Build the native libraries of nwgr_amd in-tree.

* ``libnwgr_pagerank.so`` -- HIP kernels + C ABI (include/nwgr_pagerank.h),
  compiled by hipcc for gfx950 only.
* ``libnwgr_graphgen.so`` -- host-only generators / CSR builder and graph
  ingestion (include/nwgr_graphgen.h, include/nwgr_io.h), compiled by g++.

Always recompiles (the build check runs it on a machine without a GPU;
hipcc cross-compiles).  Usage: ``python -m nwgr_amd.build``.
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
BUILD = os.path.join(ROOT, "_build")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC",
    # packed FP32 VALU ops are kept off (hipcc -O3 would form them by itself)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

HIP_SOURCES = ["graph.hip", "pagerank.hip", "weighted.hip", "tiled.hip", "ppr.hip", "multi.hip", "csr.hip",
               "topk.hip", "cache.hip", "fp64.hip", "batched.hip"]

PAGERANK_LIB = os.path.join(PKG, "libnwgr_pagerank.so")
GRAPHGEN_LIB = os.path.join(PKG, "libnwgr_graphgen.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")
    return r.stdout


def _compile_hip(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_pagerank() -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    tmp = PAGERANK_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, PAGERANK_LIB)
    return PAGERANK_LIB


def build_graphgen() -> str:
    cxx = shutil.which("g++") or "g++"
    tmp = GRAPHGEN_LIB + ".tmp"
    _run([cxx, "-std=c++20", "-O3", "-march=x86-64-v3", "-fPIC", "-shared", "-pthread",
          f"-I{INCLUDE}", f"-I{CSRC}", os.path.join(CSRC, "graphgen.cpp"), os.path.join(CSRC, "graphio.cpp"),
          "-o", tmp])
    os.replace(tmp, GRAPHGEN_LIB)
    return GRAPHGEN_LIB


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        futs = [ex.submit(build_pagerank), ex.submit(build_graphgen)]
        for f in futs:
            f.result()


if __name__ == "__main__":
    build_all()
    print("built", PAGERANK_LIB, GRAPHGEN_LIB)

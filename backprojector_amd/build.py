"""Builds backprojector_amd/libbpx.so in-tree with hipcc for gfx950.

Every translation unit is compiled with -ffp-contract=off (the reference path-length formula
rounds each operation separately) and with packed-fp32 instructions disabled.  Objects are
compiled in parallel and linked into one shared library exporting the C ABI of
include/bpx.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbpx.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

SOURCES = [
    "bpx_kernels.hip",
    "bpx_plan.cpp",
    "bpx_api.cpp",
    "bpx_pipeline.cpp",
    "bpx_render.cpp",
    "bpx_compat.cpp",
    "bpx_octree.hip",
    "bpx_octree_plan.cpp",
    "bpx_net.hip",
    "bpx_net.cpp",
]

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(HERE, '..', 'include')}", "-pthread",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found; ROCm is required to build libbpx")
    return p


def _compile(src: str) -> str:
    stem, ext = os.path.splitext(src)
    obj = os.path.join(BUILD, f"{stem}_{ext[1:]}.o")
    lang = ["-x", "hip"]
    cmd = [hipcc(), *lang, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    """Compiles every source and links libbpx.so (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp,
           f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}", "-Wl,-soname,libbpx.so"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

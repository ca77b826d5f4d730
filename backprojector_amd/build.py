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
    "bpx_batch.cpp",
    "bpx_noalloc.hip",
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


CLI_SRC = os.path.join(HERE, "cli", "backprojector.cpp")
CLI = os.path.join(HERE, "bin", "backprojector")
HDF5_PREFIXES = [os.environ.get("HDF5_DIR", ""), "/opt/conda", "/usr", "/usr/local"]


def find_hdf5():
    """(include dir, lib dir) of an HDF5 C library, or None."""
    for pre in HDF5_PREFIXES:
        if not pre:
            continue
        for inc in (os.path.join(pre, "include"), os.path.join(pre, "include", "hdf5", "serial")):
            if not os.path.exists(os.path.join(inc, "hdf5.h")):
                continue
            for lib in (os.path.join(pre, "lib"), os.path.join(pre, "lib", "x86_64-linux-gnu", "hdf5", "serial"),
                        os.path.join(pre, "lib", "x86_64-linux-gnu")):
                if os.path.exists(os.path.join(lib, "libhdf5.so")):
                    return inc, lib
    return None


def build_cli(verbose: bool = True):
    """Builds the backprojector command line (needs an HDF5 C library; skipped without)."""
    h5 = find_hdf5()
    if h5 is None:
        if verbose:
            print("HDF5 not found: backprojector CLI not built")
        return None
    inc, lib = h5
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    tmp = CLI + ".tmp"
    # host-only C++ linked by the toolchain of libbpx.  The HDF5 library is named by path (not
    # -L) and its directory comes last in the run path: a bundled, older C++ runtime there
    # must not shadow the system one that libbpx and the HIP runtime need.
    stdcxx = subprocess.run(["g++", "-print-file-name=libstdc++.so.6"], capture_output=True,
                            text=True).stdout.strip()
    sysdir = os.path.dirname(os.path.realpath(stdcxx)) if os.path.isabs(stdcxx) else "/usr/lib/x86_64-linux-gnu"
    rpath = ":".join(["$ORIGIN/..", sysdir, os.path.join(ROCM, "lib"), lib])
    cmd = [hipcc(), "-x", "c++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(HERE, '..', 'include')}",
           f"-I{inc}", CLI_SRC, "-x", "none", "-o", tmp, f"-L{HERE}", "-lbpx", os.path.join(lib, "libhdf5.so"),
           f"-Wl,-rpath,{rpath}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CLI)
    if verbose:
        print(f"built {CLI}")
    return CLI


def build(verbose: bool = True) -> str:
    """Compiles every source and links libbpx.so (always rebuilds), then the CLI."""
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
    build_cli(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

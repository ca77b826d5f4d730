"""Roofline helpers: a plain copy kernel of known size, used as the HBM bandwidth reference
and to calibrate the FETCH_SIZE / WRITE_SIZE counters that measure a kernel's HBM traffic.

    python -m backprojector_amd.roofline copy --mib 1024 --iters 5
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys

from ._native import check, lib


def copy_benchmark(nbytes: int, iterations: int = 5) -> dict:
    ms = ctypes.c_double()
    check(lib().bpx_copy_benchmark(int(nbytes), int(iterations), ctypes.byref(ms)))
    moved = 2 * (nbytes // 16 * 16)          # read + write
    return {"bytes_copied": nbytes // 16 * 16, "mean_ms": ms.value,
            "GBps_read_plus_write": moved / (ms.value * 1e-3) / 1e9}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["copy"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=5)
    a = ap.parse_args(argv)
    print(json.dumps(copy_benchmark(a.mib << 20, a.iters)), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])

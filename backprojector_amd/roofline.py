"""Roofline measurement: hardware counters of a kernel read with rocprofv3 `--pmc` passes, and the
rooflines they place it against.

* copy kernel of known size: the HBM bandwidth reference, and the calibration of the
  FETCH_SIZE / WRITE_SIZE counters that measure a kernel's HBM traffic
      python -m backprojector_amd.roofline copy --mib 1024 --iters 5
* counter passes: `collect(cmd)` runs `cmd` once per pass under
  `rocprofv3 --pmc <counters> --output-format csv` (each pass in its own process, killed after a
  time limit) and returns the mean value per dispatch of every counter for every kernel.
* `kernel_roofline(...)`: the kernel's position against each roofline of MI355X -- VALU
  instruction issue, transcendental (sqrt) rate, HBM bandwidth -- and the binding one.

MI355X figures: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz = 3.93e13 VALU lane-instructions/s
(non-packed); transcendental instructions (v_sqrt_f32) issue at a quarter of that, 9.83e12/s;
HBM3E 8 TB/s.
"""
from __future__ import annotations

import argparse
import csv
import ctypes
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile

VALU_LANE_OPS_PER_S = 256 * 4 * 16 * 2.4e9
TRANS_LANE_OPS_PER_S = VALU_LANE_OPS_PER_S / 4
HBM_BYTES_PER_S = 8.0e12

# One rocprofv3 run each: counters of different blocks within the per-run hardware limits
# (at most 8 SQ, 4 TCC, 2 GRBM counters; FETCH_SIZE takes 3 TCC, WRITE_SIZE 2).
PASSES = [
    ["SQ_INSTS_VALU", "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_WAVES",
     "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"],
    ["VALUBusy"],
    ["FETCH_SIZE"],
    ["WRITE_SIZE"],
    ["SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"],
]

# FETCH_SIZE / WRITE_SIZE (KiB per dispatch) read for a copy kernel of 1 GiB read + 1 GiB
# written: 524,298 and 1,048,576, so bytes = FETCH_SIZE x 1024 x 2.000 + WRITE_SIZE x 1024 x 1.000
# (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over `python -m backprojector_amd.roofline copy
# --mib 1024 --iters 2`).
FETCH_FACTOR = 1048576.0 / 524298.0
WRITE_FACTOR = 1.0


def copy_benchmark(nbytes: int, iterations: int = 5) -> dict:
    from ._native import check, lib
    ms = ctypes.c_double()
    check(lib().bpx_copy_benchmark(int(nbytes), int(iterations), ctypes.byref(ms)))
    moved = 2 * (nbytes // 16 * 16)          # read + write
    return {"bytes_copied": nbytes // 16 * 16, "mean_ms": ms.value,
            "GBps_read_plus_write": moved / (ms.value * 1e-3) / 1e9}


def rocprofv3():
    return shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                         else None)


def base_name(kernel: str) -> str:
    """`void bpx::confocal_lds_kernel<bpx::FusedF32, 8, 8>(bpx::ConfocalArgs)` -> `confocal_lds_kernel`."""
    k = kernel.split("(")[0].split("<")[0].strip()
    k = k.split()[-1] if k.split() else k
    return k.split("::")[-1]


def parse_counter_csv(directory: str) -> dict:
    """{kernel name: [{counter: value, "duration_ns": ...} per dispatch, in launch order]} from a
    rocprofv3 CSV output directory (a counter's rows of one dispatch are summed)."""
    rows = {}
    for path in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                try:
                    v = float(row.get("Counter_Value", "nan"))
                    d = int(row.get("Dispatch_Id") or row.get("Correlation_Id") or 0)
                except ValueError:
                    continue
                rec = rows.setdefault(d, {"kernel": row.get("Kernel_Name", ""), "counters": {}})
                rec["counters"][row.get("Counter_Name", "")] = rec["counters"].get(row.get("Counter_Name", ""), 0.0) + v
                try:
                    rec["duration_ns"] = float(row["End_Timestamp"]) - float(row["Start_Timestamp"])
                except (KeyError, TypeError, ValueError):
                    pass
    out = {}
    for d in sorted(rows):
        rec = rows[d]
        entry = dict(rec["counters"])
        if "duration_ns" in rec:
            entry["duration_ns"] = rec["duration_ns"]
        out.setdefault(rec["kernel"], []).append(entry)
    return out


def collect(cmd, passes=PASSES, timeout_s=120, log=None, env=None):
    """Runs `cmd` (a list, the program and its arguments) once per counter pass under rocprofv3 and
    merges the passes dispatch by dispatch (the k-th launch of a kernel in one run is the k-th in
    the others: the program is deterministic).  Returns ({kernel: [per-dispatch counters]},
    [pass errors])."""
    prof = rocprofv3()
    if prof is None:
        return {}, ["rocprofv3 not found"]
    merged, errors = {}, []
    for counters in passes:
        with tempfile.TemporaryDirectory(prefix="bpx_pmc_") as out:
            full = ["timeout", "-s", "KILL", str(int(timeout_s)), prof, "--pmc", *counters, "--output-format", "csv",
                    "-d", out, "-o", "run", "--", *cmd]
            if log:
                log(f"counters: {' '.join(counters)}")
            r = subprocess.run(full, capture_output=True, text=True, cwd=tempfile.gettempdir(), env=env)
            got = parse_counter_csv(out)
            if not got:
                errors.append(f"{' '.join(counters)}: rc {r.returncode}: {(r.stderr or r.stdout)[-400:]}")
            for k, ds in got.items():
                dst = merged.setdefault(k, [])
                for i, d in enumerate(ds):
                    if i == len(dst):
                        dst.append({})
                    dst[i].update(d)
    return merged, errors


def mean_dispatch(dispatches, skip_first=True):
    """Mean of each counter over a kernel's dispatches (the first one left out when there are
    several: it may include cold caches)."""
    ds = dispatches[1:] if skip_first and len(dispatches) > 1 else dispatches
    keys = set().union(*[d.keys() for d in ds]) if ds else set()
    return {k: sum(d[k] for d in ds if k in d) / max(1, sum(1 for d in ds if k in d)) for k in keys}


def find_kernel(counts: dict, name: str):
    """(full name, mean counters per dispatch) of the kernel whose base name is `name` (the one
    with the most VALU work if several instantiations match)."""
    hits = {k: mean_dispatch(v) for k, v in counts.items() if base_name(k) == name}
    if not hits:
        return None, None
    k = max(hits, key=lambda k: hits[k].get("SQ_INSTS_VALU", 0.0))
    return k, hits[k]


def signature(kernel: str) -> str:
    """`void bpx::confocal_lds_kernel<bpx::FusedF32, 8, 8>(bpx::ConfocalArgs)` ->
    `confocal_lds_kernel<FusedF32, 8, 8>`."""
    k = kernel.replace("void ", "").replace("bpx::", "")
    depth, end = 0, len(k)
    for i, ch in enumerate(k):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            end = i
            break
    return k[:end].strip()


def largest_dispatch(dispatches):
    """The dispatch with the most VALU instructions (for kernels launched at several sizes)."""
    return dict(max(dispatches, key=lambda d: d.get("SQ_INSTS_VALU", 0.0))) if dispatches else {}


def summarize(counts: dict, min_ns: float = 0.0) -> dict:
    """{kernel: {"mean": ..., "largest": ..., "dispatches": n}} of a collect() result (kernels
    whose largest dispatch lasts min_ns or more): the compact form a committed summary keeps."""
    out = {}
    for k, ds in counts.items():
        big = largest_dispatch(ds)
        if big.get("duration_ns", 0.0) >= min_ns:
            out[k] = {"mean": mean_dispatch(ds), "largest": big, "dispatches": len(ds)}
    return out


def find_signature(counts: dict, prefix: str, pick: str = "mean"):
    """(full name, counters per dispatch) of the kernel whose signature starts with `prefix`:
    the mean over its dispatches, or its largest dispatch (pick="max").  `counts` is a collect()
    result or its summarize() form."""
    def agg(v):
        if isinstance(v, dict):
            return dict(v["largest" if pick == "max" else "mean"])
        return largest_dispatch(v) if pick == "max" else mean_dispatch(v)
    hits = {k: agg(v) for k, v in counts.items() if signature(k).startswith(prefix)}
    if not hits:
        return None, None
    k = max(hits, key=lambda k: hits[k].get("SQ_INSTS_VALU", 0.0))
    return k, hits[k]


def kernel_roofline(c: dict, kernel_ms: float, algorithmic_bytes: float = None):
    """Position of one kernel (mean counters per dispatch `c`, mean duration `kernel_ms`) against
    the VALU, transcendental and HBM rooflines; `bound` is the one with the largest fraction."""
    t = kernel_ms * 1e-3
    out = {"kernel_ms": kernel_ms}
    frac = {}
    if "SQ_INSTS_VALU" in c:
        a = c["SQ_INSTS_VALU"] * 64 / t
        out["valu"] = {"achieved": a, "peak": VALU_LANE_OPS_PER_S, "unit": "VALU lane-instructions/s",
                       "frac": a / VALU_LANE_OPS_PER_S, "instructions_per_dispatch": c["SQ_INSTS_VALU"]}
        frac["valu"] = a / VALU_LANE_OPS_PER_S
    if "VALUBusy" in c:
        out["valu_busy"] = c["VALUBusy"] / 100.0
    trans = c.get("SQ_INSTS_VALU_TRANS_F32")
    if trans is not None:
        a = trans * 64 / t
        out["sqrt"] = {"achieved": a, "peak": TRANS_LANE_OPS_PER_S, "unit": "transcendental lane-ops/s",
                       "frac": a / TRANS_LANE_OPS_PER_S}
        frac["sqrt"] = a / TRANS_LANE_OPS_PER_S
    if "FETCH_SIZE" in c and "WRITE_SIZE" in c:
        traffic = c["FETCH_SIZE"] * 1024 * FETCH_FACTOR + c["WRITE_SIZE"] * 1024 * WRITE_FACTOR
        out["hbm"] = {"achieved": traffic / t, "peak": HBM_BYTES_PER_S, "unit": "B/s",
                      "frac": traffic / t / HBM_BYTES_PER_S, "traffic_bytes": traffic}
        frac["hbm"] = traffic / t / HBM_BYTES_PER_S
    if "SQ_LDS_BANK_CONFLICT" in c and c.get("SQ_LDS_IDX_ACTIVE"):
        out["lds_bank_conflict_frac"] = c["SQ_LDS_BANK_CONFLICT"] / c["SQ_LDS_IDX_ACTIVE"]
    if algorithmic_bytes:
        out["effective_gather_Bps"] = algorithmic_bytes / t
    if frac:
        out["bound"] = max(frac, key=frac.get)
        out["frac"] = frac[out["bound"]]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["copy"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=5)
    a = ap.parse_args(argv)
    print(json.dumps(copy_benchmark(a.mib << 20, a.iters)), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])

#!/usr/bin/env python3
"""Headline benchmark: voxel-pair updates/s of the single-tap GPU backprojection.

Workload: confocal 128 x 128 wall grid over [-1, 1]^2 m (16,384 laser/SPAD pairs), laser and
SPAD origin (0, -1, 0.5) m, T = 4096 bins of 1 mm, t0 = 0, hidden scene of 8 points and two
0.3 m patches (seed 1000), 128^3 voxels over [-0.5, 0.5]^2 x [0.3, 1.3] m, fp32 samples.
The capture goes through the reference pipeline (time limits -> transient chunk -> pair
table -> basis) into a device-resident plan; one step = one full reconstruction on the GPU
(row staging + threshold table + gather kernel) with the capture already in HBM.

Prints one JSON line (see the keys below).  Only the cpu_baseline leg and the correctness
check use the CPU oracle, as the checker.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel-pair updates/s 128^2->128^3; learned-octree speedup at <=2% L2 err"
HBM_PEAK_GBS = 8000.0          # MI355X published HBM3E bandwidth
P5000_UPDATES_PER_S = 16384 * 128 ** 3 / 5.993   # README.md:106, 5993 ms on a Quadro P5000


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--T", type=int, default=4096)
    ap.add_argument("--side", type=int, default=128)
    ap.add_argument("--voxels", type=int, default=128)
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU time of the bounded cpu_baseline sample (0 = skip)")
    ap.add_argument("--check-voxels", type=int, default=1024)
    ap.add_argument("--octree", type=int, default=256,
                    help="fine volume side of the octree-vs-plain comparison (0 = skip)")
    ap.add_argument("--octree-tau", type=float, default=0.25)
    ap.add_argument("--octree-proposal", type=float, default=None,
                    help="proposal threshold of the learned octree (default: the shipped choice)")
    ap.add_argument("--configs", type=int, default=1,
                    help="also time the exhaustive 65,536-pair, 16-bit and complex workloads (0 = skip)")
    ap.add_argument("--one-shot", type=int, default=3,
                    help="timed one-shot nlosbpy.backproject calls (0 = skip)")
    ap.add_argument("--batch", type=int, default=64,
                    help="uneven captures of the batched comparison (0 = skip)")
    return ap.parse_args()


def octree_comparison(bp, cap, side, tau, proposal=None, reps=3):
    """Plain dense reconstruction vs octree refinement at side^3 on the same capture: the
    sound bound rule alone, and with the shipped learned proposal net attached."""
    from backprojector_amd import proposals
    occ_frac = proposals.OCCUPIED_FRACTION
    N = (side,) * 3
    args = proposals.plan_args(cap, side)
    with bp.Plan(*args) as plain:
        tp = plain.benchmark(1, reps)
        vp = plain.volume()
        plain_kernel = plain.info()["kernel"]
    # the texture-path gather (the machinery the octree's exact voxels use), bit-identical
    with bp.Plan(*args, variant=5) as plain_tex:
        tt = plain_tex.benchmark(1, reps)
        if not np.array_equal(plain_tex.volume(), vp):
            raise RuntimeError("plain kernels disagree")
    occ = vp >= occ_frac * vp.max()
    upd_plain = len(args[1]) * side ** 3
    net, meta = proposals.load()
    if proposal is None:
        proposal = meta.get("proposal_threshold") or 0.2
    rows = {}
    with bp.Plan.octree(*args, levels=4, threshold=tau) as oct_:
        for name, p in (("bound_rule", None), ("learned", proposal)):
            if p is not None:
                proposals.apply_operating_point(oct_, net, tau, p)
            t = oct_.benchmark(1, reps)
            vo = oct_.volume()
            st = oct_.octree_stats()
            err = float(np.linalg.norm((vo - vp)[occ]) / max(np.linalg.norm(vp[occ]), 1e-300))
            rows[name] = {"octree_ms": t["per_exec_ms"], "speedup": tp["per_exec_ms"] / t["per_exec_ms"],
                          "speedup_vs_texture_gather": tt["per_exec_ms"] / t["per_exec_ms"],
                          "rel_l2_occupied": err, "fine_voxels_computed": st["active"][-1],
                          "active_per_level": st["active"],
                          "work_fraction": (st["fine_updates"] + st["bound_updates"]) / upd_plain,
                          "fp64_running_max": st["running_max"]}
            if p is not None:
                rows[name]["proposal_threshold"] = p
    return {"side": side, "levels": 4, "tau": tau, "plain_ms": tp["per_exec_ms"],
            "plain_kernel": plain_kernel,
            "plain_texture_gather_ms": tt["per_exec_ms"],
            "occupied_voxels": int(occ.sum()),
            "occupied_definition": f"plain value >= {occ_frac} x max",
            "net": {"trained_scenes": meta.get("scenes"), "description": meta.get("description"),
                    "test": meta.get("test")},
            **rows}


def batch_comparison(bp, count, seed=1004, serial=True):
    """Uneven confocal captures: one persistent work-queue launch (autotuned, hipGraph replay)
    against one plan execution per capture."""
    from backprojector_amd import proposals, scenes
    t_r = time.perf_counter()
    caps = scenes.batch_captures(count, seed)
    args = [proposals.plan_args(c, N) for c, N in caps]
    log(f"batch: rendered and prepared {count} captures in {time.perf_counter() - t_r:.1f}s")
    with bp.Batch(args) as b:
        cfg = b.tune(reps=2)
        t = b.benchmark(1, 5)
        b.execute()
        lat = b.latencies()
    out = {"captures": count, "seeds": f"SeedSequence({seed}).spawn({count})",
           "updates": t["updates"], "batch_ms": t["per_exec_ms"], "updates_per_s": t["updates_per_s"],
           "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
           "config": cfg}
    for st in ("bf16",):   # 16-bit rows (per-capture scale): half the row memory, same queue
        with bp.Batch(args, storage=st) as b:
            c16 = b.tune(reps=2)
            t16 = b.benchmark(1, 5)
        out[st] = {"batch_ms": t16["per_exec_ms"], "updates_per_s": t16["updates_per_s"], "config": c16}
    if serial:
        ms = 0.0
        for a in args:
            with bp.Plan(*a) as pl:
                ms += pl.benchmark(1, 2)["per_exec_ms"]
        out["one_launch_per_capture_ms"] = ms
        out["speedup_vs_one_launch_per_capture"] = ms / t["per_exec_ms"]
    return out


def _plan_line(pl, N, sample_bytes, workload, reps):
    t = pl.benchmark(1, reps)
    info = pl.info()
    upd = info["pairs_used"] * N[0] * N[1] * N[2]
    g = t["gather_kernel_ms"]
    return {"workload": workload, "kernel_path": info["path"], "kernel": info["kernel"],
            "lane_voxels": info["lane_voxels"],
            "updates": upd, "exec_ms": t["per_exec_ms"], "gather_ms": g, "staging_ms": t["table_kernel_ms"],
            "updates_per_s": upd / (t["per_exec_ms"] * 1e-3), "gather_updates_per_s": upd / (g * 1e-3),
            "bytes_per_update": sample_bytes,
            "gather_frac_hbm": upd * sample_bytes / (g * 1e-3) / 1e9 / HBM_PEAK_GBS}


def config_measurements(bp, cap, reps=5):
    """The other workloads, one plan each (device-resident, timed with HIP events): the
    exhaustive 16x16 x 16x16-pair capture into 64^3 with exact bins (grid kernel, and the same
    pairs shuffled for the arbitrary-pair kernel); the headline capture at 128^3 with 16-bit
    storage, in the CPU path's semantics, and phasor-filtered (complex) samples."""
    from backprojector_amd import scenes
    out = {}
    N2 = (64, 64, 64)
    c2 = scenes.render(scenes.config_scene(2), 16, 4096, confocal=False, laser_side=16, seed=1002)
    with bp.Plan.from_nlos(*c2.nlos_args(N2), wall_normal=(0, 0, 0), assume_row_major=True) as pl:
        out["exhaustive_64cube"] = _plan_line(pl, N2, 4, "exhaustive 16x16 lasers x 16x16 SPADs -> 64^3, T=4096, f32", reps)
    # the same pairs in a shuffled order: the list loses its laser x SPAD structure, so the
    # arbitrary-pair kernel (both legs exact per update) runs instead of the grid kernel
    pre2 = bp.nlosbpy.prepare(*c2.nlos_args(N2), (0, 0, 0), assume_row_major=True)
    perm = np.random.default_rng(7).permutation(len(pre2["pairs"]))
    with bp.Plan(pre2["chunk"][perm], pre2["pairs"][perm], N2, pre2["zero"], pre2["inc"], pre2["chunked_t0"],
                 c2.deltaT) as pl:
        out["exhaustive_64cube_shuffled_pairs"] = _plan_line(
            pl, N2, 4, "exhaustive 16x16 x 16x16 pairs in shuffled order (no grid structure) -> 64^3, f32", reps)
    del c2, pre2
    N = (128, 128, 128)
    for st, b in (("f16", 2), ("bf16", 2)):
        with bp.Plan.from_nlos(*cap.nlos_args(N), wall_normal=(0, 0, 0), assume_row_major=True,
                               storage=st) as pl:
            out[f"confocal_128cube_{st}"] = _plan_line(pl, N, b, f"confocal 128x128 -> 128^3, T=4096, {st} storage", reps)
    pre = bp.nlosbpy.prepare(*cap.nlos_args(N), (0, 0, 0), assume_row_major=True)
    # the reference CPU path's semantics (voxel-diagonal window, double distances) on the GPU:
    # one-shot call from host arrays, so the copy of the chunk is included
    wargs = (pre["chunk"], pre["pairs"], N, cap.volume_position, cap.volume_size, pre["chunked_t0"],
             cap.deltaT)
    bp.window_backprojection(*wargs)
    t1 = time.perf_counter()
    bp.window_backprojection(*wargs)
    wms = (time.perf_counter() - t1) * 1e3
    upd = len(pre["pairs"]) * N[0] * N[1] * N[2]
    out["confocal_128cube_cpu_path_semantics"] = {"workload": "confocal 128x128 -> 128^3, T=4096, window mode (CPU path), "
                                             "one-shot from host arrays", "exec_ms": wms,
                                 "updates": upd, "updates_per_s": upd / (wms * 1e-3)}
    cplx = bp.phasor_pulse(pre["chunk"], 0.004, cap.deltaT)
    for st, b in (("f32", 8), ("f16", 4), ("bf16", 4)):
        with bp.Plan(cplx, pre["pairs"], N, pre["zero"], pre["inc"], pre["chunked_t0"], cap.deltaT,
                     storage=st) as pl:
            name = {"f32": "complex64", "f16": "complex f16", "bf16": "complex bf16"}[st]
            out[f"phasor_128cube_{name.replace(' ', '_')}"] = _plan_line(
                pl, N, b, f"confocal 128x128 -> 128^3, T=4096, phasor-filtered (4 mm), {name} storage", reps)
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Host processor model and logical CPU count (from /proc/cpuinfo where present)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (this library runs "
              "on a single GPU; sharding across GPUs is a design note, not code)", file=sys.stderr)
        sys.exit(2)
    import backprojector_amd as bp
    from backprojector_amd import scenes

    t_r = time.perf_counter()
    cap = scenes.metric_capture(T=args.T, side=args.side)
    log(f"rendered capture {cap.transient.shape} in {time.perf_counter() - t_r:.1f}s")
    N = (args.voxels,) * 3
    plan = bp.Plan.from_nlos(*cap.nlos_args(N), wall_normal=(0, 0, 0), assume_row_major=True,
                             storage=args.storage, variant=args.variant)
    info = plan.info()
    P = info["pairs_used"]
    updates = P * N[0] * N[1] * N[2]
    log(f"plan: {info}")

    for _ in range(args.warmup):
        plan.execute()
    plan.synchronize()
    t0 = time.perf_counter()
    timing = plan.benchmark(0, args.steps)      # K executions, HIP events on the plan's stream
    plan.synchronize()
    elapsed = time.perf_counter() - t0
    value = updates * args.steps / elapsed
    gather_ms = timing["gather_kernel_ms"]
    sample_bytes = {"f32": 4, "fp32": 4, "f16": 2, "fp16": 2, "bf16": 2}[args.storage]
    achieved = updates * sample_bytes / (gather_ms * 1e-3) / 1e9
    log(f"steps {args.steps}: {elapsed * 1e3 / args.steps:.2f} ms/step, gather kernel "
        f"{gather_ms:.2f} ms, table kernel {timing['table_kernel_ms']:.2f} ms")
    vol = plan.volume()

    # ---- correctness on sampled voxels, checked by the oracle (its own pipeline restatement)
    from oracle import oracle as orc
    mn, mx = orc.time_limits(cap.laser_position, cap.laser_grid, cap.camera_position,
                             cap.camera_grid, cap.volume_size, cap.volume_position, cap.t0,
                             cap.deltaT)
    pairs = orc.precompute_pairs(cap.laser_position, cap.laser_grid, cap.camera_position,
                                 cap.camera_grid, True)
    zero, inc = orc.basis(cap.volume_position, cap.volume_size, N)
    chunk = orc.transient_chunk(cap.transient, mn, mx)
    ct0 = float(orc.chunked_t0(cap.t0, mn, cap.deltaT))
    H = chunk.reshape(len(pairs), -1)
    rng = np.random.default_rng(0)
    vox = np.stack([rng.integers(0, n, args.check_voxels) for n in N], 1).astype(np.uint32)
    got = vol[vox[:, 0], vox[:, 1], vox[:, 2]].astype(np.float64)
    exact = orc.a1_voxels_f64acc(H, pairs, vox, zero, inc, ct0, cap.deltaT)
    fp64 = orc.fp64_voxels(H, pairs, vox, zero, inc, ct0, cap.deltaT)
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    check = {"voxels": int(args.check_voxels), "rel_l2_vs_reference_fp32_formula": rel(got, exact),
             "rel_l2_vs_fp64_oracle": rel(got, fp64)}
    log(f"check: {check}")

    # ---- one-shot calls (host arrays in, host volume out; PCIe and plan set-up included)
    one_shot = None
    if args.one_shot > 0:
        nargs = cap.nlos_args(N)
        bp.nlosbpy.backproject(*nargs, (0, 0, 0), assume_row_major=True)
        t1 = time.perf_counter()
        for _ in range(args.one_shot):
            bp.nlosbpy.backproject(*nargs, (0, 0, 0), assume_row_major=True)
        dev_ms = (time.perf_counter() - t1) * 1e3 / args.one_shot
        t1 = time.perf_counter()
        for _ in range(args.one_shot):
            pre = bp.nlosbpy.prepare(*nargs, (0, 0, 0), assume_row_major=True)
            bp.backprojection(pre["chunk"], pre["pairs"], N, pre["zero"], pre["inc"], pre["chunked_t0"],
                              cap.deltaT)
        host_ms = (time.perf_counter() - t1) * 1e3 / args.one_shot
        one_shot = {"calls": args.one_shot, "nlosbpy_backproject_ms": dev_ms,
                    "updates_per_s_pcie_inclusive": updates / (dev_ms * 1e-3),
                    "host_chunk_then_backprojection_ms": host_ms,
                    "source_bytes": int(cap.transient.nbytes)}
        log(f"one-shot: {one_shot}")

    # ---- CPU baseline: the reference CPU path (oracle restatement, OpenMP), bounded sample
    cpu = None
    if args.cpu_seconds > 0:
        slab = N[1] * N[2]
        nth = orc.num_threads()
        t1 = time.perf_counter()
        orc.a2_volume(H, pairs, N, cap.volume_position, cap.volume_size, ct0, cap.deltaT,
                      vox_begin=0, vox_end=slab)
        one = time.perf_counter() - t1
        slabs = int(max(1, min(N[0], args.cpu_seconds / max(one, 1e-3))))
        t1 = time.perf_counter()
        orc.a2_volume(H, pairs, N, cap.volume_position, cap.volume_size, ct0, cap.deltaT,
                      vox_begin=0, vox_end=slab * slabs)
        dt = time.perf_counter() - t1
        cpu = {"value": slab * slabs * P / dt, "unit": "voxel-pair updates/s", "cores": nth,
               "kind": "port",
               "sample": f"reference CPU path (voxel-diagonal window, double distances) over "
                         f"{slabs} of {N[0]} x-slabs ({slab * slabs} voxels x {P} pairs) in "
                         f"{dt:.1f} s on {nth} threads",
               "host": host_cpu()}
        # like-for-like: the single-tap formula of the GPU path, restated on the CPU (one bin per
        # update, fp32 taps, double accumulation), over a bounded voxel sample
        vox_all = np.stack(np.meshgrid(*[np.arange(n) for n in N], indexing="ij"), -1).reshape(-1, 3)
        n1 = min(len(vox_all), slab)
        t1 = time.perf_counter()
        orc.a1_voxels_f64acc(H, pairs, vox_all[:n1], zero, inc, ct0, cap.deltaT)
        one = time.perf_counter() - t1
        n2 = int(max(n1, min(len(vox_all), n1 * (0.5 * args.cpu_seconds) / max(one, 1e-3))))
        t1 = time.perf_counter()
        orc.a1_voxels_f64acc(H, pairs, vox_all[:n2], zero, inc, ct0, cap.deltaT)
        dt = time.perf_counter() - t1
        cpu["single_tap"] = {"value": n2 * P / dt, "unit": "voxel-pair updates/s", "cores": nth,
                             "sample": f"single-tap formula of the GPU path (fp32 bins, double sums) "
                                       f"over the first {n2} voxels x {P} pairs in {dt:.1f} s on "
                                       f"{nth} threads"}
        log(f"cpu baseline: {cpu}")

    configs = None
    if args.configs:
        configs = config_measurements(bp, cap)
        log("configs: " + "; ".join(f"{k} {v.get('gather_ms', v['exec_ms']):.2f} ms "
                                    f"{v.get('gather_updates_per_s', v['updates_per_s']):.3g}/s"
                                    for k, v in configs.items()))

    plan.close()
    octree = None
    if args.octree > 0:
        octree = octree_comparison(bp, cap, args.octree, args.octree_tau, args.octree_proposal)
        log(f"octree: {octree}")
    batch = None
    if args.batch > 0:
        batch = batch_comparison(bp, args.batch)
        log(f"batch: {batch}")

    # ---- HBM traffic per launch of the gather kernel, from the committed rocprofv3 --pmc
    # summary (FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel calibration)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    key = f"{info['kernel']}|{args.storage}|{args.side}|{N[0]}|{args.T}"
    if os.path.exists(tfile):
        tab = json.load(open(tfile))
        if key in tab:
            traffic = tab[key]["hbm_bytes_per_launch"]

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-pair updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / P5000_UPDATES_PER_S,
        "dtype": args.storage if args.storage != "fp32" else "f32",
        "data": "synthetic",
        "config": {"workload": f"confocal {args.side}x{args.side} pairs -> {N[0]}^3 voxels, "
                               f"T={args.T} bins of 1 mm (chunk {info['Tc']}), single-tap GPU path",
                   "pairs": P, "voxels": N[0] * N[1] * N[2], "chunk_bins": info["Tc"],
                   "storage": args.storage, "kernel_path": info["path"],
                   "lane_voxels": info["lane_voxels"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_key": key,
                     "kernel": info["kernel"], "kernel_ms": gather_ms,
                     "note": "achieved = 4 B per update (algorithmic) / kernel time; the LDS-staged "
                             "kernel fetches each staged sample once per tile and reads it from LDS for "
                             "every voxel of the tile, so this rate can exceed the HBM peak; its real "
                             "HBM bytes per launch are `traffic`",
                     "table_kernel_ms": timing["table_kernel_ms"],
                     "bytes_per_update": sample_bytes},
        "cpu_baseline": cpu,
        "check": check,
        "one_shot": one_shot,
        "configs": configs,
        "octree": octree,
        "batch": batch,
        "baseline_ref": "Quadro P5000 5993 ms (README.md:106) -> 5.73e9 updates/s",
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

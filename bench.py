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
    ap.add_argument("--tune", type=int, default=1,
                    help="with --variant -1: time every gather variant on the capture first (bpx_plan_tune)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU time of the bounded cpu_baseline sample (0 = skip)")
    ap.add_argument("--check-voxels", type=int, default=1024)
    ap.add_argument("--octree", type=int, default=256,
                    help="fine volume side of the octree-vs-plain comparison (0 = skip)")
    ap.add_argument("--octree-tau", type=float, default=0.25)
    ap.add_argument("--octree-proposal", type=float, default=None,
                    help="proposal threshold of the learned octree (default: the shipped operating point)")
    ap.add_argument("--octree-weights", default=None,
                    help="proposal-net weights and operating point (path without .npy/.json; default: shipped)")
    ap.add_argument("--configs", type=int, default=1,
                    help="also time the exhaustive 65,536-pair, 16-bit and complex workloads (0 = skip)")
    ap.add_argument("--one-shot", type=int, default=3,
                    help="timed one-shot nlosbpy.backproject calls (0 = skip)")
    ap.add_argument("--batch", type=int, default=64,
                    help="uneven captures of the batched comparison (0 = skip)")
    ap.add_argument("--batch-serial", type=int, default=1,
                    help="also time one plan execution per capture of the batch (0 = skip)")
    ap.add_argument("--counters", choices=["headline", "all", "off"], default="headline",
                    help="rocprofv3 --pmc passes run by this bench before its timed run: over the headline "
                         "gather only (default), over every workload of this run, or none")
    ap.add_argument("--counters-out", default=None,
                    help="write the per-kernel counters and rooflines of the passes to this JSON file")
    ap.add_argument("--counter-timeout", type=int, default=150, help="seconds per counter pass")
    return ap.parse_args()


def octree_comparison(bp, cap, side, tau, proposal=None, reps=3, ctx=None, weights=None):
    """Plain dense reconstruction vs octree refinement at side^3 on the same capture: the
    sound bound rule alone, and with the shipped learned proposal net attached."""
    from backprojector_amd import proposals
    occ_frac = proposals.OCCUPIED_FRACTION
    N = (side,) * 3
    args = proposals.plan_args(cap, side)
    with bp.Plan(*args) as plain:
        plain.tune(reps=2)   # the fastest plain kernel for this capture (same volume)
        tp = plain.benchmark(1, reps)
        vp = plain.volume()
        plain_kernel = plain.info()["kernel"]
    # the texture-path gather (the machinery the octree's exact voxels use), bit-identical
    with bp.Plan(*args, variant=5) as plain_tex:
        tt = plain_tex.benchmark(1, reps)
        if not np.array_equal(plain_tex.volume(), vp):
            raise RuntimeError("plain kernels disagree")
    occ = vp >= occ_frac * vp.max()
    support = proposals.surface_support(cap, side) if cap.hidden is not None else None
    upd_plain = len(args[1]) * side ** 3
    net, meta = proposals.load(weights) if weights else proposals.load()
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
                          "fallback_voxels": st["fallback_voxels"], "coarse_pairs": st["coarse_pairs"],
                          "rel_l2_surface_support": proposals.rel_l2_on(vo, vp, support) if support is not None else None,
                          "fp64_running_max": st["running_max"]}
            if p is not None:
                rows[name]["proposal_threshold"] = p
    return {"side": side, "levels": 4, "tau": tau, "plain_ms": tp["per_exec_ms"],
            "plain_kernel": plain_kernel,
            "plain_texture_gather_ms": tt["per_exec_ms"],
            "occupied_voxels": int(occ.sum()),
            "occupied_definition": f"plain value >= {occ_frac} x max",
            "surface_support_voxels": int(support.sum()) if support is not None else None,
            "surface_support_definition": "voxels within 1 voxel (per axis) of a hidden surface sample; "
                                          "independent of the reconstruction and of tau",
            "net": {"trained_scenes": meta.get("scenes"), "description": meta.get("description"),
                    "operating_point": meta.get("proposal_threshold"),
                    "held_out_test": meta.get("test")},
            **rows, "kernel_rooflines": octree_rooflines(ctx)}


def octree_rooflines(ctx):
    """Counters of the octree's sparse passes (their largest dispatch; durations under the
    counter passes)."""
    if ctx is None or not ctx.get("counters"):
        return None
    out = {}
    for name in ("oct_bound_kernel<1", "oct_bound_kernel<2", "oct_fine_kernel<1", "oct_fine_kernel<2",
                 "oct_tile_kernel", "winmax_fused_kernel", "conv3_fwd_kernel"):
        r = roofline_of(ctx["counters"], name, None, None, ctx["fingerprint"], pick="max")
        if r is not None:
            out[name] = r
    return out


def batch_comparison(bp, count, seed=1004, serial=True, ctx=None):
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
    if ctx is not None and ctx.get("counters"):
        # the persistent queue kernel of the chosen configuration, and the prologue
        k, d = cfg["lane_voxels"], cfg["lds_depth"]
        if cfg["launches"] == 1:
            queue = f"batch_queue_kernel<FusedF32, FusedSource<FusedF32>, {k}, {d}, 7, 4>"
        else:   # the LDS-staged tiles' queue (its voxels per lane are the LDS tile's)
            queue = f"batch_queue_kernel<FusedF32, FusedSource<FusedF32>, 2, {d}, 2, 4>"
        names = [queue, "batch_prologue_kernel"]
        if cfg["launches"] == 3:   # the 512-lane 16^3-tile queue
            names.append("batch_queue_kernel<FusedF32, FusedSource<FusedF32>, 2, 8, 2, 8>")
        if cfg["launches"] >= 2:
            names += [f"batch_tex_kernel<{k}, FusedSource<FusedF32>", "stage_rows_kernel"]
        roofs = {name: roofline_of(ctx["counters"], name, None, None, ctx["fingerprint"], pick="max")
                 for name in names}
        out["kernel_rooflines"] = {k: v for k, v in roofs.items() if v is not None}
    return out


def _plan_line(pl, N, sample_bytes, workload, reps, storage="f32", cplx=False, ctx=None):
    t = pl.benchmark(1, reps)
    info = pl.info()
    upd = info["pairs_used"] * N[0] * N[1] * N[2]
    g = t["gather_kernel_ms"]
    line = {"workload": workload, "kernel_path": info["path"], "kernel": info["kernel"],
            "lane_voxels": info["lane_voxels"],
            "updates": upd, "exec_ms": t["per_exec_ms"], "gather_ms": g, "staging_ms": t["table_kernel_ms"],
            "updates_per_s": upd / (t["per_exec_ms"] * 1e-3), "gather_updates_per_s": upd / (g * 1e-3),
            "bytes_per_update": sample_bytes,
            "effective_gather_GBps": upd * sample_bytes / (g * 1e-3) / 1e9}
    line["row_storage"] = info["row_storage"]
    if ctx is not None:
        prefix = device_kernel(info, storage if cplx else info["row_storage"], cplx)
        if prefix.startswith("confocal_lds_kernel<FusedC64Part<0>"):   # two passes, one per component
            prefix = [prefix, prefix.replace("Part<0>", "Part<1>")]
        line["roofline"] = roofline_of(ctx["counters"], prefix, g, upd * sample_bytes, ctx["fingerprint"])
    return line


def config_measurements(bp, cap, reps=5, ctx=None):
    """The other workloads, one plan each (device-resident, timed with HIP events): the
    exhaustive 16x16 x 16x16-pair capture into 64^3 with exact bins (grid kernel, and the same
    pairs shuffled for the arbitrary-pair kernel); the headline capture at 128^3 with 16-bit
    storage, in the CPU path's semantics, and phasor-filtered (complex) samples."""
    from backprojector_amd import scenes
    out = {}
    N2 = (64, 64, 64)
    c2 = scenes.render(scenes.config_scene(2), 16, 4096, confocal=False, laser_side=16, seed=1002)
    with bp.Plan.from_nlos(*c2.nlos_args(N2), wall_normal=(0, 0, 0), assume_row_major=True) as pl:
        out["exhaustive_64cube"] = _plan_line(pl, N2, 4, "exhaustive 16x16 lasers x 16x16 SPADs -> 64^3, T=4096, f32",
                                              reps, ctx=ctx)
    # the same pairs in a shuffled order: the list loses its laser x SPAD structure, so the
    # arbitrary-pair kernel (both legs exact per update) runs instead of the grid kernel
    pre2 = bp.nlosbpy.prepare(*c2.nlos_args(N2), (0, 0, 0), assume_row_major=True)
    perm = np.random.default_rng(7).permutation(len(pre2["pairs"]))
    with bp.Plan(pre2["chunk"][perm], pre2["pairs"][perm], N2, pre2["zero"], pre2["inc"], pre2["chunked_t0"],
                 c2.deltaT) as pl:
        out["exhaustive_64cube_shuffled_pairs"] = _plan_line(
            pl, N2, 4, "exhaustive 16x16 x 16x16 pairs in shuffled order (no grid structure) -> 64^3, f32", reps,
            ctx=ctx)
    del c2, pre2
    N = (128, 128, 128)
    for st, b in (("f16", 2), ("bf16", 2)):
        with bp.Plan.from_nlos(*cap.nlos_args(N), wall_normal=(0, 0, 0), assume_row_major=True,
                               storage=st) as pl:
            out[f"confocal_128cube_{st}"] = _plan_line(pl, N, b, f"confocal 128x128 -> 128^3, T=4096, {st} storage",
                                                       reps, storage=st, ctx=ctx)
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
    if ctx is not None:   # kernel time: its duration under the counter passes
        out["confocal_128cube_cpu_path_semantics"]["roofline"] = roofline_of(
            ctx["counters"], "window_kernel", None, None, ctx["fingerprint"])
    cplx = bp.phasor_pulse(pre["chunk"], 0.004, cap.deltaT)
    if ctx is not None:   # the complex FIR filter (one-shot call; kernel time under the counter passes)
        out["phasor_fir_4mm"] = {"workload": "phasor FIR (4 mm) of the 128x128 x Tc chunk, complex64 out",
                                 "roofline": roofline_of(ctx["counters"], "phasor_fir_kernel", None,
                                                         pre["chunk"].nbytes * 3, ctx["fingerprint"])}
    for st, b in (("f32", 8), ("f16", 4), ("bf16", 4)):
        with bp.Plan(cplx, pre["pairs"], N, pre["zero"], pre["inc"], pre["chunked_t0"], cap.deltaT,
                     storage=st) as pl:
            name = {"f32": "complex64", "f16": "complex f16", "bf16": "complex bf16"}[st]
            out[f"phasor_128cube_{name.replace(' ', '_')}"] = _plan_line(
                pl, N, b, f"confocal 128x128 -> 128^3, T=4096, phasor-filtered (4 mm), {name} storage", reps,
                storage=st, cplx=True, ctx=ctx)
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


COUNTERS_FILE = os.path.join(ROOT, "profiles", "counters.json")


def source_fingerprint():
    """sha256 of the kernel sources: committed counter summaries older than the kernels are
    reported as stale."""
    import glob
    import hashlib
    h = hashlib.sha256()
    for f in sorted(glob.glob(os.path.join(ROOT, "backprojector_amd", "csrc", "*"))):
        if f.endswith((".hip", ".hpp", ".cpp")):
            h.update(os.path.basename(f).encode())
            h.update(open(f, "rb").read())
    return h.hexdigest()[:16]


def counter_passes(args):
    """Runs this bench under rocprofv3 --pmc (one process per pass, a short run of the same
    workloads) and returns {kernel: [per-dispatch counters]}, or None."""
    from backprojector_amd import roofline as rl
    child = [sys.executable, os.path.abspath(__file__), "--counters", "off", "--steps", "2", "--warmup", "1",
             "--cpu-seconds", "0", "--check-voxels", "16", "--one-shot", "0", "--T", str(args.T), "--side",
             str(args.side), "--voxels", str(args.voxels), "--storage", args.storage, "--variant", str(args.variant)]
    if args.counters == "headline":
        child += ["--octree", "0", "--configs", "0", "--batch", "0"]
    else:
        child += ["--octree", str(args.octree), "--octree-tau", str(args.octree_tau), "--configs", str(args.configs),
                  "--batch", str(args.batch), "--batch-serial", "0"]
        if args.octree_proposal is not None:
            child += ["--octree-proposal", str(args.octree_proposal)]
        if args.octree_weights:
            child += ["--octree-weights", args.octree_weights]
    env = dict(os.environ)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    t = time.perf_counter()
    counts, errors = rl.collect(child, timeout_s=args.counter_timeout, log=log, env=env)
    log(f"counter passes: {len(rl.PASSES)} runs in {time.perf_counter() - t:.0f} s, {len(counts)} kernels"
        + (f", errors: {errors}" if errors else ""))
    return {"kernels": counts, "errors": errors, "passes": rl.PASSES,
            "command": "rocprofv3 --pmc <pass> --output-format csv -d <dir> -o run -- " + " ".join(
                ["python3", "bench.py"] + child[2:]),
            "fingerprint": source_fingerprint()} if counts else None


def committed_counters():
    try:
        return json.load(open(COUNTERS_FILE))
    except (OSError, ValueError):
        return None


def device_kernel(info, storage, cplx):
    """Signature prefix (roofline.signature form) of the gather kernel a plan launches."""
    import re
    k = info["kernel"]
    m = re.match(r"confocal_lds_kernel<KZ=(\d+), waves=(\d+)>", k)
    if m:
        fz = {("f32", False): "FusedF32", ("f16", False): "FusedH16<false>", ("bf16", False): "FusedH16<true>",
              ("f32", True): "FusedC64Part<0>", ("f16", True): "FusedC16<false>",
              ("bf16", True): "FusedC16<true>"}[(storage, cplx)]
        return f"confocal_lds_kernel<{fz}, {m.group(1)}, {m.group(2)}>"
    smp = {("f32", False): "SmpF32", ("f16", False): "SmpF16", ("bf16", False): "SmpBF16",
           ("f32", True): "SmpC64", ("f16", True): "SmpCF16", ("bf16", True): "SmpCBF16"}[(storage, cplx)]
    m = re.match(r"(grid_kernel|pairs_kernel)<K=(\d+)>", k)
    if m:
        return f"{m.group(1)}<{smp}, {m.group(2)}"
    m = re.match(r"confocal_kernel<K=(\d+)>", k)
    if m:
        return "confocal_kernel<FusedSource<"
    return k.split("<")[0]


def roofline_of(counters, prefix, kernel_ms, algorithmic_bytes, fingerprint, pick="mean"):
    """Roofline block of one kernel from a counter set (live or committed), or None.  prefix: a
    signature prefix, or a list of them whose counters are summed (one launch made of several
    kernels).  kernel_ms None: the kernels' duration under the counter passes."""
    from backprojector_amd import roofline as rl
    if not counters:
        return None
    prefixes = prefix if isinstance(prefix, list) else [prefix]
    names, c = [], {}
    for pf in prefixes:
        name, ci = rl.find_signature(counters["kernels"], pf, pick)
        if ci is None:
            return None
        names.append(name)
        for k, v in ci.items():
            c[k] = c.get(k, 0.0) + v
    if "VALUBusy" in c:
        c["VALUBusy"] /= len(prefixes)
    if kernel_ms is None:
        if "duration_ns" not in c:
            return None
        kernel_ms = c["duration_ns"] * 1e-6
    r = rl.kernel_roofline(c, kernel_ms, algorithmic_bytes)
    r["kernels"] = names
    r["counters_per_dispatch"] = {k: v for k, v in c.items() if k != "duration_ns"}
    if "duration_ns" in c:
        r["kernel_ms_under_counters"] = c["duration_ns"] * 1e-6
    r["counters_source"] = counters.get("source", "live rocprofv3 --pmc passes of this bench run")
    r["dispatch"] = "largest" if pick == "max" else "mean"
    r["stale"] = counters.get("fingerprint") != fingerprint
    return r


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (this library runs "
              "on a single GPU; sharding across GPUs is a design note, not code)", file=sys.stderr)
        sys.exit(2)
    # hardware counters first, in child processes, before this process touches the GPU
    fp = source_fingerprint()
    live = counter_passes(args) if args.counters != "off" else None
    committed = committed_counters()
    if committed is not None:
        committed["source"] = "profiles/counters.json (committed rocprofv3 --pmc passes)"
    if args.counters_out and live is not None:
        with open(args.counters_out, "w") as f:
            json.dump(live, f)
    # workloads this run profiled use the live counters, the others the committed ones
    ctx = {"counters": live if args.counters == "all" and live else committed, "fingerprint": fp}
    import backprojector_amd as bp
    from backprojector_amd import scenes

    t_r = time.perf_counter()
    cap = scenes.metric_capture(T=args.T, side=args.side)
    log(f"rendered capture {cap.transient.shape} in {time.perf_counter() - t_r:.1f}s")
    N = (args.voxels,) * 3
    plan = bp.Plan.from_nlos(*cap.nlos_args(N), wall_normal=(0, 0, 0), assume_row_major=True,
                             storage=args.storage, variant=args.variant)
    if args.variant < 0 and args.tune:   # plan set-up, before the timed region
        tuned = plan.tune(reps=3)
        log(f"autotuned gather variant: {tuned}")
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
    info = plan.info()   # after the executions: the row storage they staged
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
        configs = config_measurements(bp, cap, ctx=ctx)
        log("configs: " + "; ".join(f"{k} {v.get('gather_ms', v.get('exec_ms', 0)):.2f} ms "
                                    f"{v.get('gather_updates_per_s', v.get('updates_per_s', 0)):.3g}/s"
                                    for k, v in configs.items()))

    plan.close()
    octree = None
    if args.octree > 0:
        octree = octree_comparison(bp, cap, args.octree, args.octree_tau, args.octree_proposal, ctx=ctx,
                                   weights=args.octree_weights)
        log(f"octree: {octree}")
    batch = None
    if args.batch > 0:
        batch = batch_comparison(bp, args.batch, serial=bool(args.batch_serial), ctx=ctx)
        log(f"batch: {batch}")

    # ---- the gather kernel against its rooflines: counters from the rocprofv3 --pmc passes of this
    # run (else the committed ones), kernel time from this run's HIP events; the row storage the
    # executions staged (exact fp16 rows for integer samples) names the instantiation that ran
    prefix = device_kernel(info, info["row_storage"], False)
    rl_block = roofline_of(live if live else committed, prefix, gather_ms, updates * sample_bytes, fp)
    if rl_block is not None and "valu" in rl_block:
        roof = {"bound": rl_block["bound"],
                "achieved": rl_block[rl_block["bound"]]["achieved"], "peak": rl_block[rl_block["bound"]]["peak"],
                "unit": rl_block[rl_block["bound"]]["unit"], "frac": rl_block["frac"],
                "traffic": rl_block.get("hbm", {}).get("traffic_bytes"),
                "valu_frac": rl_block["valu"]["frac"], "valu_busy": rl_block.get("valu_busy"),
                "sqrt_frac": rl_block.get("sqrt", {}).get("frac"), "hbm_frac": rl_block.get("hbm", {}).get("frac"),
                "lds_bank_conflict_frac": rl_block.get("lds_bank_conflict_frac"),
                "valu_instructions_per_update": rl_block["valu"]["instructions_per_dispatch"] * 64 / updates,
                "effective_gather_GBps": achieved, "algorithmic_bytes_per_update": sample_bytes,
                "kernel": rl_block["kernels"][0], "kernel_ms": gather_ms,
                "table_kernel_ms": timing["table_kernel_ms"],
                "counters_source": rl_block["counters_source"], "stale": rl_block["stale"],
                "counters_per_dispatch": rl_block["counters_per_dispatch"],
                "note": "bound = the roofline with the largest fraction; VALU: SQ_INSTS_VALU x 64 lanes / kernel "
                        "time against 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz; sqrt: SQ_INSTS_VALU_TRANS_F32 x 64 "
                        "against a quarter of that; HBM: (FETCH_SIZE x 2.000 + WRITE_SIZE) KiB (copy-kernel "
                        "calibration) against 8 TB/s.  effective_gather_GBps = 4 B per update / kernel time "
                        "(each staged sample is read from LDS for every voxel of a tile, so it exceeds HBM)"}
    else:
        roof = {"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None,
                "effective_gather_GBps": achieved, "kernel_ms": gather_ms,
                "note": "no counters (rocprofv3 unavailable and no committed summary for this kernel)"}

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
                   "storage": args.storage, "row_storage": info["row_storage"], "kernel_path": info["path"],
                   "lane_voxels": info["lane_voxels"],
                   "kernel_choice": "autotuned (bpx_plan_tune)" if args.variant < 0 and args.tune
                   else f"variant {args.variant}"},
        "roofline": roof,
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

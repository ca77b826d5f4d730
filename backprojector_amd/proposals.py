"""Learned octree proposals: training on synthetic captures, validation, and the shipped weights.

No reference counterpart (the reference's octree, backproject_cuda.cu:183-367, prunes with a
fixed rule and does not run).  The net (libbpx, bpx_net_*) sees each coarse level of a
bounds-only octree execution -- per block: U / max U, V / running max, evaluated flag,
U / rms U -- and is trained with weighted binary cross-entropy to predict, for each of the 8
children, whether the child holds an "occupied" voxel of the dense reconstruction (value >=
OCCUPIED_FRACTION x max).  At inference a child is refined iff its parent passes the sound
bound rule (U >= tau x running max) AND its probability exceeds the proposal threshold.

Everything runs on the GPU through the library's own kernels: dense and octree plans,
features, labels (FP64 max, max-pool pyramid), forward, backward, Adam.

    python -m backprojector_amd.proposals train --scenes 2000 --out weights/octree_proposal
    python -m backprojector_amd.proposals validate --weights weights/octree_proposal
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import math
import os
import sys
import time

import numpy as np

from . import core, nlosbpy, scenes

OCCUPIED_FRACTION = 0.25
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "octree_proposal")
TRAIN_SEED_BASE = 50000         # training scenes 50000 ... 50000 + n - 1
VALIDATION_SEED_BASE = 55000    # captures that choose the proposal thresholds
TEST_SEED_BASE = 60000          # held-out test captures 60000 ...

# Proposal configurations tried on the validation captures: a probability for every level or
# one per level (coarse levels cautious: a veto there drops a large block), the coarse-level
# pair stride (features from every 2nd wall point per axis: a quarter of the pairs), and the
# band of low-confidence children computed at full resolution instead (level 2: children of
# 2 x 2 x 2 fine voxels; at the last coarse level a child is one voxel, so a band there equals
# a lower probability).
CANDIDATES = [0.2, 0.3,
              {"p": 0.2, "stride": 2}, {"p": 0.3, "stride": 2},
              {"p": 0.15, "stride": 2, "fallback": [0, 0, 0.08, 0]},
              {"p": 0.2, "stride": 2, "fallback": [0, 0, 0.1, 0]},
              {"p": 0.25, "stride": 2, "fallback": [0, 0, 0.1, 0]},
              {"p": 0.3, "stride": 2, "fallback": [0, 0, 0.1, 0]},
              {"p": 0.3, "stride": 2, "fallback": [0, 0, 0.15, 0]},
              {"p": [0.1, 0.15, 0.2, 0.3], "stride": 2, "fallback": [0, 0, 0.1, 0]},
              {"p": [0.1, 0.2, 0.3, 0.25], "stride": 2, "fallback": [0, 0.1, 0.15, 0]},
              # narrower bands just below the proposal threshold
              {"p": 0.2, "stride": 2, "fallback": [0, 0, 0.15, 0]},
              {"p": 0.2, "stride": 2, "fallback": [0, 0, 0.17, 0]},
              {"p": 0.22, "stride": 2, "fallback": [0, 0, 0.15, 0]},
              {"p": 0.22, "stride": 2, "fallback": [0, 0, 0.18, 0]},
              {"p": 0.25, "stride": 2, "fallback": [0, 0, 0.15, 0]},
              {"p": 0.25, "stride": 2, "fallback": [0, 0, 0.2, 0]}]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def training_scene(rng):
    """Random hidden geometry: 0-12 points, 0-3 patches (0.1-0.4 m), 0-2 spheres."""
    parts = []
    npts = int(rng.integers(0, 13))
    if npts:
        parts.append(scenes.points(rng, npts))
    for _ in range(int(rng.integers(0, 4))):
        parts.append(scenes.patch(rng, size=float(rng.uniform(0.1, 0.4)), spacing=0.004))
    for _ in range(int(rng.integers(0, 3))):
        parts.append(scenes.sphere(rng, spacing=0.006))
    if not parts:
        parts.append(scenes.points(rng, 1))
    return np.concatenate(parts)


def plan_args(cap, voxels):
    """(chunk, pairs, N, zero, inc, t0, deltaT) through the reference pipeline's preparation."""
    N = (voxels,) * 3
    prep = nlosbpy.prepare(*cap.nlos_args(N))
    return (prep["chunk"], prep["pairs"], N, prep["zero"], prep["inc"],
            float(prep["chunked_t0"]), cap.deltaT)


def training_capture(index, side, voxels, T=4096):
    seed = TRAIN_SEED_BASE + index
    rng = np.random.default_rng(seed)
    cap = scenes.render(training_scene(rng), side, T, seed=seed)
    return plan_args(cap, voxels)


def train(scenes_n=2000, steps_per_scene=3, side=64, voxels=256, levels=4, lr=2e-3,
          pos_weight=8.0, seed=0, log_every=50, init=None, tau=0.25, on_policy=0.5,
          policy_threshold=0.1, label_fraction=0.2, coarse_pair_stride=1):
    """Trains a net on `scenes_n` distinct synthetic scenes (Adam, cosine learning rate).

    Each scene's octree runs its coarse levels with the bound rule (threshold tau), so the net
    sees the same evaluated sets and features as at inference; for the last `on_policy`
    fraction of the scenes the net being trained is attached (proposal threshold
    policy_threshold) and the training data follow its own selections.  Children are labelled
    occupied at label_fraction x max, below the evaluation's OCCUPIED_FRACTION, so that voxels
    near the occupancy boundary lie inside the positive class (a safety margin: one missed
    occupied voxel costs far more error than one extra computed voxel costs time).
    coarse_pair_stride: the coarse levels' features from every s-th wall point per grid axis,
    as at inference with that stride.  Returns (net, history)."""
    net = core.Net(seed, params=init)
    hist = []
    total = scenes_n * steps_per_scene
    step = 0
    t_start = time.perf_counter()
    window = []
    with cf.ThreadPoolExecutor(1) as pool:     # renders the next capture while the GPU trains
        fut = pool.submit(training_capture, 0, side, voxels)
        for i in range(scenes_n):
            args = fut.result()
            if i + 1 < scenes_n:
                fut = pool.submit(training_capture, i + 1, side, voxels)
            attach = i >= scenes_n * (1.0 - on_policy)
            with core.Plan.octree(*args, levels=levels, threshold=tau, bounds_only=True,
                                  proposal_threshold=policy_threshold) as op, \
                    core.Plan(*args) as dp:
                if coarse_pair_stride > 1:
                    op.set_octree_options(tau, policy_threshold, bounds_only=True,
                                          coarse_pair_stride=coarse_pair_stride)
                if attach:
                    op.attach_net(net)
                try:
                    op.execute()
                    dp.execute()
                except Exception as e:
                    raise RuntimeError(f"scene {i}: chunk {args[0].shape}, {len(args[1])} pairs, "
                                       f"N {args[2]}, t0 {args[5]}: {e}") from e
                for _ in range(steps_per_scene):
                    lr_t = lr * (0.1 + 0.9 * 0.5 * (1 + math.cos(math.pi * step / total)))
                    loss = net.train_octree(op, dp, label_fraction, pos_weight, lr_t)
                    step += 1
                    window.append(loss)
            if (i + 1) % log_every == 0 or i + 1 == scenes_n:
                m = float(np.mean(window))
                hist.append({"scenes": i + 1, "loss": m})
                log(f"scene {i + 1}/{scenes_n}: mean loss {m:.5f} "
                    f"({time.perf_counter() - t_start:.0f} s)")
                window = []
    return net, hist


def rel_l2_occupied(vol, dense):
    occ = dense >= OCCUPIED_FRACTION * dense.max()
    return float(np.linalg.norm((vol - dense)[occ]) / max(np.linalg.norm(dense[occ]), 1e-300))


def surface_support(cap, voxels, dilate=1):
    """Voxels within `dilate` voxels (per axis) of the hidden geometry's surface samples: an
    occupancy that does not depend on the reconstruction or on any threshold."""
    N = np.asarray((voxels,) * 3)
    lo = np.asarray(cap.volume_position, np.float64) - 0.5 * np.asarray(cap.volume_size, np.float64)
    step = np.asarray(cap.volume_size, np.float64) / (N - 1)
    idx = np.rint((np.asarray(cap.hidden[:, :3], np.float64) - lo) / step).astype(np.int64)
    idx = idx[np.all((idx >= 0) & (idx < N), axis=1)]
    m = np.zeros(tuple(N), bool)
    for dx in range(-dilate, dilate + 1):
        for dy in range(-dilate, dilate + 1):
            for dz in range(-dilate, dilate + 1):
                j = idx + (dx, dy, dz)
                j = j[np.all((j >= 0) & (j < N), axis=1)]
                m[j[:, 0], j[:, 1], j[:, 2]] = True
    return m


def rel_l2_on(vol, dense, mask):
    return float(np.linalg.norm((vol - dense)[mask]) / max(np.linalg.norm(dense[mask]), 1e-300))


def apply_operating_point(plan, net, tau, cfg):
    """Attaches net and sets one proposal configuration on an octree plan: a probability, one
    per level, or {"p": either, "keep": keep_bound_ratio, "fallback": full-resolution fallback
    probability (one or per level), "stride": coarse-level pair stride}."""
    plan.attach_net(net)
    if isinstance(cfg, dict):
        plan.set_octree_options(tau, cfg["p"], keep_bound_ratio=cfg.get("keep", 0.0),
                                fallback_threshold=cfg.get("fallback", 0.0),
                                coarse_pair_stride=cfg.get("stride", 1))
    else:
        plan.set_octree_options(tau, cfg)


def compare(cap, voxels, net, configs, tau=0.25, levels=4, reps=3):
    """Dense vs bound-rule octree vs learned octree (each proposal configuration) on one
    capture; one octree plan is reused with its thresholds replaced."""
    args = plan_args(cap, voxels)
    with core.Plan(*args) as dp:
        td = dp.benchmark(1, reps)["per_exec_ms"]
        dense = dp.volume()
    support = surface_support(cap, voxels) if cap.hidden is not None else None
    rows = []
    with core.Plan.octree(*args, levels=levels, threshold=tau) as op:
        for cfg in [None] + list(configs):
            if cfg is None:
                op.attach_net(None)
            else:
                apply_operating_point(op, net, tau, cfg)
            t = op.benchmark(1, reps)["per_exec_ms"]
            vol = op.volume()
            st = op.octree_stats()
            rows.append({"proposal_threshold": cfg, "ms": t, "speedup": td / t,
                         "rel_l2_occupied": rel_l2_occupied(vol, dense),
                         "rel_l2_surface_support": rel_l2_on(vol, dense, support) if support is not None else None,
                         "fine_voxels": st["active"][-1], "active": st["active"]})
    return {"dense_ms": td, "occupied_voxels": int((dense >= OCCUPIED_FRACTION * dense.max()).sum()),
            "rows": rows}


def evaluate(net, seeds, voxels=256, configs=CANDIDATES, side=128, tau=0.25, label="validation"):
    """Captures like the benchmark's (128^2 confocal, T=4096, 8 points + 2 patches, 256^3) for
    each seed."""
    per = []
    for k, seed in enumerate(seeds):
        cap = scenes.metric_capture(side=side, seed=seed)
        r = compare(cap, voxels, net, configs, tau)
        per.append(r)
        log(f"{label} capture {seed}: dense {r['dense_ms']:.1f} ms; " + "; ".join(
            f"p={row['proposal_threshold']}: x{row['speedup']:.2f} err {row['rel_l2_occupied']:.4f}"
            for row in r["rows"]))
    summary = []
    for j, cfg in enumerate([None] + list(configs)):
        sp = [r["rows"][j]["speedup"] for r in per]
        er = [r["rows"][j]["rel_l2_occupied"] for r in per]
        es = [r["rows"][j]["rel_l2_surface_support"] for r in per if r["rows"][j]["rel_l2_surface_support"] is not None]
        summary.append({"proposal_threshold": cfg, "mean_speedup": float(np.mean(sp)),
                        "min_speedup": float(np.min(sp)), "mean_err": float(np.mean(er)),
                        "max_err": float(np.max(er)),
                        "mean_err_surface_support": float(np.mean(es)) if es else None,
                        "max_err_surface_support": float(np.max(es)) if es else None,
                        "per_capture_err": [float(e) for e in er], "per_capture_speedup": [float(x) for x in sp]})
    return summary, per


def validate(net, n=10, voxels=256, configs=CANDIDATES, side=128, tau=0.25):
    return evaluate(net, range(VALIDATION_SEED_BASE, VALIDATION_SEED_BASE + n), voxels, configs,
                    side, tau, "validation")


def has_fallback(cfg):
    return isinstance(cfg, dict) and any(float(x) > 0 for x in np.atleast_1d(cfg.get("fallback", 0)))


def pick_threshold(summary, max_err=0.015, mean_err=0.0075, need_fallback=True):
    """Fastest proposal configuration whose worst validation error stays <= max_err and whose
    mean error stays <= mean_err (three quarters of the 2 % / 1 % targets: a margin for captures
    the validation set does not hold), among those with a full-resolution fallback band when
    need_fallback (the plain reconstruction as the low-confidence fallback)."""
    ok = [s for s in summary if s["proposal_threshold"] is not None and s["max_err"] <= max_err
          and s["mean_err"] <= mean_err and (has_fallback(s["proposal_threshold"]) or not need_fallback)]
    if not ok:
        return None
    return max(ok, key=lambda s: s["mean_speedup"])["proposal_threshold"]


def save(net, path, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.save(path + ".npy", net.params())
    with open(path + ".json", "w") as f:
        json.dump(meta, f, indent=1)


def load(path=WEIGHTS):
    """(Net with the saved parameters, metadata dict)."""
    params = np.load(path + ".npy")
    meta = json.load(open(path + ".json")) if os.path.exists(path + ".json") else {}
    return core.Net(0, params=params), meta


def select_and_test(net, meta, n_val=10, n_test=20, voxels=256, configs=CANDIDATES):
    """Chooses the proposal configuration on the validation captures, then reports it (and
    the bound rule alone) on the held-out test captures."""
    summary, _ = validate(net, n_val, voxels, configs)
    meta["validation"] = summary
    pick = pick_threshold(summary)
    meta["proposal_threshold"] = pick
    if pick is not None and n_test:
        test, _ = evaluate(net, range(TEST_SEED_BASE, TEST_SEED_BASE + n_test), voxels, [pick],
                           label="test")
        meta["test"] = {"captures": n_test, "seeds": f"{TEST_SEED_BASE}..{TEST_SEED_BASE + n_test - 1}",
                        "bound_rule": test[0], "learned": test[1]}
    return meta


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--scenes", type=int, default=2000)
    t.add_argument("--steps-per-scene", type=int, default=3)
    t.add_argument("--side", type=int, default=64)
    t.add_argument("--voxels", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--pos-weight", type=float, default=8.0)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--validate", type=int, default=10)
    t.add_argument("--test", type=int, default=20)
    t.add_argument("--log-every", type=int, default=50)
    t.add_argument("--label-fraction", type=float, default=0.2)
    t.add_argument("--on-policy", type=float, default=0.5,
                   help="fraction of the scenes (the last ones) trained with the net attached")
    t.add_argument("--init", default=None, help="start from these weights (another epoch)")
    t.add_argument("--out", default=WEIGHTS)
    v = sub.add_parser("validate")
    v.add_argument("--weights", default=WEIGHTS)
    v.add_argument("--n", type=int, default=10)
    v.add_argument("--test", type=int, default=20)
    v.add_argument("--voxels", type=int, default=256)
    v.add_argument("--out", default=None, help="write weights + updated metadata here")
    v.add_argument("--candidates", default=None,
                   help="JSON list of proposal thresholds (a number, or one per level) to choose from")
    a = ap.parse_args(argv)
    if a.cmd == "train":
        t0 = time.perf_counter()
        init = np.load(a.init + ".npy") if a.init else None
        net, hist = train(a.scenes, a.steps_per_scene, a.side, a.voxels, 4, a.lr, a.pos_weight,
                          a.seed, a.log_every, init=init, on_policy=a.on_policy,
                          label_fraction=a.label_fraction)
        meta = {"scenes": a.scenes, "seeds": f"{TRAIN_SEED_BASE}..{TRAIN_SEED_BASE + a.scenes - 1}",
                "steps_per_scene": a.steps_per_scene,
                "capture": f"confocal {a.side}x{a.side}, T=4096, 1 mm bins",
                "voxels": a.voxels, "levels": 4, "lr": a.lr, "pos_weight": a.pos_weight,
                "occupied_fraction": OCCUPIED_FRACTION, "label_fraction": a.label_fraction,
                "on_policy": a.on_policy, "init": os.path.basename(a.init) if a.init else None,
                "train_seconds": time.perf_counter() - t0,
                "history": hist}
        save(net, a.out, meta)
        if a.validate:
            select_and_test(net, meta, a.validate, a.test, a.voxels)
            save(net, a.out, meta)
        print(json.dumps({k: meta[k] for k in meta if k != "history"}), flush=True)
    else:
        net, meta = load(a.weights)
        configs = json.loads(a.candidates) if a.candidates else CANDIDATES
        select_and_test(net, meta, a.n, a.test, a.voxels, configs)
        if a.out:
            save(net, a.out, meta)
        print(json.dumps({k: meta[k] for k in meta if k != "history"}), flush=True)


if __name__ == "__main__":
    main()

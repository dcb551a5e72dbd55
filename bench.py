#!/usr/bin/env python3
"""ForceAtlas2 throughput on one MI355X.

Workload (default): R-MAT over 2^20 ids, 10,000,000 distinct undirected non-loop edges,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 1, loaded with the reference's edge-list rules
(isolated ids vanish); initial layout from glibc random() after srandom(1234) in a square
of side 2*sqrt(V); weak gravity k_g = 1, k_r = 10, theta = 1. One step = one FA2
iteration (tree + walk + attraction + gravity + speed + displacement), all inputs
resident in HBM. W untimed warm-up iterations, then K timed ones, bracketed by a device
synchronisation on both sides.

Also reported: per-kernel-group times from HIP events recorded on the engine's stream
during the timed region, the roofline of the dominant kernel (the Barnes-Hut walk, FP32
VALU: 12 flops per body-node interaction, interactions counted by the kernel itself),
HBM rates of the streaming groups, and the serial CPU baseline (the oracle's restatement
of the reference's CPUForceAtlas2, approximate mode) timed on this host on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FA2 iterations/s (1M-node R-MAT) + time-to-equal-stress, learned vs plain"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def build_workload(name):
    from gpugraphlayout_amd import graphs
    if name == "rmat20":
        s, t = graphs.rmat(20, 10_000_000, seed=1)
        cfg = {"workload": "rmat-scale20-10M-edges", "rmat_abcd": [0.57, 0.19, 0.19, 0.05], "seed": 1,
               "gravity": "weak", "k_g": 1.0, "k_r": 10.0, "theta": 1.0}
        strong, k_g, k_r = False, 1.0, 10.0
    elif name == "ba10k":
        s, t = graphs.barabasi_albert(10_000, 5, 0)
        cfg = {"workload": "barabasi-albert-10k-m5", "seed": 0, "gravity": "strong", "k_g": 1.0, "k_r": 2.0,
               "theta": 1.0}
        strong, k_g, k_r = True, 1.0, 2.0
    else:
        raise SystemExit(f"unknown workload {name}")
    g = graphs.edges_to_graph(s, t)
    side = 2.0 * np.sqrt(g.n)
    pos = graphs.glibc_random_positions(g.n, side, side, seed=1234)
    cfg.update({"nodes": g.n, "edges": g.m, "init": "glibc random() after srandom(1234), square side 2*sqrt(V)"})
    return g, pos, strong, k_g, k_r, cfg


def streaming_bytes(V, E, C):
    """Compulsory bytes per iteration of each streaming group (each array once)."""
    return {
        "attraction": 4 * (V + 1) + 8 * E + 8 * V + 4 * V + 8 * V,   # rowptr, col(2E), pos, mass, fga
        "keys": 8 * V + 8 * V + 4 * V,                               # pos in, keys + ids out
        "sort": 8 * (8 * V + 4 * V) * 2,                             # 8 passes: read + write keys/ids
        "displacement": 8 * V * 3 + 8 * V,                           # pos rw, F, F' rw
    }


def cpu_baseline(g, pos, strong, k_g, k_r, iters):
    from oracle import oracle as orc
    cb = orc.CPUBaseline(g.n, g.src, g.dst, pos, strong, k_g, k_r, use_bh=True)
    cb.step(1)                                   # warm-up
    t0 = time.perf_counter()
    cb.step(iters)
    dt = time.perf_counter() - t0
    cb.close()
    return {"value": iters / dt, "unit": "iterations/s", "cores": 1, "kind": "port",
            "sample": f"{iters} iterations after 1 warm-up of the serial CPUForceAtlas2 restatement "
                      f"(approximate mode) on the same graph and initial layout; {dt:.1f} s on 1 host core",
            "seconds": dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="rmat20")
    ap.add_argument("--group", type=int, default=32)
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", type=int, default=10,
                    help="iterations per candidate launch shape before timing (0: default shapes)")
    ap.add_argument("--graph-replay", type=int, default=100,
                    help="extra iterations replayed as one captured hipGraph per iteration (reported apart)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; a multi-GPU run is not part of this build "
              "(see DESIGN.md for how the path would shard).", file=sys.stderr)
        sys.exit(2)

    from gpugraphlayout_amd import Engine, device_info

    print(f"[bench] building workload {args.workload}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    g, pos, strong, k_g, k_r, cfg = build_workload(args.workload)
    print(f"[bench] V={g.n} E={g.m} built in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    dev = device_info(0)
    eng = Engine(g.n, g.src, g.dst, strong_gravity=int(strong), gravity=k_g, scale=k_r, group_size=args.group)
    eng.set_positions(pos)
    eng.step(args.warmup)
    eng.sync()
    launch = eng.launch_config()
    tune_s = None
    if args.autotune > 0:
        t0 = time.perf_counter()
        launch = eng.autotune(args.autotune)
        tune_s = time.perf_counter() - t0
        print(f"[bench] autotune {launch} in {tune_s:.2f}s", file=sys.stderr, flush=True)
    s0 = eng.stats()
    print(f"[bench] warm-up done ({args.warmup} it), timing {args.steps} it", file=sys.stderr, flush=True)

    eng.set_timing(True)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    dt = time.perf_counter() - t0
    groups = eng.timing()
    eng.set_timing(False)
    s1 = eng.stats()
    value = args.steps / dt
    print(f"[bench] {value:.1f} it/s; groups(ms) {json.dumps({k: round(v, 4) for k, v in groups.items()})}",
          file=sys.stderr, flush=True)

    V, E, C = g.n, g.m, s1["cells"]
    inter = (s1["interactions_total"] - s0["interactions_total"]) / args.steps
    walk_ms = groups["walk"]
    walk_tflops = 12.0 * inter / (walk_ms * 1e-3) / 1e12
    sb = streaming_bytes(V, E, C)
    hbm = {k: {"bytes": b, "ms": groups[k], "GB/s": b / (groups[k] * 1e-3) / 1e9} for k, b in sb.items()}

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "walk_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    replay = None
    if args.graph_replay > 0:
        eng.set_graph_capture(True)
        eng.step(1)
        eng.sync()
        t1 = time.perf_counter()
        eng.step(args.graph_replay)
        eng.sync()
        replay = args.graph_replay / (time.perf_counter() - t1)
        eng.set_graph_capture(False)
    final = eng.stats()
    eng.close()

    cpu = None
    if not args.no_cpu_baseline:
        print(f"[bench] CPU baseline: {args.cpu_iters} iterations", file=sys.stderr, flush=True)
        cpu = cpu_baseline(g, pos, strong, k_g, k_r, args.cpu_iters)
        cpu["cpu_model"] = _cpu_model()
        cpu["nproc"] = os.cpu_count()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(cfg, group_size=args.group, parallelism="single-gpu",
                       walk_block=launch["walk_block"], blocks_per_cu=launch["blocks_per_cu"]),
        "roofline": {
            "bound": "mfma",
            "engine": "fp32 VALU (no MFMA: the walk is not a dense contraction)",
            "kernel": "k_walk (Barnes-Hut force walk)",
            "achieved": walk_tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": walk_tflops / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "flops_per_launch": 12.0 * inter,
            "interactions_per_iteration": inter,
            "avg_launch_ms": walk_ms,
        },
        "cpu_baseline": cpu,
        "nodes_edges_per_s": (V + E) * value,
        "kernel_groups_ms": groups,
        "hbm_streaming": hbm,
        "graph_replay_iterations_per_s": replay,
        "autotune_s": tune_s,
        "cells": C,
        "max_level": s1["max_level"],
        "final_swing": final["swing"],
        "device": dev,
    }
    print(json.dumps(out))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()

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

Other workloads (--workload): ba10k (Barabasi-Albert n = 10,000, m = 5, strong gravity,
k_r = 2, 500 iterations), rmat24 (R-MAT over 2^24 ids, 128,000,000 edges, seed 3, 200
iterations) and ba4096 (4,096 Barabasi-Albert graphs laid out as one batch, 300 iterations,
with the estimate of laying them out one after another).
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


# Per-workload defaults: (timed iterations, CPU-baseline iterations, CPU warm-up iterations).
WORKLOAD_DEFAULTS = {"rmat20": (500, 10, 1), "ba10k": (500, 500, 0), "rmat24": (200, 1, 0), "ba4096": (300, 10, 1)}


def _host_graph(s, t):
    """Reference loader semantics through the C++ host layer (O(E), hash-based)."""
    from gpugraphlayout_amd.host import HostGraph
    hg = HostGraph.from_edges(s, t)
    return hg.arrays()


def build_workload(name):
    from gpugraphlayout_amd import graphs
    if name == "rmat20":
        s, t = graphs.rmat(20, 10_000_000, seed=1)
        cfg = {"workload": "rmat-scale20-10M-edges", "rmat_abcd": [0.57, 0.19, 0.19, 0.05], "seed": 1,
               "gravity": "weak", "k_g": 1.0, "k_r": 10.0, "theta": 1.0}
        strong, k_g, k_r = False, 1.0, 10.0
    elif name == "rmat24":
        s, t = graphs.rmat(24, 128_000_000, seed=3)
        cfg = {"workload": "rmat-scale24-128M-edges", "rmat_abcd": [0.57, 0.19, 0.19, 0.05], "seed": 3,
               "gravity": "weak", "k_g": 1.0, "k_r": 10.0, "theta": 1.0}
        strong, k_g, k_r = False, 1.0, 10.0
    elif name == "ba10k":
        s, t = graphs.barabasi_albert(10_000, 5, 0)
        cfg = {"workload": "barabasi-albert-10k-m5", "seed": 0, "gravity": "strong", "k_g": 1.0, "k_r": 2.0,
               "theta": 1.0}
        strong, k_g, k_r = True, 1.0, 2.0
    else:
        raise SystemExit(f"unknown workload {name}")
    g = _host_graph(s, t)
    del s, t
    side = 2.0 * np.sqrt(g.n)
    pos = graphs.glibc_random_positions(g.n, side, side, seed=1234)
    cfg.update({"nodes": g.n, "edges": g.m, "init": "glibc random() after srandom(1234), square side 2*sqrt(V)"})
    return g, pos, strong, k_g, k_r, cfg


def build_batch(count=4096, seed=2):
    """4,096 Barabasi-Albert graphs: n = round(exp(U[ln 100, ln 20000])), m ~ U{2..8}
    (numpy default_rng(seed)), graph k generated with BA seed seed*2^32 + k and laid out
    from i.i.d. uniform positions in a square of side 2*sqrt(n) (default_rng(seed*2^32 + k))."""
    from gpugraphlayout_amd import graphs
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(100.0), np.log(20000.0), count))).astype(np.int64)
    ms = rng.integers(2, 9, count)
    out = []
    for k in range(count):
        n, m = int(ns[k]), int(ms[k])
        s, t = graphs.barabasi_albert(n, m, (seed << 32) + k)
        g = _host_graph(s, t)
        pos = graphs.uniform_positions(g.n, 2.0 * np.sqrt(g.n), (seed << 32) + k)
        out.append((g, pos))
    return out


def streaming_bytes(V, E, C):
    """Compulsory bytes per iteration of each streaming group (each array once)."""
    return {
        "attraction": 4 * (V + 1) + 8 * E + 8 * V + 4 * V + 8 * V,   # rowptr, col(2E), pos, mass, fga
        "keys": 8 * V + 8 * V + 4 * V,                               # pos in, keys + ids out
        "sort": 8 * (8 * V + 4 * V) * 2,                             # 8 passes: read + write keys/ids
        "displacement": 8 * V * 3 + 8 * V,                           # pos rw, F, F' rw
    }


def cpu_baseline(g, pos, strong, k_g, k_r, iters, warm=1):
    from oracle import oracle as orc
    cb = orc.CPUBaseline(g.n, g.src, g.dst, pos, strong, k_g, k_r, use_bh=True)
    if warm:
        cb.step(warm)
    t0 = time.perf_counter()
    cb.step(iters)
    dt = time.perf_counter() - t0
    cb.close()
    return {"value": iters / dt, "unit": "iterations/s", "cores": 1, "kind": "port",
            "sample": f"{iters} iterations after {warm} warm-up of the serial CPUForceAtlas2 restatement "
                      f"(approximate mode) on the same graph and initial layout; {dt:.1f} s on 1 host core",
            "seconds": dt}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_run(eng, args, steps):
    """Autotune (optional), then `steps` timed iterations with per-group HIP events."""
    tune_s = None
    if args.walk_block > 0:
        eng.set_launch_config(args.walk_block, args.blocks_per_cu)
    launch = eng.launch_config()
    if args.autotune > 0 and args.walk_block <= 0:
        t0 = time.perf_counter()
        launch = eng.autotune(args.autotune)
        tune_s = time.perf_counter() - t0
        log(f"autotune {launch} in {tune_s:.2f}s")
    s0 = eng.stats()
    eng.set_timing(True)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    dt = time.perf_counter() - t0
    groups = eng.timing()
    eng.set_timing(False)
    s1 = eng.stats()
    log(f"{steps / dt:.1f} it/s; groups(ms) {json.dumps({k: round(v, 4) for k, v in groups.items()})}")
    return dt, groups, s0, s1, launch, tune_s


def walk_roofline(groups, s0, s1, steps):
    inter = (s1["interactions_total"] - s0["interactions_total"]) / steps
    walk_ms = groups["walk"]
    walk_tflops = 12.0 * inter / (walk_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "walk_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    return {
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
    }


def base_record(args, steps, dt, value, cfg):
    return {
        "metric": METRIC,
        "value": value,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * dt / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
    }


def run_single(args):
    from gpugraphlayout_amd import Engine, device_info
    steps = args.steps or WORKLOAD_DEFAULTS[args.workload][0]
    cpu_iters = args.cpu_iters or WORKLOAD_DEFAULTS[args.workload][1]
    cpu_warm = WORKLOAD_DEFAULTS[args.workload][2]
    log(f"building workload {args.workload}")
    t0 = time.perf_counter()
    g, pos, strong, k_g, k_r, cfg = build_workload(args.workload)
    log(f"V={g.n} E={g.m} built in {time.perf_counter() - t0:.1f}s")

    dev = device_info(0)
    eng = Engine(g.n, g.src, g.dst, strong_gravity=int(strong), gravity=k_g, scale=k_r, group_size=args.group)
    eng.set_positions(pos)
    eng.step(args.warmup)
    eng.sync()
    log(f"warm-up done ({args.warmup} it), timing {steps} it")
    dt, groups, s0, s1, launch, tune_s = timed_run(eng, args, steps)
    value = steps / dt

    V, E, C = g.n, g.m, s1["cells"]
    sb = streaming_bytes(V, E, C)
    hbm = {k: {"bytes": b, "ms": groups[k], "GB/s": b / (groups[k] * 1e-3) / 1e9} for k, b in sb.items()}

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
        log(f"CPU baseline: {cpu_iters} iterations")
        cpu = cpu_baseline(g, pos, strong, k_g, k_r, cpu_iters, cpu_warm)
        cpu["cpu_model"] = _cpu_model()
        cpu["nproc"] = os.cpu_count()

    out = base_record(args, steps, dt, value,
                      dict(cfg, group_size=args.group, parallelism="single-gpu",
                           walk_block=launch["walk_block"], blocks_per_cu=launch["blocks_per_cu"]))
    out.update({
        "roofline": walk_roofline(groups, s0, s1, steps),
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
    })
    print(json.dumps(out))


def run_batch(args):
    """Batched layout of the 4,096-graph set: one iteration advances every graph. Compared
    with laying out a seeded sample of the graphs one after another (each alone on the
    device, its iteration replayed as a captured hipGraph) scaled to the whole set."""
    from gpugraphlayout_amd import BatchEngine, Engine, device_info
    steps = args.steps or WORKLOAD_DEFAULTS["ba4096"][0]
    strong, k_g, k_r = True, 1.0, 2.0
    log("building 4096 Barabasi-Albert graphs")
    t0 = time.perf_counter()
    items = build_batch()
    V = sum(g.n for g, _ in items)
    E = sum(g.m for g, _ in items)
    log(f"sum V={V} sum E={E} built in {time.perf_counter() - t0:.1f}s")
    dev = device_info(0)
    eng = BatchEngine([(g.n, g.src, g.dst) for g, _ in items], strong_gravity=int(strong), gravity=k_g, scale=k_r,
                      group_size=args.group)
    eng.set_positions(np.concatenate([p for _, p in items]))
    eng.step(args.warmup)
    eng.sync()
    log(f"warm-up done ({args.warmup} it), timing {steps} it")
    dt, groups, s0, s1, launch, tune_s = timed_run(eng, args, steps)
    value = steps / dt
    replay = None
    if args.graph_replay > 0:
        # the whole batch's iteration captured once as a hipGraph, then replayed
        eng.set_graph_capture(True)
        eng.step(1)
        eng.sync()
        t1 = time.perf_counter()
        eng.step(args.graph_replay)
        eng.sync()
        replay = args.graph_replay / (time.perf_counter() - t1)
        eng.set_graph_capture(False)
        log(f"hipGraph replay: {replay:.1f} it/s")
    eng.close()

    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(len(items), args.batch_sample, replace=False))
    seq_s = 0.0
    for k in sample:
        g, pos = items[k]
        e1 = Engine(g.n, g.src, g.dst, strong_gravity=int(strong), gravity=k_g, scale=k_r, group_size=args.group)
        e1.set_graph_capture(True)
        e1.set_positions(pos)
        e1.step(args.warmup)
        e1.sync()
        t1 = time.perf_counter()
        e1.step(steps)
        e1.sync()
        seq_s += time.perf_counter() - t1
        e1.close()
    seq_total = seq_s * len(items) / len(sample)
    log(f"sequential sample: {len(sample)} graphs, {seq_s:.2f}s -> {seq_total:.1f}s for the set "
        f"(batched {dt:.2f}s)")

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc  # noqa: F401  (checked import before timing)
        csample = sample[: args.batch_cpu_sample]
        cpu_iters = args.cpu_iters or WORKLOAD_DEFAULTS["ba4096"][1]
        tsum = 0.0
        for k in csample:
            g, pos = items[k]
            r = cpu_baseline(g, pos, strong, k_g, k_r, cpu_iters, 1)
            tsum += r["seconds"] / cpu_iters
        per_batch_it = tsum * len(items) / len(csample)
        cpu = {"value": 1.0 / per_batch_it, "unit": "iterations/s", "cores": 1, "kind": "port",
               "sample": f"{cpu_iters} iterations (after 1 warm-up) of the serial CPUForceAtlas2 restatement on "
                         f"each of {len(csample)} sampled graphs; mean seconds per graph-iteration x 4096 = "
                         f"{per_batch_it:.2f} s per batch iteration on 1 host core",
               "seconds": tsum * cpu_iters, "cpu_model": _cpu_model(), "nproc": os.cpu_count()}

    cfg = {"workload": "barabasi-albert-4096-graphs", "graphs": len(items), "n": "round(exp(U[ln100, ln20000]))",
           "m": "U{2..8}", "seed": 2, "nodes": V, "edges": E, "gravity": "strong", "k_g": k_g, "k_r": k_r,
           "theta": 1.0, "init": "uniform square of side 2*sqrt(n) per graph", "group_size": args.group,
           "parallelism": "single-gpu", "walk_block": launch["walk_block"], "blocks_per_cu": launch["blocks_per_cu"]}
    out = base_record(args, steps, dt, value, cfg)
    out.update({
        "roofline": walk_roofline(groups, s0, s1, steps),
        "cpu_baseline": cpu,
        "graph_iterations_per_s": value * len(items),
        "nodes_edges_per_s": (V + E) * value,
        "kernel_groups_ms": groups,
        "sequential_estimate_s": seq_total,
        "sequential_sample_graphs": int(len(sample)),
        "batched_s": dt,
        "batched_speedup_vs_sequential": seq_total / dt,
        "graph_replay_iterations_per_s": replay,
        "autotune_s": tune_s,
        "device": dev,
    })
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed iterations (0: the workload's default)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="rmat20", choices=sorted(WORKLOAD_DEFAULTS))
    ap.add_argument("--group", type=int, default=32)
    ap.add_argument("--cpu-iters", type=int, default=0, help="CPU-baseline iterations (0: the workload's default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", type=int, default=30,
                    help="iterations per candidate launch shape before timing (0: default shapes)")
    ap.add_argument("--walk-block", type=int, default=0, help="fixed walk block size (skips the autotuner)")
    ap.add_argument("--blocks-per-cu", type=int, default=8, help="with --walk-block: grid-stride blocks per CU")
    ap.add_argument("--graph-replay", type=int, default=100,
                    help="extra iterations replayed as one captured hipGraph per iteration (reported apart)")
    ap.add_argument("--batch-sample", type=int, default=64, help="graphs laid out alone for the sequential estimate")
    ap.add_argument("--batch-cpu-sample", type=int, default=16, help="graphs timed on the CPU baseline")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only stream MIB MiB through torch.neg (4 B read + 4 B written per element), "
                         "5 times: the known-size kernel that scales the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; a run on several GPUs is not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy > 0:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "calibrate_copy_bytes_written": 4 * n, "launches": 5}))
        return
    if args.workload == "ba4096":
        run_batch(args)
    else:
        run_single(args)


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

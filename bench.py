#!/usr/bin/env python3
"""ForceAtlas2 throughput on one MI355X.

Workload (default): R-MAT over 2^20 ids, 10,000,000 distinct undirected non-loop edges,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 1, loaded with the reference's edge-list rules
(isolated ids vanish); initial layout from glibc random() after srandom(1234) in a square
of side 2*sqrt(V); weak gravity k_g = 1, k_r = 10, theta = 1. One step = one FA2
iteration (tree + walk + attraction + gravity + speed + displacement), all inputs
resident in HBM. W untimed warm-up iterations, then K timed ones, bracketed by a device
synchronisation on both sides: eager launches with HIP events around the force walk only (the
kernel time of the roofline; every event costs a little), or for ba10k, whose iteration is
bound by launch latency, K replays of the iteration captured as one hipGraph, after an equal
eager pass with the walk's events. The per-group breakdown comes from a further K eager
iterations with events around every kernel group.

Also reported:
  roofline      the Barnes-Hut walk's (FP32 VALU: 12 flops per body-node interaction,
                interactions counted by the kernel itself), its time from HIP events on the
                engine's stream over the timed region, and its HBM bytes per launch measured
                in this run by two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over the first
                50 timed iterations of a child run of the same workload and launch shape, each
                scaled by what the counter reads for a copy of known size;
  cpu_baseline  the serial CPU restatement of the reference's CPUForceAtlas2 (approximate
                mode) timed on this host on a bounded sample of the same workload;
  time_to_equal_stress  (default workload) the learned initialiser on held-out stochastic
                block model graphs (50,000 nodes): plain FA2 from a random start for 1,000
                iterations against the network's start refined by FA2 until it reaches the
                plain run's normalized stress, wall times including features and inference;
  quality       final normalized stress (10^6 seeded node pairs) and swing of the GPU run and,
                for ba10k, of the CPU baseline's full 500-iteration run from the same start.

Other workloads (--workload): ba10k (Barabasi-Albert n = 10,000, m = 5, strong gravity,
k_r = 2, 500 iterations), rmat24 (R-MAT over 2^24 ids, 128,000,000 edges, seed 3, 200
iterations) and ba4096 (4,096 Barabasi-Albert graphs laid out as one batch, 300 iterations,
with the same graphs laid out one after another for comparison).
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FA2 iterations/s (1M-node R-MAT) + time-to-equal-stress, learned vs plain"
FP32_PEAK_TFLOPS = 157.3          # MI355X FP32 vector, packed
FP32_UNPACKED_TFLOPS = 78.6       # without packed FP32 (this build compiles none)
HBM_PEAK_GBS = 8000.0
MODEL = os.path.join(ROOT, "gpugraphlayout_amd", "data", "mpnn_sbm50k.npy")

# Per-workload defaults: (timed iterations, CPU-baseline iterations, CPU warm-up iterations).
WORKLOAD_DEFAULTS = {"rmat20": (500, 10, 1), "ba10k": (500, 500, 0), "rmat24": (200, 1, 0), "ba4096": (300, 10, 1)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def streaming_bytes(V, E):
    """Compulsory bytes per iteration of each streaming group (each array once)."""
    return {
        "attraction": 4 * (V + 1) + 8 * E + 8 * V + 4 * V + 8 * V,   # rowptr, col(2E), pos, mass, fga
        "keys": 8 * V + 8 * V + 4 * V,                               # pos in, keys + ids out
        "sort": 8 * (8 * V + 4 * V) * 2,                             # 8 passes: read + write keys/ids
        "displacement": 8 * V * 3 + 8 * V,                           # pos rw, F, F' rw
    }


def walk_algorithmic_bytes(V, C, split):
    """The walk's compulsory bytes per launch: every record (V bodies + C cells, 16 B) read once
    and every cell's child list (16 B) once, each body's own {x, y, m} (16 B, the sorted copy)
    and id (4 B), and F written (8 B) per body; the fused epilogue (one stream) also reads F_ga
    and F' (8 B each), which the split walk leaves to the combine kernel."""
    return 16 * (V + C) + 16 * C + 20 * V + (8 if split else 24) * V


def cpu_baseline(g, pos, strong, k_g, k_r, iters, warm=1, keep=False):
    from oracle import oracle as orc
    cb = orc.CPUBaseline(g.n, g.src, g.dst, pos, strong, k_g, k_r, use_bh=True)
    if warm:
        cb.step(warm)
    t0 = time.perf_counter()
    cb.step(iters)
    dt = time.perf_counter() - t0
    out = {"value": iters / dt, "unit": "iterations/s", "cores": 1, "kind": "port",
           "sample": f"{iters} iterations after {warm} warm-up of the serial CPUForceAtlas2 restatement "
                     f"(approximate mode) on the same graph and initial layout; {dt:.1f} s on 1 host core",
           "seconds": dt, "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
    if keep:
        out["_final"] = (cb.positions(), float(cb.state()[4]))
    cb.close()
    return out


def timed_run(eng, args, steps):
    """Autotune (optional), then `steps` timed eager iterations with HIP events around the
    walk (its average time per launch, in ms, is returned with the region's wall time)."""
    tune_s = None
    if args.walk_block > 0:
        eng.set_launch_config(args.walk_block, args.blocks_per_cu, args.attraction_stream,
                              args.attraction_blocks_per_cu)
    launch = eng.launch_config()
    if args.autotune > 0 and args.walk_block <= 0:
        t0 = time.perf_counter()
        launch = eng.autotune(args.autotune)
        tune_s = time.perf_counter() - t0
        log(f"autotune {launch} in {tune_s:.2f}s")
    s0 = eng.stats()
    eng.set_timing("walk")
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    dt = time.perf_counter() - t0
    walk_ms = eng.timing()["walk"]
    eng.set_timing(False)
    s1 = eng.stats()
    log(f"{steps / dt:.1f} it/s; walk {walk_ms:.4f} ms")
    return dt, walk_ms, s0, s1, launch, tune_s


def group_breakdown(eng, steps):
    """`steps` more eager iterations with HIP events around every kernel group: the average ms
    per iteration of each group, and the pass's own rate (the events slow it a little)."""
    eng.set_timing(True)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    rate = steps / (time.perf_counter() - t0)
    groups = eng.timing()
    eng.set_timing(False)
    log(f"groups(ms) {json.dumps({k: round(v, 4) for k, v in groups.items()})} at {rate:.1f} it/s")
    return groups, rate


def walk_roofline(walk_ms, s0, s1, steps, V, split):
    inter = (s1["interactions_total"] - s0["interactions_total"]) / steps
    tflops = 12.0 * inter / (walk_ms * 1e-3) / 1e12
    return {
        "bound": "valu",
        "engine": "fp32 VALU (no MFMA: the walk is not a dense contraction)",
        "kernel": "k_walk (Barnes-Hut force walk)",
        "achieved": tflops,
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": tflops / FP32_PEAK_TFLOPS,
        "frac_attainable": tflops / FP32_UNPACKED_TFLOPS,
        "attainable_peak": FP32_UNPACKED_TFLOPS,
        "attainable_note": "the build compiles no packed FP32 (v_pk_*_f32), which halves the issuable FP32 rate",
        "traffic": None,
        "flops_per_launch": 12.0 * inter,
        "interactions_per_iteration": inter,
        "avg_launch_ms": walk_ms,
        "algorithmic_bytes_per_launch": walk_algorithmic_bytes(V, s1["cells"], split),
    }


# ---- HBM traffic per walk launch, measured in this run ------------------------------------
def _counter_rows(outdir):
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _per_dispatch(rows, counter, name_part):
    """Counter value (summed over its rows) per dispatch of kernels whose name contains name_part,
    in dispatch order."""
    acc = {}
    for r in rows:
        if r.get("Counter_Name") != counter or name_part not in r.get("Kernel_Name", ""):
            continue
        d = int(r.get("Dispatch_Id") or r.get("Correlation_Id") or len(acc))
        acc[d] = acc.get(d, 0.0) + float(r["Counter_Value"])
    return [acc[k] for k in sorted(acc)]


class ProfilerCrash(RuntimeError):
    """The profiled child ended on a signal: no further GPU work in this process."""


def _profile(counter, child_args, tag, kernel_regex):
    exe = shutil.which("rocprofv3")
    if not exe:
        raise RuntimeError("rocprofv3 not found")
    outdir = tempfile.mkdtemp(prefix=f"bench_pmc_{tag}_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    # counters for the one kernel only (a counter record for every dispatch of a long run is
    # more than the profiler holds)
    cmd = ["timeout", "-s", "KILL", "240", exe, "--pmc", counter, "--kernel-include-regex", kernel_regex,
           "-d", outdir, "-o", "run", "--output-format", "csv", "--", sys.executable,
           os.path.abspath(__file__)] + child_args
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
    rows = _counter_rows(outdir)
    shutil.rmtree(outdir, ignore_errors=True)
    if not rows:
        msg = f"rocprofv3 --pmc {counter} ({tag}) gave no counters (exit {r.returncode})"
        if r.returncode < 0 or r.returncode >= 128:
            raise ProfilerCrash(msg)
        raise RuntimeError(f"{msg}: {(r.stderr or '')[-300:]}")
    return rows


TRAFFIC_STEPS = 50   # the profiler does not survive a counter collection over hundreds of iterations


def measure_traffic(args, launch, steps):
    """Walk HBM bytes per launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) averaged over the
    first TRAFFIC_STEPS timed iterations of the same workload, start and launch shape in a child
    run, each scaled by the ratio of the known bytes a 256 MiB copy moves to what the same
    counter reads for it."""
    steps = min(steps, TRAFFIC_STEPS)
    child = ["--traffic-child", "--workload", args.workload, "--group", str(args.group), "--warmup", str(args.warmup),
             "--steps", str(steps), "--walk-block", str(launch["walk_block"]),
             "--blocks-per-cu", str(launch["blocks_per_cu"]),
             "--attraction-stream", str(launch["attraction_stream"]),
             "--attraction-blocks-per-cu", str(launch["attraction_blocks_per_cu"])]
    cal = ["--calibrate-copy", "256"]
    copy_bytes = 256 * (1 << 20)
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        walk = _per_dispatch(_profile(counter, child, "walk", "k_walk"), counter, "k_walk")[-steps:]
        copy = _per_dispatch(_profile(counter, cal, "copy", "neg"), counter, "neg")
        if not walk or not copy:
            raise RuntimeError(f"no {counter} rows for the walk or the copy")
        factor = copy_bytes / (1024.0 * float(np.median(copy)))
        out[counter] = {"kib_per_dispatch": float(np.mean(walk)), "copy_kib": float(np.median(copy)),
                        "factor": factor, "bytes": float(np.mean(walk)) * 1024.0 * factor,
                        "dispatches": len(walk)}
    out["bytes_per_launch"] = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    return out


def traffic_child(args):
    """The workload and launch shape of the parent's timed run, under a counter collection."""
    from gpugraphlayout_amd import Engine, workloads
    g, pos, prm, _ = workloads.single(args.workload)
    eng = Engine(g.n, g.src, g.dst, group_size=args.group, **prm)
    eng.set_launch_config(args.walk_block, args.blocks_per_cu, args.attraction_stream, args.attraction_blocks_per_cu)
    eng.set_positions(pos)
    eng.step(args.warmup + args.steps)
    eng.sync()
    eng.close()


# ---- time to equal stress (the learned initialiser) ------------------------------------
def time_to_equal_stress(ngraphs, starts, per_start):
    from gpugraphlayout_amd import learned
    if not os.path.exists(MODEL):
        return {"error": f"no trained network at {os.path.relpath(MODEL, ROOT)}"}
    m, meta = learned.load_model(MODEL)
    seeds = list(range(5000, 5000 + ngraphs))
    t0 = time.perf_counter()
    res = learned.evaluate_sbm(m, meta["target_scale"], seeds, meta["kappa"], meta["feature_steps"], meta["check"],
                               multistart=(starts, per_start) if starts > 1 else None,
                               log=lambda s: log("learned " + s[:160]))
    m.close()
    single = learned.summarise(res)
    plain_s = [r["plain_seconds"] for r in res]
    out = {
        "workload": "sbm-50k-32-blocks-dirichlet2 (p_in = 8/|B|, p_out = 0.5/V), weak gravity k_g = 1, k_r = 2",
        "held_out_seeds": [seeds[0], seeds[-1]],
        "graphs": len(seeds),
        "plain": "uniform random start, 1,000 FA2 iterations (hipGraph replay); NS on 10^6 seeded pairs",
        "learned": "features + inference written into the engine, then FA2 until NS <= the plain run's final NS; "
                   "wall time of a rerun without the checks (features, inference and iterations included)",
        "model": os.path.relpath(MODEL, ROOT),
        "trained_on": {k: meta.get(k) for k in ("train_seeds", "train_steps", "target_iterations", "command")},
        "target_time_ratio": 0.5,
        "plain_seconds_median": float(np.median(plain_s)),
        "single_start": single,
        "seconds": time.perf_counter() - t0,
    }
    if starts > 1:
        out["multi_start"] = {
            "starts": starts, "iterations_per_start": per_start,
            "protocol": "network proposals from different feature seeds in turn, each refined up to the per-start "
                        "budget with stress checks, until one reaches the plain run's NS; timed as run (checks "
                        "included)",
            "learned": learned.summarise_multistart(res, "multistart_learned"),
            "random_starts_same_protocol": learned.summarise_multistart(res, "multistart_random"),
        }
    return out


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
    from gpugraphlayout_amd import Engine, device_info, workloads
    steps = args.steps or WORKLOAD_DEFAULTS[args.workload][0]
    cpu_iters = args.cpu_iters or WORKLOAD_DEFAULTS[args.workload][1]
    cpu_warm = WORKLOAD_DEFAULTS[args.workload][2]
    log(f"building workload {args.workload}")
    t0 = time.perf_counter()
    g, pos, prm, cfg = workloads.single(args.workload)
    log(f"V={g.n} E={g.m} built in {time.perf_counter() - t0:.1f}s")

    dev = device_info(0)
    eng = Engine(g.n, g.src, g.dst, group_size=args.group, **prm)
    eng.set_positions(pos)
    eng.step(args.warmup)
    eng.sync()
    log(f"warm-up done ({args.warmup} it), timing {steps} it")
    dt, walk_ms, s0, s1, launch, tune_s = timed_run(eng, args, steps)
    value = steps / dt
    eager_value = value
    timed = args.timed or ("replay" if args.workload == "ba10k" else "eager")

    V, E = g.n, g.m
    roof = walk_roofline(walk_ms, s0, s1, steps, V, launch["attraction_stream"] == 1)

    replay = None
    replay_iters = steps if timed == "replay" else args.graph_replay
    if replay_iters > 0:
        # the iteration captured once as a hipGraph (one launch per iteration), then replayed
        eng.set_graph_capture(True)
        eng.step(1)
        eng.sync()
        t1 = time.perf_counter()
        eng.step(replay_iters)
        eng.sync()
        replay_dt = time.perf_counter() - t1
        replay = replay_iters / replay_dt
        eng.set_graph_capture(False)
        if timed == "replay":
            # the timed region: K replayed iterations; the roofline comes from the K eager
            # iterations with the walk's HIP events just before
            dt, value = replay_dt, replay
            log(f"hipGraph replay (timed region): {value:.1f} it/s")
    final = eng.stats()
    groups, groups_rate = group_breakdown(eng, steps)
    sb = streaming_bytes(V, E)
    hbm = {k: {"bytes": b, "ms": groups[k], "GB/s": b / (groups[k] * 1e-3) / 1e9} for k, b in sb.items()}

    # layout quality: a fresh run of the workload's own iteration count from the same start, so
    # the GPU's and the CPU baseline's final layouts are comparable (ba10k runs the CPU path to
    # the end as well)
    quality = None
    cpu = None
    if not args.no_quality:
        iters = WORKLOAD_DEFAULTS[args.workload][0]
        eng.set_positions(pos)
        eng.reset()
        eng.step(iters)
        eng.sync()
        eng.stress_setup(1024, 977, seed=0)
        ns_gpu, _ = eng.normalized_stress()
        quality = {"iterations": iters, "pairs": "1,024 sources x 977 targets, seed 0, hop distances by BFS",
                   "gpu": {"normalized_stress": ns_gpu, "final_swing": eng.stats()["swing"]}}
        if args.workload == "ba10k" and not args.no_cpu_baseline:
            log(f"CPU baseline: {cpu_iters} iterations (full run)")
            cpu = cpu_baseline(g, pos, prm["strong_gravity"], prm["gravity"], prm["scale"], cpu_iters, cpu_warm,
                               keep=True)
            cpos, cswing = cpu.pop("_final")
            quality["cpu"] = {"normalized_stress": eng.normalized_stress(cpos)[0], "final_swing": cswing}
    eng.close()

    if cpu is None and not args.no_cpu_baseline:
        log(f"CPU baseline: {cpu_iters} iterations")
        cpu = cpu_baseline(g, pos, prm["strong_gravity"], prm["gravity"], prm["scale"], cpu_iters, cpu_warm)

    tte = None
    if args.workload == "rmat20" and args.learned_graphs > 0:
        log(f"time to equal stress: {args.learned_graphs} held-out SBM graphs")
        tte = time_to_equal_stress(args.learned_graphs, args.learned_starts, args.learned_per_start)

    # last: the counter passes run in child processes (after a child's crash nothing else runs)
    if args.traffic:
        try:
            t0 = time.perf_counter()
            tr = measure_traffic(args, launch, steps)
            roof["traffic"] = tr["bytes_per_launch"]
            roof["traffic_counters"] = {k: tr[k] for k in ("FETCH_SIZE", "WRITE_SIZE")}
            roof["traffic_window"] = (f"timed iterations 1-{min(steps, TRAFFIC_STEPS)} of a child run of the same "
                                      f"workload, start and launch shape")
            log(f"walk traffic {tr['bytes_per_launch'] / 1e6:.1f} MB per launch in {time.perf_counter() - t0:.0f}s")
        except Exception as exc:   # the line still reports everything else
            roof["traffic_error"] = str(exc)[:400]
            log(f"traffic not measured: {exc}")

    out = base_record(args, steps, dt, value,
                      dict(cfg, group_size=args.group, parallelism="single-gpu",
                           walk_block=launch["walk_block"], blocks_per_cu=launch["blocks_per_cu"],
                           attraction_stream=launch["attraction_stream"],
                           attraction_blocks_per_cu=launch["attraction_blocks_per_cu"],
                           timed=("hipGraph replay, one launch per iteration; the roofline from the same number "
                                  "of eager iterations with HIP events around the walk just before"
                                  if timed == "replay" else "eager launches with HIP events around the walk"),
                           kernel_groups=("a further pass of the same number of eager iterations with HIP events "
                                          "around every group")))
    out.update({
        "eager_iterations_per_s": eager_value,
        "groups_pass_iterations_per_s": groups_rate,
        "roofline": roof,
        "cpu_baseline": cpu,
        "time_to_equal_stress": tte,
        "quality": quality,
        "nodes_edges_per_s": (V + E) * value,
        "kernel_groups_ms": groups,
        "hbm_streaming": hbm,
        "graph_replay_iterations_per_s": replay,
        "autotune_s": tune_s,
        "cells": s1["cells"],
        "max_level": s1["max_level"],
        "final_swing": final["swing"],
        "device": dev,
    })
    print(json.dumps(out))


def run_batch(args):
    """Batched layout of the 4,096-graph set: one iteration advances every graph. Compared
    with laying out a seeded sample of the graphs one after another (each alone on the
    device, its iteration replayed as a captured hipGraph) scaled to the whole set; the
    sampled graphs' final layouts are compared with the batch's bit for bit."""
    from gpugraphlayout_amd import BatchEngine, Engine, device_info, workloads
    steps = args.steps or WORKLOAD_DEFAULTS["ba4096"][0]
    prm = workloads.BATCH_PARAMS
    log("building 4096 Barabasi-Albert graphs")
    t0 = time.perf_counter()
    items = workloads.batch4096()
    V = sum(g.n for g, _ in items)
    E = sum(g.m for g, _ in items)
    off = np.concatenate([[0], np.cumsum([g.n for g, _ in items])])
    log(f"sum V={V} sum E={E} built in {time.perf_counter() - t0:.1f}s")
    dev = device_info(0)
    eng = BatchEngine([(g.n, g.src, g.dst) for g, _ in items], group_size=args.group, **prm)
    eng.set_positions(np.concatenate([p for _, p in items]))
    eng.step(args.warmup)
    eng.sync()
    log(f"warm-up done ({args.warmup} it), timing {steps} it")
    dt, walk_ms, s0, s1, launch, tune_s = timed_run(eng, args, steps)
    value = steps / dt
    batch_final = eng.positions()
    groups, groups_rate = group_breakdown(eng, steps)
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
    same = 0
    for k in sample:
        g, pos = items[k]
        e1 = Engine(g.n, g.src, g.dst, group_size=args.group, **prm)
        e1.set_graph_capture(True)
        e1.set_positions(pos)
        e1.step(args.warmup)
        e1.sync()
        t1 = time.perf_counter()
        e1.step(steps)
        e1.sync()
        seq_s += time.perf_counter() - t1
        same += int(np.array_equal(e1.positions().view(np.uint32), batch_final[off[k]:off[k + 1]].view(np.uint32)))
        e1.close()
    seq_total = seq_s * len(items) / len(sample)
    log(f"sequential sample: {len(sample)} graphs, {seq_s:.2f}s -> {seq_total:.1f}s for the set "
        f"(batched {dt:.2f}s); {same}/{len(sample)} bit-identical to the batch")

    cpu = None
    if not args.no_cpu_baseline:
        csample = sample[: args.batch_cpu_sample]
        cpu_iters = args.cpu_iters or WORKLOAD_DEFAULTS["ba4096"][1]
        tsum = 0.0
        for k in csample:
            g, pos = items[k]
            r = cpu_baseline(g, pos, prm["strong_gravity"], prm["gravity"], prm["scale"], cpu_iters, 1)
            tsum += r["seconds"] / cpu_iters
        per_batch_it = tsum * len(items) / len(csample)
        cpu = {"value": 1.0 / per_batch_it, "unit": "iterations/s", "cores": 1, "kind": "port",
               "sample": f"{cpu_iters} iterations (after 1 warm-up) of the serial CPUForceAtlas2 restatement on "
                         f"each of {len(csample)} sampled graphs; mean seconds per graph-iteration x 4096 = "
                         f"{per_batch_it:.2f} s per batch iteration on 1 host core",
               "seconds": tsum * cpu_iters, "cpu_model": _cpu_model(), "nproc": os.cpu_count()}

    cfg = {"workload": "barabasi-albert-4096-graphs", "graphs": len(items), "n": "round(exp(U[ln100, ln20000]))",
           "m": "U{2..8}", "seed": 2, "nodes": V, "edges": E, "gravity": "strong", "k_g": prm["gravity"],
           "k_r": prm["scale"], "theta": 1.0, "init": "uniform square of side 2*sqrt(n) per graph",
           "group_size": args.group, "parallelism": "single-gpu", "walk_block": launch["walk_block"],
           "blocks_per_cu": launch["blocks_per_cu"], "attraction_stream": launch["attraction_stream"],
           "attraction_blocks_per_cu": launch["attraction_blocks_per_cu"]}
    out = base_record(args, steps, dt, value, cfg)
    out.update({
        "roofline": walk_roofline(walk_ms, s0, s1, steps, V, launch["attraction_stream"] == 1),
        "cpu_baseline": cpu,
        "graph_iterations_per_s": value * len(items),
        "nodes_edges_per_s": (V + E) * value,
        "kernel_groups_ms": groups,
        "groups_pass_iterations_per_s": groups_rate,
        "sequential_estimate_s": seq_total,
        "sequential_sample_graphs": int(len(sample)),
        "sequential_sample_bit_identical": int(same),
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
    ap.add_argument("--no-quality", action="store_true", help="skip the final-layout quality run")
    ap.add_argument("--autotune", type=int, default=30,
                    help="iterations per candidate launch shape before timing (0: default shapes)")
    ap.add_argument("--walk-block", type=int, default=0, help="fixed walk block size (skips the autotuner)")
    ap.add_argument("--blocks-per-cu", type=int, default=8, help="with --walk-block: grid-stride blocks per CU")
    ap.add_argument("--attraction-stream", type=int, default=0, choices=(0, 1, 2),
                    help="with --walk-block: 0 attraction on the engine's stream; 1 on a second stream beside "
                         "the walk; 2 on a second stream beside keys/sort/tree")
    ap.add_argument("--attraction-blocks-per-cu", type=int, default=0,
                    help="with --walk-block: attraction grid per compute unit (0: one block per work item)")
    ap.add_argument("--graph-replay", type=int, default=100,
                    help="extra iterations replayed as one captured hipGraph per iteration (reported apart)")
    ap.add_argument("--timed", choices=("eager", "replay"), default=None,
                    help="single graphs: the timed region as eager launches with HIP events around the walk "
                         "(default) or as hipGraph replays (default for ba10k, whose iteration is launch-bound), "
                         "the walk's time then from an equal eager pass")
    ap.add_argument("--traffic", type=int, default=None,
                    help="1: measure the walk's HBM bytes with rocprofv3 counters in child runs (default: on for "
                         "rmat20)")
    ap.add_argument("--learned-graphs", type=int, default=20,
                    help="held-out SBM graphs for time-to-equal-stress (rmat20 only; 0: skip)")
    ap.add_argument("--learned-starts", type=int, default=4, help="network proposals in the multi-start protocol")
    ap.add_argument("--learned-per-start", type=int, default=100, help="FA2 iterations per proposal")
    ap.add_argument("--batch-sample", type=int, default=64, help="graphs laid out alone for the sequential estimate")
    ap.add_argument("--batch-cpu-sample", type=int, default=16, help="graphs timed on the CPU baseline")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only stream MIB MiB through torch.neg (4 B read + 4 B written per element), "
                         "5 times: the known-size kernel that scales the FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
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
    if args.traffic_child:
        traffic_child(args)
        return
    if args.traffic is None:
        args.traffic = 1 if args.workload == "rmat20" else 0
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

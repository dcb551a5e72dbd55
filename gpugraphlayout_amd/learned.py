"""Learned initial layouts for ForceAtlas2: training data, training, and the
time-to-equal-stress comparison against plain (random-start) FA2.

Workload: stochastic block model graphs (50,000 nodes, 32 blocks with Dirichlet(2) sizes,
p_in = 8/|B|, p_out = 0.5/V), training seeds 1000.., held-out seeds 5000..; targets are
plain FA2 layouts after 1,000 iterations (weak gravity k_g = 1, k_r = 2, theta = 1),
computed as batched layouts. The network (gfa2_mpnn_*) is trained on the device.

Evaluation per held-out graph:
  plain   : random start (uniform square of side 2*sqrt(V)), N FA2 iterations; wall time
            of the iterations; final normalized stress NS_plain on 10^6 seeded pairs.
  learned : features + forward pass on the device written into the engine, then FA2
            until NS <= NS_plain. The iteration count k* is found with stress checks every
            `check` iterations; the wall time is then measured on a clean rerun of
            inference + k* iterations (no checks inside the timed region).
Usage: python -m gpugraphlayout_amd.learned --train 200 --test 20 --out result.json
"""
import argparse
import json
import sys
import time

import numpy as np

from . import BatchEngine, Engine, Mpnn, graphs

FA2 = dict(scale=2.0, gravity=1.0, strong_gravity=0)


def sbm_graph(seed, n=50_000, blocks=32, alpha=2.0):
    s, t, _ = graphs.sbm(n, blocks, alpha, seed)
    return graphs.edges_to_graph(s, t)


def start_positions(g, seed):
    return graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), seed)


def fa2_targets(gs, iters, chunk_nodes=4_000_000, seed0=0, log=None):
    """Plain FA2 layouts after `iters` iterations, many graphs per batched launch chain."""
    out = [None] * len(gs)
    k = 0
    while k < len(gs):
        j, total = k, 0
        while j < len(gs) and (j == k or total + gs[j].n <= chunk_nodes):
            total += gs[j].n
            j += 1
        be = BatchEngine([(g.n, g.src, g.dst) for g in gs[k:j]], **FA2)
        be.set_graph_capture(True)
        be.set_positions(np.concatenate([start_positions(g, seed0 + k + q) for q, g in enumerate(gs[k:j])]))
        be.step(iters)
        be.sync()
        for q, p in enumerate(be.split(be.positions())):
            out[k + q] = p.copy()
        be.close()
        if log:
            log(f"targets {j}/{len(gs)}")
        k = j
    return out


def train(gs, targets, steps, lr=3e-3, R=8, seed=0, feature_steps=2, log=None):
    m = Mpnn(seed=seed)
    for k, (g, t) in enumerate(zip(gs, targets)):
        rowptr, col = g.csr()
        m.add_graph(g.n, rowptr, col, target=t, feature_seed=k, steps=feature_steps)
    done = 0
    while done < steps:
        step = min(500, steps - done)
        loss = m.train(step, lr=lr, R=R, seed=seed + done)
        done += step
        if log:
            log(f"train {done}/{steps} loss {loss:.5f}")
    return m


def evaluate(m, g, seed, iters_plain=1000, check=10, feature_steps=2, nsources=1024, ntargets=977):
    rowptr, col = g.csr()
    # plain run
    e = Engine(g.n, g.src, g.dst, **FA2)
    e.set_graph_capture(True)
    e.stress_setup(nsources, ntargets, seed=seed)
    e.set_positions(start_positions(g, 10_000 + seed))
    e.sync()
    t0 = time.perf_counter()
    e.step(iters_plain)
    e.sync()
    t_plain = time.perf_counter() - t0
    ns_plain, _ = e.normalized_stress()
    ns_random_start, _ = e.normalized_stress(start_positions(g, 10_000 + seed))
    # learned: find k*
    gid = m.add_graph(g.n, rowptr, col, feature_seed=10_000 + seed, steps=feature_steps)
    scale = m.target_scale()
    e.reset()
    m.infer_to_engine(gid, scale, e)
    ns_start, _ = e.normalized_stress()
    kstar = None
    it = 0
    while it <= iters_plain:
        if e.normalized_stress()[0] <= ns_plain:
            kstar = it
            break
        e.step(check)
        it += check
    result = {"seed": seed, "nodes": g.n, "edges": g.m, "plain_iterations": iters_plain, "plain_seconds": t_plain,
              "ns_plain": ns_plain, "ns_random_start": ns_random_start, "ns_learned_start": ns_start,
              "learned_iterations": kstar}
    if kstar is not None:
        # clean timed rerun: features + inference + k* iterations
        e.reset()
        e.sync()
        t0 = time.perf_counter()
        gid2 = m.add_graph(g.n, rowptr, col, feature_seed=10_000 + seed, steps=feature_steps)
        m.infer_to_engine(gid2, scale, e)
        e.step(kstar)
        e.sync()
        t_learned = time.perf_counter() - t0
        result["learned_seconds"] = t_learned
        result["ns_learned"] = e.normalized_stress()[0]
        result["time_ratio"] = t_learned / t_plain
    e.close()
    return result


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200, help="training graphs (seeds 1000..)")
    ap.add_argument("--test", type=int, default=20, help="held-out graphs (seeds 5000..)")
    ap.add_argument("--nodes", type=int, default=50_000)
    ap.add_argument("--target-iters", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[learned] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    train_gs = [sbm_graph(1000 + k, n=args.nodes) for k in range(args.train)]
    log(f"{len(train_gs)} training graphs in {time.perf_counter() - t0:.1f}s")
    t1 = time.perf_counter()
    targets = fa2_targets(train_gs, args.target_iters, log=log)
    log(f"targets in {time.perf_counter() - t1:.1f}s")
    t2 = time.perf_counter()
    m = train(train_gs, targets, args.steps, lr=args.lr, log=log)
    train_seconds = time.perf_counter() - t2
    del train_gs, targets
    results = []
    for k in range(args.test):
        g = sbm_graph(5000 + k, n=args.nodes)
        r = evaluate(m, g, 5000 + k, iters_plain=args.target_iters)
        results.append(r)
        log(json.dumps(r))
    ratios = [r["time_ratio"] for r in results if "time_ratio" in r]
    summary = {"train_graphs": args.train, "test_graphs": args.test, "nodes": args.nodes,
               "target_iterations": args.target_iters, "train_steps": args.steps, "train_seconds": train_seconds,
               "reached": len(ratios), "median_time_ratio": float(np.median(ratios)) if ratios else None,
               "mean_time_ratio": float(np.mean(ratios)) if ratios else None, "results": results}
    print(json.dumps(summary))
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(summary, fh, indent=1)


if __name__ == "__main__":
    main()

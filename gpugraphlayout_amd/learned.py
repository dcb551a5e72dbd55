"""Learned initial layouts for ForceAtlas2: training data, training, and the
time-to-equal-stress comparison against plain (random-start) FA2.

Two workloads:

  sbm50k  stochastic block model graphs (50,000 nodes, 32 blocks with Dirichlet(2) sizes,
          p_in = 8/|B|, p_out = 0.5/V); training seeds 1000.., held-out seeds 5000..;
          FA2 with weak gravity k_g = 1, k_r = 2, theta = 1; plain run 1,000 iterations.
  rmat    the scale test: trained on small R-MAT graphs (2^14..2^17 ids, 8 edges per id,
          (0.57, 0.19, 0.19, 0.05), seeds 1000..), tested on R-MAT 2^24 with 128M edges
          (seed 3) and on R-MAT 2^20 with 10M edges (seed 1); FA2 with weak gravity
          k_g = 1, k_r = 10; plain run 200 (2^24) or 500 (2^20) iterations from the
          reference's glibc start (srandom(1234), square of side 2*sqrt(V)).

Targets are plain FA2 layouts after 1,000 iterations from a uniform random start, computed
as batched layouts; the network (gfa2_mpnn_*) is trained on the device.

Evaluation per test graph:
  plain   : random start, N FA2 iterations (hipGraph replay); wall time of the iterations;
            final normalized stress NS_plain on 10^6 seeded pairs.
  learned : features (from the engine's device CSR) + forward pass written into the
            engine, then FA2 until NS <= NS_plain. The iteration count k* is found with
            stress checks every `check` iterations; the wall time is then measured on a
            clean rerun of features + inference + k* iterations (no checks inside the
            timed region).

The learned start is placed at RMS radius kappa * c * sqrt(k_r * E): c is the mean of
rms / sqrt(k_r * E) over the training targets (FA2's equilibrium size grows like the square
root of k_r times the edge count), kappa a start compression picked on validation graphs
(`--validate`: R-MAT 2^24 with 128M edges, seed 4, 200 iterations), never on the test
graphs.

Usage: python -m gpugraphlayout_amd.learned --workload sbm50k --train 200 --test 20 --out r.json
       python -m gpugraphlayout_amd.learned --workload rmat --out r.json
       python -m gpugraphlayout_amd.learned --workload rmat --validate
"""
import argparse
import json
import sys
import time

import numpy as np

from . import BatchEngine, Engine, Mpnn, graphs

FA2 = dict(scale=2.0, gravity=1.0, strong_gravity=0)          # sbm50k
FA2_RMAT = dict(scale=10.0, gravity=1.0, strong_gravity=0)    # rmat

# defaults picked by validation (feature smoothing rounds, start compression kappa)
DEFAULTS = {"sbm50k": {"feature_steps": 2, "kappa": 0.5, "train": 200, "check": 10, "test": 100},
            "rmat": {"feature_steps": 1, "kappa": 0.5, "train": 96, "check": 5, "test": 0}}


def sbm_graph(seed, n=50_000, blocks=32, alpha=2.0):
    s, t, _ = graphs.sbm(n, blocks, alpha, seed)
    return graphs.edges_to_graph(s, t)


def rmat_graph(scale, edges, seed):
    """R-MAT graph with UGraph semantics through the C++ host loader (ids in first-seen order)."""
    from .host import HostGraph
    s, t = graphs.rmat(scale, edges, seed=seed)
    hg = HostGraph.from_edges(s, t)
    return hg.arrays()


def rmat_training_graph(k, seed0=1000):
    sc = 14 + (k % 4)
    return rmat_graph(sc, 8 << sc, seed0 + k)


def start_positions(g, seed):
    return graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), seed)


def fa2_targets(gs, iters, params=None, chunk_nodes=4_000_000, seed0=0, log=None):
    """Plain FA2 layouts after `iters` iterations, many graphs per batched launch chain."""
    params = FA2 if params is None else params
    out = [None] * len(gs)
    k = 0
    while k < len(gs):
        j, total = k, 0
        while j < len(gs) and (j == k or total + gs[j].n <= chunk_nodes):
            total += gs[j].n
            j += 1
        be = BatchEngine([(g.n, g.src, g.dst) for g in gs[k:j]], **params)
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


def size_constant(gs, targets, k_r):
    """Mean over the training targets of rms radius / sqrt(k_r * E)."""
    c = [np.sqrt(((p - p.mean(0)) ** 2).sum(1).mean()) / np.sqrt(k_r * g.m) for g, p in zip(gs, targets)]
    return float(np.mean(c))


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


def plain_run(e, start, iters_plain):
    """Plain FA2 from `start`: wall time of the iterations and the final normalized stress."""
    e.set_positions(start)
    e.reset()
    e.sync()
    t0 = time.perf_counter()
    e.step(iters_plain)
    e.sync()
    t_plain = time.perf_counter() - t0
    return {"plain_iterations": iters_plain, "plain_seconds": t_plain, "ns_plain": e.normalized_stress()[0],
            "ns_random_start": e.normalized_stress(start)[0]}


def learned_run(m, e, plain, start_scale, feature_seed, feature_steps, check):
    """Learned start + FA2 on the same engine until NS <= the plain run's (see module doc)."""
    ns_plain, iters_plain = plain["ns_plain"], plain["plain_iterations"]
    e.reset()
    gid = m.add_engine_graph(e, feature_seed=feature_seed, steps=feature_steps)
    m.infer_to_engine(gid, start_scale, e)
    m.release_graph(gid)
    trajectory = [e.normalized_stress()[0]]
    kstar = None
    it = 0
    while True:
        if trajectory[-1] <= ns_plain:
            kstar = it
            break
        if it + check > iters_plain:
            break
        e.step(check)
        it += check
        trajectory.append(e.normalized_stress()[0])
    r = dict(plain)
    r.update({"ns_learned_start": trajectory[0], "start_scale": start_scale, "feature_steps": feature_steps,
              "learned_iterations": kstar, "learned_trajectory": trajectory, "check": check})
    if kstar is not None:
        e.reset()
        e.sync()
        t0 = time.perf_counter()
        gid2 = m.add_engine_graph(e, feature_seed=feature_seed, steps=feature_steps)
        m.infer_to_engine(gid2, start_scale, e)
        e.step(kstar)
        e.sync()
        t_learned = time.perf_counter() - t0
        m.release_graph(gid2)
        r["learned_seconds"] = t_learned
        r["ns_learned"] = e.normalized_stress()[0]
        r["time_ratio"] = t_learned / plain["plain_seconds"]
    return r


def multistart_run(m, e, plain, start_scale, feature_seeds, feature_steps, per_start, check, random_seeds=None):
    """Several starts in turn on the same engine, each refined by FA2 for up to `per_start`
    iterations with a stress check every `check` iterations, until one reaches NS <= the
    plain run's NS or the plain run's iteration budget is spent. The starts are the network's
    proposals for the given feature seeds (the features' random-walk fields differ per seed),
    or, with random_seeds, uniform random starts (the same protocol without the network).
    Timed as run: every start, inference, FA2 iteration and stress check is inside the
    measured wall time."""
    ns_plain, budget = plain["ns_plain"], plain["plain_iterations"]
    tries = random_seeds if random_seeds is not None else feature_seeds
    e.sync()
    t0 = time.perf_counter()
    spent, reached, used, ns = 0, None, 0, None
    for j, sd in enumerate(tries):
        if spent >= budget:
            break
        e.reset()
        if random_seeds is not None:
            e.set_positions(start_positions_n(e.n, sd))
        else:
            gid = m.add_engine_graph(e, feature_seed=sd, steps=feature_steps)
            m.infer_to_engine(gid, start_scale, e)
            m.release_graph(gid)
        used = j + 1
        it = 0
        ns = e.normalized_stress()[0]
        while ns > ns_plain and it < per_start and spent < budget:
            step = min(check, per_start - it, budget - spent)
            e.step(step)
            it += step
            spent += step
            ns = e.normalized_stress()[0]
        if ns <= ns_plain:
            reached = spent
            break
    e.sync()
    t = time.perf_counter() - t0
    out = {"starts_used": used, "iterations": spent, "reached": reached is not None, "seconds": t,
           "time_ratio": t / plain["plain_seconds"], "ns_end": ns}
    return out


def start_positions_n(n, seed):
    rng = np.random.default_rng(seed)
    side = 2 * np.sqrt(n)
    return ((rng.random((n, 2)) - 0.5) * side).astype(np.float32)


def save_model(m, path, meta):
    """Parameters (FP32 masters, .npy) and a JSON sidecar with the settings they were trained
    with and the training targets' mean RMS radius."""
    np.save(path, m.params())
    meta = dict(meta, target_scale=m.target_scale(), num_params=int(m.nparams))
    with open(path[:-4] + ".json" if path.endswith(".npy") else path + ".json", "w") as fh:
        json.dump(meta, fh, indent=1)


def load_model(path, device=0):
    """(Mpnn with the saved parameters, metadata dict)."""
    with open(path[:-4] + ".json" if path.endswith(".npy") else path + ".json") as fh:
        meta = json.load(fh)
    p = np.load(path, allow_pickle=False)
    m = Mpnn(seed=0, device=device)
    if len(p) != m.nparams:
        m.close()
        raise ValueError(f"{path}: {len(p)} parameters, the network has {m.nparams}")
    m.set_params(p)
    return m, meta


def evaluate_sbm(m, target_scale, seeds, kappa, feature_steps, check, iters_plain=1000, multistart=None, log=None):
    """Time to equal stress on held-out SBM graphs: plain FA2 (random start, `iters_plain`
    iterations) against the learned start (single start, the protocol of learned_run) and,
    with multistart = (starts, per_start), several learned proposals and, for comparison,
    as many random starts under the same protocol."""
    out = []
    for seed in seeds:
        g = sbm_graph(seed)
        e = Engine(g.n, g.src, g.dst, **FA2)
        e.set_graph_capture(True)
        e.stress_setup(1024, 977, seed=seed)
        plain = plain_run(e, start_positions(g, 10_000 + seed), iters_plain)
        r = learned_run(m, e, plain, kappa * target_scale, 10_000 + seed, feature_steps, check)
        r.pop("learned_trajectory", None)
        r.update({"seed": seed, "nodes": g.n, "edges": g.m})
        if multistart:
            starts, per_start = multistart
            fseeds = [10_000 + seed + 7919 * j for j in range(starts)]
            r["multistart_learned"] = multistart_run(m, e, plain, kappa * target_scale, fseeds, feature_steps,
                                                     per_start, check)
            rseeds = [20_000 + seed + 7919 * j for j in range(starts)]
            r["multistart_random"] = multistart_run(m, e, plain, None, None, feature_steps, per_start, check,
                                                    random_seeds=rseeds)
        e.close()
        out.append(r)
        if log:
            log(json.dumps({k: v for k, v in r.items()}))
    return out


def summarise_multistart(results, key):
    rs = [r[key] for r in results if key in r]
    ok = [x["time_ratio"] for x in rs if x["reached"]]
    return {"graphs": len(rs), "reached": len(ok), "within_half": sum(1 for x in ok if x <= 0.5),
            "median_time_ratio": float(np.median(ok)) if ok else None}


def rmat_engine(g):
    """Engine for an R-MAT test graph with its stress pairs, and the reference's glibc start."""
    e = Engine(g.n, g.src, g.dst, **FA2_RMAT)
    e.set_graph_capture(True)
    e.stress_setup(1024, 977, seed=0)
    side = 2.0 * np.sqrt(g.n)
    return e, graphs.glibc_random_positions(g.n, side, side, seed=1234)


def summarise(results):
    ratios = [r["time_ratio"] for r in results if "time_ratio" in r]
    return {"reached": len(ratios), "within_half": sum(1 for x in ratios if x <= 0.5),
            "median_time_ratio": float(np.median(ratios)) if ratios else None,
            "mean_time_ratio": float(np.mean(ratios)) if ratios else None}


def run_sbm(args, log):
    fs_list = [int(x) for x in args.sweep_feature_steps.split(",")] if args.validate else [args.feature_steps]
    kappas = [float(x) for x in args.sweep_kappa.split(",")] if args.validate else [args.kappa]
    t0 = time.perf_counter()
    train_gs = [sbm_graph(1000 + k, n=args.nodes) for k in range(args.train)]
    log(f"{len(train_gs)} training graphs in {time.perf_counter() - t0:.1f}s")
    t1 = time.perf_counter()
    targets = fa2_targets(train_gs, args.target_iters, log=log)
    log(f"targets in {time.perf_counter() - t1:.1f}s")
    models, train_seconds = {}, {}
    for fs in fs_list:
        t2 = time.perf_counter()
        models[fs] = train(train_gs, targets, args.steps, lr=args.lr, feature_steps=fs, log=log)
        train_seconds[fs] = time.perf_counter() - t2
    del train_gs, targets
    if args.save_params and len(models) == 1:
        fs = next(iter(models))
        save_model(models[fs], args.save_params,
                   {"workload": "sbm50k", "train_seeds": [1000, 1000 + args.train - 1], "nodes": args.nodes,
                    "target_iterations": args.target_iters, "train_steps": args.steps, "lr": args.lr,
                    "feature_steps": fs, "kappa": args.kappa, "check": args.check,
                    "command": "python -m gpugraphlayout_amd.learned " + " ".join(sys.argv[1:])})
        log(f"saved {args.save_params}")
    if args.test == 0:
        for m in models.values():
            m.close()
        return {"workload": "sbm50k", "train_graphs": args.train, "train_seconds": train_seconds}
    # validation graphs come from the training seed range past the graphs trained on
    seed0 = 2000 if args.validate else 5000
    results = []
    for k in range(args.test):
        seed = seed0 + k
        g = sbm_graph(seed, n=args.nodes)
        e = Engine(g.n, g.src, g.dst, **FA2)
        e.set_graph_capture(True)
        e.stress_setup(1024, 977, seed=seed)
        plain = plain_run(e, start_positions(g, 10_000 + seed), args.target_iters)
        for fs, m in models.items():
            for kappa in kappas:
                r = learned_run(m, e, plain, kappa * m.target_scale(), 10_000 + seed, fs, args.check)
                r.update({"seed": seed, "nodes": g.n, "edges": g.m, "kappa": kappa})
                results.append(r)
                log(json.dumps({q: v for q, v in r.items() if q != "learned_trajectory"}))
        e.close()
    for m in models.values():
        m.close()
    summary = {"workload": "sbm50k", "validate": bool(args.validate), "train_graphs": args.train,
               "test_graphs": args.test, "first_seed": seed0, "nodes": args.nodes,
               "target_iterations": args.target_iters, "train_steps": args.steps, "train_seconds": train_seconds,
               **summarise(results), "results": results}
    if args.validate:
        summary["by_setting"] = {f"{fs},{kappa}": summarise([r for r in results if r["feature_steps"] == fs
                                                             and r["kappa"] == kappa])
                                 for fs in fs_list for kappa in kappas}
    return summary


def run_rmat(args, log):
    fs_list = [int(x) for x in args.sweep_feature_steps.split(",")] if args.validate else [args.feature_steps]
    kappas = [float(x) for x in args.sweep_kappa.split(",")] if args.validate else [args.kappa]
    t0 = time.perf_counter()
    train_gs = [rmat_training_graph(k) for k in range(args.train)]
    log(f"{len(train_gs)} training graphs in {time.perf_counter() - t0:.1f}s")
    t1 = time.perf_counter()
    targets = fa2_targets(train_gs, args.target_iters, params=FA2_RMAT, log=log)
    log(f"targets in {time.perf_counter() - t1:.1f}s")
    k_r = FA2_RMAT["scale"]
    c = size_constant(train_gs, targets, k_r)
    log(f"size constant c = {c:.4f}")
    models, train_seconds = {}, {}
    for fs in fs_list:
        t2 = time.perf_counter()
        models[fs] = train(train_gs, targets, args.steps, lr=args.lr, feature_steps=fs, log=log)
        train_seconds[fs] = time.perf_counter() - t2
    del train_gs, targets
    if args.validate:    # the test's size and iteration budget, another seed
        tests = [("rmat-scale24-128M-edges-seed4", 24, 128_000_000, 4, 200)]
    else:
        tests = [("rmat-scale24-128M-edges-seed3", 24, 128_000_000, 3, 200),
                 ("rmat-scale20-10M-edges-seed1", 20, 10_000_000, 1, 500)]
    results = []
    for name, scale, edges, seed, iters in tests:
        g = rmat_graph(scale, edges, seed)
        log(f"{name}: V={g.n} E={g.m}")
        e, start = rmat_engine(g)
        plain = plain_run(e, start, iters)
        log(f"plain: {plain}")
        for fs, m in models.items():
            for kappa in kappas:
                r = learned_run(m, e, plain, float(kappa * c * np.sqrt(k_r * g.m)), 77, fs, args.check)
                r.update({"graph": name, "nodes": g.n, "edges": g.m, "kappa": kappa})
                results.append(r)
                log(json.dumps({q: v for q, v in r.items() if q != "learned_trajectory"}))
        e.close()
        del g
    for m in models.values():
        m.close()
    return {"workload": "rmat", "validate": bool(args.validate), "train_graphs": args.train,
            "target_iterations": args.target_iters, "train_steps": args.steps, "train_seconds": train_seconds,
            "size_constant": c, **summarise(results), "results": results}


def inference_timing(reps, log):
    """Wall time of features (from the engine's CSR) and of inference into the engine on
    R-MAT 2^24 with 128M edges (seed 3); the weights are untrained, which does not change the
    work."""
    g = rmat_graph(24, 128_000_000, 3)
    log(f"rmat-scale24-128M-edges-seed3: V={g.n} E={g.m}")
    e = Engine(g.n, g.src, g.dst, **FA2_RMAT)
    m = Mpnn(seed=1)
    out = []
    for rep in range(reps):
        e.sync()
        t0 = time.perf_counter()
        gid = m.add_engine_graph(e, feature_seed=77, steps=DEFAULTS["rmat"]["feature_steps"])
        t1 = time.perf_counter()
        m.infer_to_engine(gid, 1000.0, e)
        t2 = time.perf_counter()
        m.release_graph(gid)
        out.append({"features_ms": 1e3 * (t1 - t0), "inference_ms": 1e3 * (t2 - t1)})
        log(json.dumps(out[-1]))
    m.close()
    e.close()
    return {"workload": "rmat", "inference_timing": out, "nodes": g.n, "edges": g.m}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=sorted(DEFAULTS), default="sbm50k")
    ap.add_argument("--train", type=int, default=None, help="training graphs (seeds 1000..)")
    ap.add_argument("--test", type=int, default=None, help="held-out SBM graphs (seeds 5000..)")
    ap.add_argument("--nodes", type=int, default=50_000)
    ap.add_argument("--target-iters", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--feature-steps", type=int, default=None)
    ap.add_argument("--kappa", type=float, default=None)
    ap.add_argument("--check", type=int, default=None, help="iterations between stress checks")
    ap.add_argument("--validate", action="store_true", help="sweep settings on validation graphs, not the tests")
    ap.add_argument("--sweep-feature-steps", default="0,1,2")
    ap.add_argument("--sweep-kappa", default="0.25,0.5,1.0")
    ap.add_argument("--inference-timing", type=int, default=0,
                    help="rmat: time features + inference on R-MAT 2^24 this many times, then exit")
    ap.add_argument("--save-params", default=None, help="sbm50k: save the trained network here (.npy + .json)")
    ap.add_argument("--evaluate", default=None, help="sbm50k: evaluate a saved network on held-out graphs")
    ap.add_argument("--first-seed", type=int, default=5000, help="with --evaluate: first held-out seed")
    ap.add_argument("--starts", type=int, default=1, help="with --evaluate: learned proposals per graph (> 1: "
                                                          "also the multi-start protocol, and random starts under it)")
    ap.add_argument("--per-start", type=int, default=100, help="with --starts: FA2 iterations per proposal")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    d = DEFAULTS[args.workload]
    for key in ("feature_steps", "kappa", "train", "check", "test"):
        if getattr(args, key) is None:
            setattr(args, key, d[key])

    def log(msg):
        print(f"[learned] {msg}", file=sys.stderr, flush=True)

    if args.inference_timing:
        summary = inference_timing(args.inference_timing, log)
    elif args.evaluate:
        m, meta = load_model(args.evaluate)
        seeds = list(range(args.first_seed, args.first_seed + args.test))
        ms = (args.starts, args.per_start) if args.starts > 1 else None
        res = evaluate_sbm(m, meta["target_scale"], seeds, meta["kappa"], meta["feature_steps"], meta["check"],
                           multistart=ms, log=log)
        m.close()
        summary = {"workload": "sbm50k", "model": args.evaluate, "first_seed": args.first_seed, "test_graphs": len(seeds),
                   "single_start": summarise(res), "results": res}
        if ms:
            summary["multistart"] = {"starts": args.starts, "per_start": args.per_start,
                                     "learned": summarise_multistart(res, "multistart_learned"),
                                     "random": summarise_multistart(res, "multistart_random")}
    else:
        summary = run_sbm(args, log) if args.workload == "sbm50k" else run_rmat(args, log)
    print(json.dumps({k: v for k, v in summary.items() if k != "results"}))
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(summary, fh, indent=1)


if __name__ == "__main__":
    main()

"""The five seeded workloads the engine is measured and checked on, built the same way for the
benchmark and the tests.

  rmat20   R-MAT over 2^20 ids, 10,000,000 distinct undirected non-loop edges, (a,b,c,d) =
           (0.57,0.19,0.19,0.05), seed 1; weak gravity k_g = 1, k_r = 10; 500 iterations.
  ba10k    Barabasi-Albert n = 10,000, m = 5, seed 0; strong gravity k_g = 1, k_r = 2;
           500 iterations.
  batch4096  4,096 Barabasi-Albert graphs, n = round(exp(U[ln 100, ln 20000])), m ~ U{2..8},
           seed 2, laid out together; strong gravity k_g = 1, k_r = 2; 300 iterations.
  sbm50k   stochastic block model, 50,000 nodes, 32 blocks of Dirichlet(2) sizes, p_in =
           8/|B|, p_out = 0.5/V; weak gravity k_g = 1, k_r = 2 (the learned initialiser's set).
  rmat24   R-MAT over 2^24 ids, 128,000,000 edges, seed 3; weak gravity k_g = 1, k_r = 10;
           200 iterations.

Single graphs are loaded with the reference's edge-list rules through the C++ host loader
(ids in first-appearance order, isolated ids vanish, src/RPGraph.cpp:35-102) and start from
the reference's initial layout: glibc random() after srandom(1234) in a square of side
2*sqrt(V) (src/graph_viewer.cpp:49, src/RPGraphLayout.cpp:45-52).
"""
from __future__ import annotations

import numpy as np

from . import graphs

SINGLE = {
    "rmat20": {"iterations": 500, "params": dict(strong_gravity=0, gravity=1.0, scale=10.0),
               "config": {"workload": "rmat-scale20-10M-edges", "rmat_abcd": [0.57, 0.19, 0.19, 0.05], "seed": 1,
                          "gravity": "weak", "k_g": 1.0, "k_r": 10.0, "theta": 1.0}},
    "rmat24": {"iterations": 200, "params": dict(strong_gravity=0, gravity=1.0, scale=10.0),
               "config": {"workload": "rmat-scale24-128M-edges", "rmat_abcd": [0.57, 0.19, 0.19, 0.05], "seed": 3,
                          "gravity": "weak", "k_g": 1.0, "k_r": 10.0, "theta": 1.0}},
    "ba10k": {"iterations": 500, "params": dict(strong_gravity=1, gravity=1.0, scale=2.0),
              "config": {"workload": "barabasi-albert-10k-m5", "seed": 0, "gravity": "strong", "k_g": 1.0,
                         "k_r": 2.0, "theta": 1.0}},
}
BATCH_PARAMS = dict(strong_gravity=1, gravity=1.0, scale=2.0)
BATCH_ITERATIONS = 300
SBM_PARAMS = dict(strong_gravity=0, gravity=1.0, scale=2.0)


def host_graph(s, t):
    """Reference loader semantics through the C++ host layer (O(E), hash-based)."""
    from .host import HostGraph
    hg = HostGraph.from_edges(s, t)
    g = hg.arrays()
    hg.close()
    return g


def reference_start(n: int) -> np.ndarray:
    side = 2.0 * np.sqrt(n)
    return graphs.glibc_random_positions(n, side, side, seed=1234)


def single(name: str):
    """(graph, initial positions, engine parameters, config dict) of a single-graph workload."""
    if name == "rmat20":
        s, t = graphs.rmat(20, 10_000_000, seed=1)
    elif name == "rmat24":
        s, t = graphs.rmat(24, 128_000_000, seed=3)
    elif name == "ba10k":
        s, t = graphs.barabasi_albert(10_000, 5, 0)
    else:
        raise KeyError(name)
    g = host_graph(s, t)
    del s, t
    spec = SINGLE[name]
    cfg = dict(spec["config"], nodes=g.n, edges=g.m,
               init="glibc random() after srandom(1234), square side 2*sqrt(V)")
    return g, reference_start(g.n), dict(spec["params"]), cfg


def batch4096(count: int = 4096, seed: int = 2):
    """The batch: graph k has n = round(exp(U[ln 100, ln 20000])), m ~ U{2..8} (numpy
    default_rng(seed)), Barabasi-Albert seed seed*2^32 + k, and starts from i.i.d. uniform
    positions in a square of side 2*sqrt(n) (default_rng(seed*2^32 + k)). Returns a list of
    (graph, positions)."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(100.0), np.log(20000.0), count))).astype(np.int64)
    ms = rng.integers(2, 9, count)
    out = []
    for k in range(count):
        n, m = int(ns[k]), int(ms[k])
        s, t = graphs.barabasi_albert(n, m, (seed << 32) + k)
        g = host_graph(s, t)
        pos = graphs.uniform_positions(g.n, 2.0 * np.sqrt(g.n), (seed << 32) + k)
        out.append((g, pos))
    return out


def sbm50k(seed: int, n: int = 50_000, blocks: int = 32, alpha: float = 2.0):
    """One graph of the learned initialiser's set (training seeds 1000-1999, held-out
    5000-5099)."""
    s, t, _ = graphs.sbm(n, blocks, alpha, seed)
    return graphs.edges_to_graph(s, t)

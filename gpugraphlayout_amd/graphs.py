"""Graph inputs for the engine: seeded synthetic generators, the reference's edge-list
semantics, and its initial layout.

Edge lists follow the reference's UGraph loader (src/RPGraph.cpp:35-102): lines starting
with '#' are skipped, self-loops and repeated undirected edges are dropped, node ids are
numbered by first appearance (source before target on each line), and each undirected
edge is kept once at its smaller internal id, in file order. ``edges_to_graph`` applies
exactly those rules to an in-memory edge list; ``load_edgelist`` reads a file.

Generators draw from numpy's PCG64 seeded per call, so a (generator, parameters, seed)
triple always yields the same edge list in the same order.
"""
from __future__ import annotations

import ctypes
import ctypes.util
from dataclasses import dataclass

import numpy as np


@dataclass
class Graph:
    n: int                    # nodes (ids 0..n-1 in first-appearance order)
    src: np.ndarray           # COO, each undirected edge once, src < dst, reference order
    dst: np.ndarray
    ids: np.ndarray           # internal id -> id in the edge list

    @property
    def m(self) -> int:
        return int(len(self.src))

    def degrees(self) -> np.ndarray:
        return (np.bincount(self.src, minlength=self.n) + np.bincount(self.dst, minlength=self.n)).astype(np.int64)

    def mass(self) -> np.ndarray:
        """ForceAtlas2::mass: degree + 1 (src/RPForceAtlas2.cpp:74-77)."""
        return (self.degrees() + 1).astype(np.float32)

    def csr(self):
        """Symmetric CSR (rowptr int64, col int32)."""
        from . import coo_to_csr
        return coo_to_csr(self.n, self.src, self.dst)


def edges_to_graph(s: np.ndarray, t: np.ndarray) -> Graph:
    """UGraph semantics over an edge list given as two id arrays in file order."""
    s = np.asarray(s, dtype=np.int64)
    t = np.asarray(t, dtype=np.int64)
    keep = s != t
    s, t = s[keep], t[keep]
    if len(s) == 0:
        return Graph(0, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int64))
    # drop repeated undirected edges, keeping the first occurrence (src/RPGraph.cpp:52)
    lo, hi = np.minimum(s, t), np.maximum(s, t)
    base = int(hi.max()) + 1 if len(hi) else 1
    if int(lo.min()) >= 0 and base < (1 << 31):
        _, first = np.unique(lo * base + hi, return_index=True)
    else:
        _, first = np.unique(np.stack([lo, hi], 1), axis=0, return_index=True)
    first.sort()
    s, t = s[first], t[first]
    # ids by first appearance, source before target (src/RPGraph.cpp:78-87, :89-94)
    seq = np.stack([s, t], 1).ravel()
    uniq, idx = np.unique(seq, return_index=True)
    order = np.argsort(idx, kind="stable")
    ids = uniq[order]
    new_of_sorted = np.empty(len(uniq), np.int64)
    new_of_sorted[order] = np.arange(len(uniq))
    ms = new_of_sorted[np.searchsorted(uniq, s)]
    mt = new_of_sorted[np.searchsorted(uniq, t)]
    # adjacency at min id in insertion order, listed by ascending min id (src/RPGraph.cpp:97,
    # src/RPGPUForceAtlas2.cu:73-84)
    a, b = np.minimum(ms, mt), np.maximum(ms, mt)
    o = np.argsort(a, kind="stable")
    return Graph(len(ids), a[o].astype(np.int32), b[o].astype(np.int32), ids)


def load_edgelist(path: str) -> Graph:
    """Reads a whitespace-separated edge list with the reference's rules."""
    s, t = [], []
    with open(path) as fh:
        for line in fh:
            if not line or line[0] == "#":
                continue
            parts = line.split()
            if len(parts) < 2:
                continue
            s.append(int(parts[0]))
            t.append(int(parts[1]))
    return edges_to_graph(np.array(s, np.int64), np.array(t, np.int64))


def write_edgelist(path: str, s, t):
    with open(path, "w") as fh:
        for a, b in zip(np.asarray(s).tolist(), np.asarray(t).tolist()):
            fh.write(f"{a} {b}\n")


def barabasi_albert(n: int, m: int, seed: int):
    """Preferential attachment from a star K_{1,m}: nodes m+1..n-1 each attach to m
    distinct earlier nodes chosen with probability proportional to degree (the C++
    generator of the host library, seeded splitmix64). Returns (s, t) in creation order;
    m*(n-m-1)+m edges."""
    from .host import lib as host_lib
    count = m * (n - m - 1) + m
    s = np.zeros(max(count, 0), np.int32)
    t = np.zeros(max(count, 0), np.int32)
    got = host_lib().gfa2host_barabasi_albert(int(n), int(m), int(seed) & (2 ** 64 - 1),
                                              s.ctypes.data_as(ctypes.c_void_p), t.ctypes.data_as(ctypes.c_void_p))
    return s[:got].astype(np.int64), t[:got].astype(np.int64)


def rmat(scale: int, n_edges: int, a=0.57, b=0.19, c=0.19, seed: int = 1):
    """R-MAT edges over 2^scale ids, drawn until n_edges distinct undirected non-loop
    edges exist; returned in first-occurrence order with their drawn orientation (the C++
    generator of the host library: counter-based SplitMix64, one double per level)."""
    from .host import lib as host_lib
    s = np.zeros(n_edges, np.int32)
    t = np.zeros(n_edges, np.int32)
    got = host_lib().gfa2host_rmat(int(scale), int(n_edges), float(a), float(b), float(c),
                                   int(seed) & (2 ** 64 - 1), s.ctypes.data_as(ctypes.c_void_p),
                                   t.ctypes.data_as(ctypes.c_void_p))
    if got < 0:
        raise ValueError("invalid R-MAT arguments")
    return s[:got].astype(np.int64), t[:got].astype(np.int64)


def sbm(n: int, blocks: int, alpha: float, seed: int):
    """Stochastic block model: block sizes ~ Dirichlet(alpha) (at least 2 nodes each),
    p_in = 8/|B| within a block, p_out = 0.5/n between blocks. Edges are sampled as
    Bernoulli counts per block pair followed by uniform endpoint draws (duplicates are
    dropped by the loader rules). Returns (s, t, block_of_node)."""
    rng = np.random.default_rng(seed)
    w = rng.dirichlet([alpha] * blocks)
    sizes = np.maximum(2, np.floor(w * n).astype(np.int64))
    sizes[-1] = max(2, n - sizes[:-1].sum())
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    block = np.repeat(np.arange(blocks), sizes)
    ss, tt = [], []
    for i in range(blocks):
        for j in range(i, blocks):
            if i == j:
                pairs = sizes[i] * (sizes[i] - 1) // 2
                p = min(1.0, 8.0 / sizes[i])
            else:
                pairs = sizes[i] * sizes[j]
                p = 0.5 / total
            k = rng.binomial(pairs, p)
            if k == 0:
                continue
            a = starts[i] + rng.integers(sizes[i], size=k)
            b = starts[j] + rng.integers(sizes[j], size=k)
            ss.append(a)
            tt.append(b)
    s = np.concatenate(ss)
    t = np.concatenate(tt)
    perm = rng.permutation(len(s))
    return s[perm], t[perm], block


_libc = None


def glibc_random_positions(n: int, width: float, height: float, seed: int = 1234) -> np.ndarray:
    """The reference's initial layout: srandom(seed) (src/graph_viewer.cpp:49), then per
    node x then y = lb + (ub-lb)*float(random())/float(RAND_MAX) over [-w/2, w/2] and
    [-h/2, h/2] (src/RPGraphLayout.cpp:45-52, src/RPCommon.cpp:50-53), in float."""
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL(ctypes.util.find_library("c"))
        _libc.random.restype = ctypes.c_long
        _libc.srandom.argtypes = [ctypes.c_uint]
    _libc.srandom(seed)
    r = np.array([_libc.random() for _ in range(2 * n)], dtype=np.int64).reshape(n, 2)
    RAND_MAX = np.float32(2147483647)
    lb = np.array([-width / 2.0, -height / 2.0], np.float64).astype(np.float32)
    ub = np.array([width / 2.0, height / 2.0], np.float64).astype(np.float32)
    # ((ub - lb) * float(r)) / float(RAND_MAX), then + lb, each step in float
    return (lb + ((ub - lb) * r.astype(np.float32)) / RAND_MAX).astype(np.float32)


def uniform_positions(n: int, side: float, seed: int) -> np.ndarray:
    """i.i.d. uniform positions in a square of the given side centred at the origin."""
    rng = np.random.default_rng(seed)
    return ((rng.random((n, 2)) - 0.5) * side).astype(np.float32)

"""A batch of graphs laid out together against each graph laid out alone (bit for bit, trees
and trajectories) and against the oracle."""
import numpy as np
import pytest

from gpugraphlayout_amd import BatchEngine, Engine, GFA2_FORCE_REPULSION, graphs

pytestmark = pytest.mark.gpu


def batch_case(seed, big=False):
    out = []
    sizes = [(300, 3), (50, 2), (1000, 5), (2, 1), (120, 4), (7, 1)]
    if big:
        # one graph over 32,768 nodes: the batch sorts globally, then by graph
        sizes.insert(2, (40000, 2))
    for k, (n, m) in enumerate(sizes):
        s, t = graphs.barabasi_albert(n, m, seed + k)
        g = graphs.edges_to_graph(s, t)
        out.append(g)
    s, t = graphs.rmat(11, 6000, seed=seed)
    out.append(graphs.edges_to_graph(s, t))
    return out


@pytest.mark.parametrize("group", [32, 64])
def test_batch_equals_single_runs(group):
    print(f"case batch G={group}", flush=True)
    gs = batch_case(3)
    pos = [graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 40 + k) for k, g in enumerate(gs)]
    params = dict(scale=2.0, gravity=1.0, strong_gravity=1, group_size=group)
    be = BatchEngine([(g.n, g.src, g.dst) for g in gs], **params)
    be.set_positions(np.concatenate(pos))
    be.step(25)
    be.sync()
    got = be.split(be.positions())
    st = be.batch_stats()
    for k, g in enumerate(gs):
        e = Engine(g.n, g.src, g.dst, **params)
        e.set_positions(pos[k])
        e.step(25)
        e.sync()
        np.testing.assert_array_equal(got[k].view(np.uint32), e.positions().view(np.uint32))
        s1 = e.stats()
        assert st["speed"][k] == np.float32(s1["global_speed"])
        assert st["swing"][k] == s1["swing"]
        e.close()
    be.close()


def test_batch_with_a_graph_over_the_per_graph_sort_limit():
    """A 40,000-node graph sends the batch through the global sort and then the graph sort.
    Every graph still follows its single run bit for bit: the cell moments are exact integer
    sums, so nothing depends on the graphs sorted before it or on where the tree tiles fall."""
    print("case batch with a 40,000-node graph", flush=True)
    gs = batch_case(3, big=True)
    pos = [graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 40 + k) for k, g in enumerate(gs)]
    params = dict(scale=2.0, gravity=1.0, strong_gravity=1, group_size=32)
    be = BatchEngine([(g.n, g.src, g.dst) for g in gs], **params)
    be.set_positions(np.concatenate(pos))
    be.step(12)
    be.sync()
    got = be.split(be.positions())
    st = be.batch_stats()
    for k, g in enumerate(gs):
        e = Engine(g.n, g.src, g.dst, **params)
        e.set_positions(pos[k])
        e.step(12)
        e.sync()
        np.testing.assert_array_equal(got[k].view(np.uint32), e.positions().view(np.uint32), err_msg=f"graph {k}")
        s1 = e.stats()
        assert st["swing"][k] == s1["swing"]
        assert int(st["interactions"][k]) == s1["interactions"]
        e.close()
    be.close()


def test_batch_graphs_against_oracle(oracle):
    """Each graph of a batch against the oracle directly: tree bit-exact (levels, counts,
    masses, permutation), centroids to float rounding, repulsion within 1e-4 relative L2."""
    print("case batch vs oracle", flush=True)
    gs = batch_case(8, big=True)
    pos = [graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 60 + k) for k, g in enumerate(gs)]
    be = BatchEngine([(g.n, g.src, g.dst) for g in gs], scale=2.0)
    be.set_capture(True)
    be.set_positions(np.concatenate(pos))
    be.step(1)
    be.sync()
    T = be.tree()
    rep = be.split(be.forces(GFA2_FORCE_REPULSION))
    off = np.concatenate([[0], np.cumsum([g.n for g in gs])])
    roots = np.append(np.flatnonzero(T["level"] == 0), len(T["level"]))
    assert len(roots) == len(gs) + 1
    for k, g in enumerate(gs):
        ref = oracle.tree(pos[k], g.mass())
        c0, c1 = roots[k], roots[k + 1]
        np.testing.assert_array_equal(T["level"][c0:c1], ref["level"], err_msg=f"graph {k}")
        np.testing.assert_array_equal(T["count"][c0:c1], ref["count"], err_msg=f"graph {k}")
        np.testing.assert_array_equal(T["mass"][c0:c1].view(np.uint32), ref["mass"].view(np.uint32))
        np.testing.assert_array_equal(T["perm"][off[k]:off[k + 1]] - off[k], ref["perm"])
        scale = max(float(ref["root"][2]) * 2, 1e-30)
        assert np.max(np.abs(T["xy"][c0:c1] - ref["xy"])) <= 1e-6 * scale
        want, _ = oracle.bh_repulsion(pos[k], g.mass(), 2.0, group=32)
        den = np.linalg.norm(want)
        err = np.linalg.norm(rep[k].astype(np.float64) - want) / (den if den > 0 else 1.0)
        assert err <= 1e-4, (k, err)
    be.close()


def test_batch_with_single_node_graph():
    print("case batch single node", flush=True)
    g1 = graphs.edges_to_graph(*graphs.barabasi_albert(80, 2, 1))
    be = BatchEngine([(1, np.zeros(0, np.int32), np.zeros(0, np.int32)), (g1.n, g1.src, g1.dst)], scale=2.0)
    pos = np.concatenate([np.array([[3.0, 4.0]], np.float32), graphs.uniform_positions(g1.n, 18, 2)])
    be.set_positions(pos)
    be.step(10)
    be.sync()
    p = be.positions()
    assert np.all(np.isfinite(p))
    # a lone node only feels gravity: it moves straight toward the origin
    assert abs(p[0, 0] / p[0, 1] - 0.75) < 1e-5 and np.linalg.norm(p[0]) < 5.0
    e = Engine(g1.n, g1.src, g1.dst, scale=2.0)
    e.set_positions(pos[1:])
    e.step(10)
    e.sync()
    np.testing.assert_array_equal(p[1:].view(np.uint32), e.positions().view(np.uint32))


def test_batch_tree_matches_single_trees_across_tile_boundaries():
    """Graphs of 2,047..12,000 nodes (one to six sort tiles each, sorted one block per graph):
    the in-order permutation and every cell's level, body count, mass and centroid equal the
    graph's own tree, bit for bit."""
    print("case batch tree across tile boundaries", flush=True)
    sizes = [(2047, 2), (40, 2), (2048, 3), (2049, 2), (3, 1), (5000, 4), (12000, 2)]
    gs = [graphs.edges_to_graph(*graphs.barabasi_albert(n, m, 11 + k)) for k, (n, m) in enumerate(sizes)]
    pos = [graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 70 + k) for k, g in enumerate(gs)]
    be = BatchEngine([(g.n, g.src, g.dst) for g in gs], scale=2.0)
    be.set_positions(np.concatenate(pos))
    be.step(1)
    be.sync()
    T = be.tree()
    off = np.concatenate([[0], np.cumsum([g.n for g in gs])])
    c0 = 0
    for k, g in enumerate(gs):
        e = Engine(g.n, g.src, g.dst, scale=2.0)
        e.set_positions(pos[k])
        e.step(1)
        e.sync()
        S = e.tree()
        np.testing.assert_array_equal(T["perm"][off[k]:off[k + 1]] - off[k], S["perm"])
        c1 = c0 + len(S["level"])
        for key in ("level", "count", "mass"):
            np.testing.assert_array_equal(T[key][c0:c1], S[key], err_msg=f"graph {k} {key}")
        np.testing.assert_array_equal(T["xy"][c0:c1].view(np.uint32), S["xy"].view(np.uint32),
                                      err_msg=f"graph {k} centroids")
        assert S["level"][0] == 0 and T["level"][c0] == 0
        c0 = c1
        e.close()
    assert c0 == len(T["level"])
    be.close()


def test_batch_graph_replay_matches_eager():
    """The batch's iteration captured as one hipGraph and replayed gives the eager result."""
    print("case batch hipGraph replay", flush=True)
    gs = batch_case(5)
    pos = np.concatenate([graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 90 + k) for k, g in enumerate(gs)])
    out = []
    for capture in (False, True):
        be = BatchEngine([(g.n, g.src, g.dst) for g in gs], scale=2.0)
        be.set_graph_capture(capture)
        be.set_positions(pos)
        for _ in range(3):
            be.step(4)
            be.sync()
        out.append((be.positions(), be.batch_stats()))
        be.close()
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1]["speed"], out[1][1]["speed"])

"""Engine options beyond the reference's GPU path, against the oracle: exact O(V^2)
repulsion (the reference's GPU backend has none), linlog attraction (the reference's CPU
rule) and edge weights w^delta (declared, never read, in the reference)."""
import numpy as np
import pytest

from gpugraphlayout_amd import BatchEngine, Engine, GFA2_FORCE_ATTRACTION_GRAVITY, GFA2_FORCE_REPULSION, graphs

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case(seed, n=2000, m=4):
    s, t = graphs.barabasi_albert(n, m, seed)
    g = graphs.edges_to_graph(s, t)
    pos = graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), seed + 5)
    return g, pos


def test_exact_repulsion_fp64_matches_oracle(oracle):
    print("case exact fp64", flush=True)
    g, pos = case(1)
    eng = Engine(g.n, g.src, g.dst, scale=2.0)
    eng.set_positions(pos)
    got = eng.exact_repulsion(fp64=True)
    want = oracle.exact_repulsion_f64(pos, g.mass(), 2.0, np.arange(g.n))
    assert rel_l2(got, want) <= 1e-12
    got32 = eng.exact_repulsion(fp64=False)
    assert rel_l2(got32, want) <= 1e-5
    eng.close()


def test_exact_mode_iteration_matches_oracle(oracle):
    print("case exact mode step", flush=True)
    g, pos = case(2)
    eng = Engine(g.n, g.src, g.dst, scale=2.0, use_barneshut=0)
    eng.set_capture(True)
    eng.set_positions(pos)
    eng.step(1)
    eng.sync()
    rep = eng.forces(GFA2_FORCE_REPULSION)
    want = oracle.exact_repulsion_f64(pos, g.mass(), 2.0, np.arange(g.n))
    assert rel_l2(rep, want) <= 1e-5
    st = eng.stats()
    assert st["interactions"] == g.n * g.n
    # a few more iterations stay finite and deterministic
    eng.step(10)
    eng.sync()
    a = eng.positions()
    eng.close()
    eng = Engine(g.n, g.src, g.dst, scale=2.0, use_barneshut=0)
    eng.set_positions(pos)
    eng.step(11)
    eng.sync()
    np.testing.assert_array_equal(a.view(np.uint32), eng.positions().view(np.uint32))
    eng.close()


def test_barnes_hut_tends_to_exact_as_theta_falls():
    print("case theta -> 0", flush=True)
    g, pos = case(3, n=1500)
    eng = Engine(g.n, g.src, g.dst, scale=2.0, theta=1e-3)
    eng.set_capture(True)
    eng.set_positions(pos)
    eng.step(1)
    eng.sync()
    rep = eng.forces(GFA2_FORCE_REPULSION)
    eng.close()
    ex = Engine(g.n, g.src, g.dst, scale=2.0)
    ex.set_positions(pos)
    want = ex.exact_repulsion(fp64=True)
    ex.close()
    assert rel_l2(rep, want) <= 1e-5


def test_exact_batch_equals_single_runs():
    print("case exact batch", flush=True)
    items = [case(10 + k, n=n) for k, n in enumerate((300, 57, 1200))]
    be = BatchEngine([(g.n, g.src, g.dst) for g, _ in items], scale=2.0, use_barneshut=0)
    be.set_positions(np.concatenate([p for _, p in items]))
    be.step(5)
    be.sync()
    parts = be.split(be.positions())
    be.close()
    for (g, pos), got in zip(items, parts):
        e = Engine(g.n, g.src, g.dst, scale=2.0, use_barneshut=0)
        e.set_positions(pos)
        e.step(5)
        e.sync()
        np.testing.assert_array_equal(got.view(np.uint32), e.positions().view(np.uint32))
        e.close()


@pytest.mark.parametrize("linlog,delta,weighted", [(1, 0.0, False), (0, 1.0, True), (0, 0.5, True), (1, 2.0, True)])
def test_attraction_options_match_oracle(oracle, linlog, delta, weighted):
    print(f"case attraction linlog={linlog} delta={delta} weighted={weighted}", flush=True)
    g, pos = case(4)
    rng = np.random.default_rng(9)
    w = rng.uniform(0.25, 4.0, g.m).astype(np.float32) if weighted else None
    eng = Engine(g.n, g.src, g.dst, weights=w, scale=2.0, linlog=linlog, edge_weight_influence=delta)
    eng.set_positions(pos)
    eng.step(1)
    eng.sync()
    fga = eng.forces(GFA2_FORCE_ATTRACTION_GRAVITY)
    want = oracle.gravity(pos, g.mass(), False, 1.0)
    want = oracle.attraction_general(pos, g.src, g.dst, w=w, delta=delta, linlog=bool(linlog), f=want)
    assert rel_l2(fga, want) <= 1e-5
    # changing delta afterwards re-derives the weights
    if weighted:
        eng.set_params(edge_weight_influence=1.0)
        eng.set_positions(pos)
        eng.reset()
        eng.step(1)
        eng.sync()
        want = oracle.attraction_general(pos, g.src, g.dst, w=w, delta=1.0, linlog=bool(linlog),
                                         f=oracle.gravity(pos, g.mass(), False, 1.0))
        assert rel_l2(eng.forces(GFA2_FORCE_ATTRACTION_GRAVITY), want) <= 1e-5
    eng.close()


@pytest.mark.parametrize("linlog,weighted", [(0, False), (1, True)])
def test_attraction_on_hub_rows_matches_oracle(oracle, linlog, weighted):
    """Rows over 2,048 entries (attraction in 4,096-entry chunks, partials added in chunk
    order) against the oracle, with and without linlog and weights; the hubs' own forces
    are checked one by one."""
    print(f"case attraction hub rows linlog={linlog} weighted={weighted}", flush=True)
    n = 6000
    s, t = graphs.barabasi_albert(n, 2, 8)
    hs = np.concatenate([np.full(n, n), np.full(n + 1, n + 1)])
    ht = np.concatenate([np.arange(n), np.arange(n + 1)])
    g = graphs.edges_to_graph(np.concatenate([s, hs]), np.concatenate([t, ht]))
    deg = np.diff(g.csr()[0])
    hubs = np.nonzero(deg > 4096)[0]
    assert len(hubs) == 2
    pos = graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 12)
    rng = np.random.default_rng(10)
    w = rng.uniform(0.25, 4.0, g.m).astype(np.float32) if weighted else None
    eng = Engine(g.n, g.src, g.dst, weights=w, scale=2.0, linlog=linlog, edge_weight_influence=1.0)
    eng.set_positions(pos)
    eng.step(1)
    eng.sync()
    fga = eng.forces(GFA2_FORCE_ATTRACTION_GRAVITY)
    want = oracle.gravity(pos, g.mass(), False, 1.0)
    want = oracle.attraction_general(pos, g.src, g.dst, w=w, delta=1.0, linlog=bool(linlog), f=want)
    assert rel_l2(fga, want) <= 1e-5
    for h in hubs:
        assert rel_l2(fga[h], want[h]) <= 1e-5
    eng.close()

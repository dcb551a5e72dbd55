"""The learned initialiser on the device against its FP64 oracle, and what it gains."""
import numpy as np
import pytest

from gpugraphlayout_amd import Engine, GFA2Error, Mpnn, graphs, learned
from oracle import mpnn_oracle as mo

pytestmark = pytest.mark.gpu


def graph_case(seed, n=600):
    s, t, _ = graphs.sbm(n, 6, 2.0, seed)
    g = graphs.edges_to_graph(s, t)
    rowptr, col = g.csr()
    return g, rowptr, col


def hub_case(seed, n=1500):
    """An SBM graph plus two hubs adjacent to every other node: rows of about n entries take
    the chunked aggregation path (rows over 512 entries, chunks of 512)."""
    s, t, _ = graphs.sbm(n, 6, 2.0, seed)
    hubs_s = np.concatenate([np.full(n, n), np.full(n + 1, n + 1)])
    hubs_t = np.concatenate([np.arange(n), np.arange(n + 1)])
    g = graphs.edges_to_graph(np.concatenate([s, hubs_s]), np.concatenate([t, hubs_t]))
    rowptr, col = g.csr()
    assert np.diff(rowptr).max() > 2 * 512
    return g, rowptr, col


def normalised(T):
    T = T.astype(np.float64) - T.mean(0)
    return T / np.sqrt((T ** 2).sum(1).mean())


def test_features_match_oracle():
    print("case mpnn features", flush=True)
    g, rowptr, col = graph_case(1)
    m = Mpnn(seed=0)
    gid = m.add_graph(g.n, rowptr, col, feature_seed=1234, steps=8)
    X = m.features(gid)
    ref = mo.features(rowptr, col, 1234, 8)
    assert np.max(np.abs(X - ref)) < 1e-4
    m.close()


def test_hub_rows_match_oracle():
    """Features, loss and gradient on a graph with hub rows (chunked aggregation)."""
    print("case mpnn hub rows", flush=True)
    g, rowptr, col = hub_case(3)
    X_ref = mo.features(rowptr, col, 4321, 3)
    rng = np.random.default_rng(6)
    T = rng.normal(size=(g.n, 2)).astype(np.float32)
    m = Mpnn(seed=5)
    gid = m.add_graph(g.n, rowptr, col, target=T, feature_seed=4321, steps=3)
    X = m.features(gid)
    assert np.max(np.abs(X - X_ref)) < 1e-4
    loss, grad = m.loss_grad(gid, [5, 77])
    L, G = mo.loss_grad(m.params().astype(np.float64), rowptr, col, X, normalised(T), [5, 77], round_bf16=True)
    assert loss == pytest.approx(L, rel=1e-5)
    assert np.linalg.norm(grad - G) / np.linalg.norm(G) < 1e-4
    m.close()


def test_engine_graph_features_match_host_graph_features():
    """Features from an engine's device CSR equal those from the same CSR uploaded from the
    host, bit for bit."""
    print("case mpnn engine-graph features", flush=True)
    g, rowptr, col = graph_case(4)
    e = Engine(g.n, rowptr=rowptr, col=col)
    m = Mpnn(seed=0)
    a = m.add_graph(g.n, rowptr, col, feature_seed=77, steps=1)
    b = m.add_engine_graph(e, feature_seed=77, steps=1)
    assert np.array_equal(m.features(a), m.features(b))
    assert np.array_equal(m.infer(a, 5.0), m.infer(b, 5.0))
    m.release_graph(b)
    with pytest.raises(GFA2Error):
        m.features(b)
    assert np.array_equal(m.features(a), m.features(m.add_engine_graph(e, feature_seed=77, steps=1)))
    e.close()
    m.close()


def test_loss_and_gradient_match_oracle():
    print("case mpnn gradient", flush=True)
    g, rowptr, col = graph_case(2)
    rng = np.random.default_rng(5)
    T = rng.normal(size=(g.n, 2)).astype(np.float32) * 7 + 3
    m = Mpnn(seed=3)
    gid = m.add_graph(g.n, rowptr, col, target=T, feature_seed=9, steps=4)
    p = m.params()
    offs = [17, 101, 250]
    loss, grad = m.loss_grad(gid, offs)
    X = m.features(gid)
    L, G = mo.loss_grad(p.astype(np.float64), rowptr, col, X, normalised(T), offs, round_bf16=True)
    assert loss == pytest.approx(L, rel=1e-5)
    assert np.linalg.norm(grad - G) / np.linalg.norm(G) < 1e-4
    m.close()


def test_training_reduces_loss_and_is_deterministic():
    print("case mpnn training", flush=True)
    losses, params = [], []
    for rep in range(2):
        m = Mpnn(seed=11)
        gids = []
        for k in range(4):
            g, rowptr, col = graph_case(20 + k)
            pos = graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 30 + k)
            e = Engine(g.n, g.src, g.dst, scale=2.0)
            e.set_positions(pos)
            e.step(150)
            e.sync()
            gids.append(m.add_graph(g.n, rowptr, col, target=e.positions(), feature_seed=k, steps=8))
            e.close()
        first, _ = m.loss_grad(gids[0], [5, 77])
        m.train(300, lr=3e-3, R=8, seed=1)
        last, _ = m.loss_grad(gids[0], [5, 77])
        losses.append((first, last))
        params.append(m.params())
        m.close()
    assert losses[0][1] < 0.7 * losses[0][0]
    np.testing.assert_array_equal(params[0].view(np.uint32), params[1].view(np.uint32))


def test_infer_to_engine_places_positions():
    print("case mpnn infer", flush=True)
    g, rowptr, col = graph_case(40)
    m = Mpnn(seed=2)
    gid = m.add_graph(g.n, rowptr, col, feature_seed=3)
    xy = m.infer(gid, 25.0)
    assert np.allclose(xy.mean(0), 0, atol=1e-3)
    assert np.sqrt((xy.astype(np.float64) ** 2).sum(1).mean()) == pytest.approx(25.0, rel=1e-4)
    e = Engine(g.n, g.src, g.dst, scale=2.0)
    m.infer_to_engine(gid, 25.0, e)
    np.testing.assert_array_equal(e.positions().view(np.uint32), xy.view(np.uint32))
    e.step(3)
    e.sync()
    assert np.all(np.isfinite(e.positions()))
    e.close()
    m.close()


def test_learned_start_beats_random_start():
    """Trained on FA2 layouts of 64 seeded SBM graphs, the network's start has lower
    normalized stress than the random start on held-out graphs, on average and on most
    graphs. Whether FA2 from it then reaches the plain run's final stress sooner is reported
    (printed), not asserted: on these graphs plain FA2 plateaus within ~50 iterations and
    runs settle in basins whose stress differs by a few 1e-4 either way."""
    print("case mpnn gain", flush=True)
    n, it_plain, fsteps = 1500, 300, 2
    train = []
    for k in range(64):
        s, t, _ = graphs.sbm(n, 8, 2.0, 1000 + k)
        train.append(graphs.edges_to_graph(s, t))
    targets = learned.fa2_targets(train, it_plain)
    m = Mpnn(seed=0)
    for k, (g, target) in enumerate(zip(train, targets)):
        rowptr, col = g.csr()
        m.add_graph(g.n, rowptr, col, target=target, feature_seed=k, steps=fsteps)
    m.train(1500, lr=3e-3, R=8, seed=2)
    scale = m.target_scale()
    rand_ns, learned_ns = [], []
    for k in range(8):
        s, t, _ = graphs.sbm(n, 8, 2.0, 5000 + k)
        g = graphs.edges_to_graph(s, t)
        rowptr, col = g.csr()
        rand_start = graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 100 + k)
        plain = Engine(g.n, g.src, g.dst, **learned.FA2)
        plain.stress_setup(200, 100, seed=k)
        ns_random_start, _ = plain.normalized_stress(rand_start)
        plain.set_positions(rand_start)
        plain.step(it_plain)
        plain.sync()
        ns_plain, _ = plain.normalized_stress()
        plain.close()
        gid = m.add_graph(g.n, rowptr, col, feature_seed=100 + k, steps=fsteps)
        e = Engine(g.n, g.src, g.dst, **learned.FA2)
        m.infer_to_engine(gid, scale, e)
        e.stress_setup(200, 100, seed=k)
        ns_learned_start, _ = e.normalized_stress()
        done = None
        for it in range(0, it_plain + 1, 10):
            if it:
                e.step(10)
            if e.normalized_stress()[0] <= ns_plain:
                done = it
                break
        e.close()
        print(f"held-out {k}: NS random start {ns_random_start:.4f}, learned start {ns_learned_start:.4f}; "
              f"plain after {it_plain} it {ns_plain:.4f}, learned run reaches it at {done}", flush=True)
        rand_ns.append(ns_random_start)
        learned_ns.append(ns_learned_start)
    m.close()
    assert np.mean(learned_ns) < np.mean(rand_ns) - 0.02
    assert sum(a < b for a, b in zip(learned_ns, rand_ns)) >= 6

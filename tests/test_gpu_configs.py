"""Every benchmark workload on the engine against the oracle.

  ba10k     (BA 10k/m5, strong gravity, k_r = 2): at iterations 0, 100 and 499 the tree
            bit-exact, Barnes-Hut repulsion within 1e-4 relative L2 of the oracle's walk on
            the same positions, gravity + attraction within 1e-5; every speed update of the
            500-iteration run equal to the oracle's update from the same S and T.
  batch4096 (4,096 BA graphs in one batch): 300 iterations with no error flag and finite
            positions; a seeded sample of graphs checked against the oracle (tree, repulsion)
            at the start and bit for bit against the same graphs laid out alone at the end.
  sbm50k    (the learned initialiser's graphs): features, loss and gradient of one 50,000-node
            graph against the FP64 oracle.
  rmat24    (R-MAT 2^24 / 128M edges): at iterations 0 and 199 no error flag, the whole tree
            bit-exact against the oracle's (root count V, root mass the float sum of 2E + V), a
            valid permutation, and the repulsion of a seeded sample of
            walk groups within 1e-4 of the oracle's walk of those groups over the whole tree;
            the Barnes-Hut error against FP64 exact sums on a node sample is reported.
Progress lines go to the real stderr (the large cases spend minutes building their graphs).
"""
import sys
import time

import numpy as np
import pytest

from gpugraphlayout_amd import (BatchEngine, Engine, GFA2_FORCE_ATTRACTION_GRAVITY, GFA2_FORCE_REPULSION, Mpnn,
                                workloads)

pytestmark = pytest.mark.gpu


def progress(msg):
    print(f"[configs] {msg}", file=sys.__stderr__, flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def check_tree(T, ref, what):
    assert len(T["level"]) == len(ref["level"]), what
    np.testing.assert_array_equal(T["level"], ref["level"], err_msg=what)
    np.testing.assert_array_equal(T["count"], ref["count"], err_msg=what)
    np.testing.assert_array_equal(T["mass"].view(np.uint32), ref["mass"].view(np.uint32), err_msg=what)
    np.testing.assert_array_equal(T["perm"], ref["perm"], err_msg=what)
    scale = max(float(ref["root"][2]) * 2, 1e-30)
    assert np.max(np.abs(T["xy"] - ref["xy"])) <= 1e-6 * scale, what


def test_ba10k_against_oracle_at_checkpoints(oracle):
    progress("case ba10k checkpoints 0/100/499")
    g, pos, prm, _ = workloads.single("ba10k")
    mass = g.mass()
    eng = Engine(g.n, g.src, g.dst, **prm)
    eng.set_positions(pos)
    rng = np.random.default_rng(5)
    sample = rng.choice(g.n, 2048, replace=False)
    done = 0
    for check in (0, 100, 499):
        eng.step(check - done)
        done = check
        eng.sync()
        p = eng.positions()
        assert np.all(np.isfinite(p))
        eng.set_capture(True)
        eng.step(1)
        done += 1
        eng.sync()
        eng.set_capture(False)
        rep = eng.forces(GFA2_FORCE_REPULSION)
        fga = eng.forces(GFA2_FORCE_ATTRACTION_GRAVITY)
        T = eng.tree()
        st = eng.stats()
        check_tree(T, oracle.tree(p, mass), f"tree at iteration {check}")
        ref, info = oracle.bh_repulsion(p, mass, prm["scale"], group=32)
        err = rel_l2(rep, ref)
        want_fga = oracle.attraction_coo(p, g.src, g.dst, oracle.gravity(p, mass, 1, prm["gravity"]))
        err_fga = rel_l2(fga, want_fga)
        bh = rel_l2(rep[sample], oracle.exact_repulsion_f64(p, mass, prm["scale"], sample))
        progress(f"ba10k iteration {check}: repulsion rel-L2 {err:.2e}, gravity+attraction {err_fga:.2e}, "
                 f"Barnes-Hut error vs FP64 exact {bh:.2e}")
        assert err <= 1e-4
        assert err_fga <= 1e-5
        assert bh <= 5e-2
        assert abs(st["interactions"] - info["interactions"]) <= 1e-5 * info["interactions"]
        assert st["error_flags"] == 0
    eng.close()


def test_ba10k_speed_updates_match_oracle_every_iteration(oracle):
    """The adaptive speed step (src/RPFA2Kernels.cu:167-219) on the device against the oracle's
    restatement, from the device's own S and T, at every iteration of the 500-iteration run and
    of a run started at a far too high speed; between them the erratic (S/T > 2), efficiency-down
    and efficiency-up branches all occur, and the speed both rises and falls."""
    g, pos, prm, _ = workloads.single("ba10k")
    seen = set()
    for gs0 in (1.0, 400.0):
        progress(f"case ba10k speed updates from global speed {gs0}")
        eng = Engine(g.n, g.src, g.dst, global_speed=gs0, **prm)
        eng.set_positions(pos)
        st = eng.stats()
        state = np.array([st["global_speed"], st["speed_efficiency"], 1.0, 10.0], np.float32)
        for it in range(500 if gs0 == 1.0 else 120):
            eng.step(1)
            st = eng.stats()
            want = oracle.speed_update(g.n, st["swing"], st["traction"], state)
            got = np.array([st["global_speed"], st["speed_efficiency"]], np.float32)
            assert got[0] == want[0] and got[1] == want[1], (gs0, it, got, want[:2])
            S, T = np.float32(st["swing"]), np.float32(st["traction"])
            if S / T > 2.0:
                seen.add("erratic")
            ratio = float(want[1]) / float(state[1])
            if ratio < 1.0 and not np.isclose(ratio, 0.5):
                seen.add("efficiency down")
            if ratio > 1.0 or np.isclose(ratio, 0.65):
                seen.add("efficiency up")
            seen.add("rise" if want[0] > state[0] else "fall")
            state = want
        eng.close()
    progress(f"speed-update branches seen: {sorted(seen)}")
    assert {"erratic", "efficiency down", "efficiency up", "rise", "fall"} <= seen


def test_interaction_totals_add_up():
    """stats()['interactions_total'] grows by each iteration's own count (single graph), and a
    batch's total is the sum of its graphs' per-iteration counts."""
    progress("case interaction totals")
    g, pos, prm, _ = workloads.single("ba10k")
    eng = Engine(g.n, g.src, g.dst, **prm)
    eng.set_positions(pos)
    t0 = eng.stats()["interactions_total"]
    acc = 0
    for _ in range(7):
        eng.step(1)
        st = eng.stats()
        assert st["interactions"] > 0
        acc += st["interactions"]
    assert eng.stats()["interactions_total"] - t0 == acc
    eng.close()
    items = workloads.batch4096(count=12, seed=9)
    be = BatchEngine([(q.n, q.src, q.dst) for q, _ in items], **workloads.BATCH_PARAMS)
    be.set_positions(np.concatenate([p for _, p in items]))
    t0 = be.stats()["interactions_total"]
    acc = 0
    for _ in range(5):
        be.step(1)
        acc += int(be.batch_stats()["interactions"].sum())
    assert be.stats()["interactions_total"] - t0 == acc
    be.close()


def split_cells(T, ngraphs):
    """A batch's pre-order cell list split per graph (each graph's cells start at its root)."""
    roots = np.flatnonzero(T["level"] == 0)
    assert len(roots) == ngraphs
    bounds = np.append(roots, len(T["level"]))
    return [{k: T[k][bounds[q]:bounds[q + 1]] for k in ("level", "count", "mass", "xy")} for q in range(ngraphs)]


def test_batch4096_against_oracle_and_single_runs(oracle):
    progress("case batch4096: building 4,096 graphs")
    t0 = time.time()
    items = workloads.batch4096()
    G = len(items)
    off = np.concatenate([[0], np.cumsum([g.n for g, _ in items])])
    be = BatchEngine([(g.n, g.src, g.dst) for g, _ in items], **workloads.BATCH_PARAMS)
    start = np.concatenate([p for _, p in items])
    be.set_positions(start)
    progress(f"batch4096 built in {time.time() - t0:.1f}s: sum V = {off[-1]}")
    be.set_capture(True)
    be.step(1)
    be.sync()
    be.set_capture(False)
    rep = be.forces(GFA2_FORCE_REPULSION)
    T = be.tree()
    st = be.stats()
    assert st["error_flags"] == 0
    cells = split_cells(T, G)
    for k, (g, _) in enumerate(items):
        assert cells[k]["count"][0] == g.n
        assert cells[k]["mass"][0] == np.float32(2 * g.m + g.n)
        perm = T["perm"][off[k]:off[k + 1]] - off[k]
        assert np.array_equal(np.sort(perm), np.arange(g.n))
    rng = np.random.default_rng(11)
    sample = np.sort(rng.choice(G, 8, replace=False))
    for k in sample:
        g, p = items[k]
        ref = oracle.tree(p, g.mass())
        got = dict(cells[k], perm=T["perm"][off[k]:off[k + 1]] - off[k])
        check_tree(got, ref, f"graph {k} tree")
        want, _ = oracle.bh_repulsion(p, g.mass(), workloads.BATCH_PARAMS["scale"], group=32)
        err = rel_l2(rep[off[k]:off[k + 1]], want)
        assert err <= 1e-4, (k, err)
    progress("batch4096: sampled graphs match the oracle at iteration 0; running 299 more iterations")
    be.step(workloads.BATCH_ITERATIONS - 1)
    be.sync()
    final = be.positions()
    assert np.all(np.isfinite(final))
    assert be.stats()["error_flags"] == 0
    bstats = be.batch_stats()
    for k in sample:
        g, p = items[k]
        e = Engine(g.n, g.src, g.dst, **workloads.BATCH_PARAMS)
        e.set_positions(p)
        e.step(workloads.BATCH_ITERATIONS)
        e.sync()
        np.testing.assert_array_equal(final[off[k]:off[k + 1]].view(np.uint32), e.positions().view(np.uint32),
                                      err_msg=f"graph {k} after {workloads.BATCH_ITERATIONS} iterations")
        s1 = e.stats()
        assert bstats["swing"][k] == s1["swing"] and bstats["speed"][k] == np.float32(s1["global_speed"])
        e.close()
    progress("batch4096: sampled graphs bit-identical to their single runs after 300 iterations")
    be.close()


def test_sbm50k_features_and_gradient_match_oracle():
    from oracle import mpnn_oracle as mo
    progress("case sbm50k features and gradient")
    g = workloads.sbm50k(5000)
    rowptr, col = g.csr()
    rng = np.random.default_rng(8)
    target = rng.normal(size=(g.n, 2)).astype(np.float32) * 40.0
    m = Mpnn(seed=4)
    gid = m.add_graph(g.n, rowptr, col, target=target, feature_seed=5000, steps=2)
    X = m.features(gid)
    X_ref = mo.features(rowptr, col, 5000, 2)
    assert np.max(np.abs(X - X_ref)) < 1e-4
    offs = [13, 977, 20011]
    loss, grad = m.loss_grad(gid, offs)
    T = target.astype(np.float64) - target.mean(0)
    T /= np.sqrt((T ** 2).sum(1).mean())
    L, Gr = mo.loss_grad(m.params().astype(np.float64), rowptr, col, X, T, offs, round_bf16=True)
    progress(f"sbm50k: loss {loss:.6g} (oracle {L:.6g}), gradient rel-L2 {rel_l2(grad, Gr):.2e}")
    assert loss == pytest.approx(L, rel=1e-5)
    assert rel_l2(grad, Gr) < 1e-4
    m.close()


def test_rmat24_properties_and_sampled_groups(oracle):
    progress("case rmat24: generating R-MAT 2^24 / 128M edges (minutes)")
    t0 = time.time()
    g, pos, prm, _ = workloads.single("rmat24")
    progress(f"rmat24: V = {g.n}, E = {g.m}, built in {time.time() - t0:.0f}s")
    mass = g.mass()
    eng = Engine(g.n, g.src, g.dst, **prm)
    eng.set_positions(pos)
    rng = np.random.default_rng(3)
    groups = np.sort(rng.choice((g.n + 31) // 32, 48, replace=False)) * 32
    sample = rng.choice(g.n, 512, replace=False)
    done = 0
    for check in (0, 199):
        eng.step(check - done)
        done = check
        eng.sync()
        p = eng.positions()
        assert np.all(np.isfinite(p))
        eng.set_capture(True)
        eng.step(1)
        done += 1
        eng.sync()
        eng.set_capture(False)
        rep = eng.forces(GFA2_FORCE_REPULSION)
        st = eng.stats()
        assert st["error_flags"] == 0
        T = eng.tree()
        assert T["level"][0] == 0 and T["count"][0] == g.n
        # 2E + V is past 2^24: the root's float mass is the quadrant-order float sum, as the
        # reference's (the oracle's tree is checked bit for bit below)
        assert abs(float(T["mass"][0]) - (2 * g.m + g.n)) <= 1e-6 * (2 * g.m + g.n)
        assert np.array_equal(np.bincount(T["perm"], minlength=g.n), np.ones(g.n, np.int64))
        check_tree(T, oracle.tree(p, mass), f"rmat24 tree at iteration {check}")
        ref, info = oracle.bh_repulsion_groups(p, mass, prm["scale"], groups, group=32)
        np.testing.assert_array_equal(T["perm"], info["perm"])
        ids = np.concatenate([info["perm"][s:s + 32] for s in groups])
        err = rel_l2(rep[ids], ref[ids])
        bh = rel_l2(rep[sample], oracle.exact_repulsion_f64(p, mass, prm["scale"], sample))
        progress(f"rmat24 iteration {check}: sampled-group repulsion rel-L2 {err:.2e} over {len(ids)} bodies, "
                 f"Barnes-Hut error vs FP64 exact (512 nodes) {bh:.2e}, cells {st['cells']}, "
                 f"max level {st['max_level']}")
        assert err <= 1e-4
        assert bh <= 5e-2
    eng.close()

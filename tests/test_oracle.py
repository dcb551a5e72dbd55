"""Pins the oracle (CPU only): each oracle function is checked against something other than
its own formula — the reference's own code (golden vectors made by tests/golden/make_golden.py
from the compiled reference sources, and live when oracle/_ref is built), an independent
construction (a Python trie for the quadtree), brute force (exact pairwise as theta -> 0),
closed forms and invariants.
"""
import os

import numpy as np
import pytest

from gpugraphlayout_amd import graphs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F32 = np.float32
EPS = F32(0.05 * 0.05)


# ---------------------------------------------------------------- independent quadtree
def path_digits(pos, rootx, rooty, radius, levels=48):
    """Quadrant digits along each body's path, replaying float32 centre arithmetic:
    r halves per level, the child centre is the parent centre +/- r."""
    px, py = pos[:, 0].astype(F32), pos[:, 1].astype(F32)
    r = F32(radius) * F32(0.5)
    bx, by = F32(rootx) < px, F32(rooty) < py
    x = (F32(rootx) + np.where(bx, r, -r)).astype(F32)
    y = (F32(rooty) + np.where(by, r, -r)).astype(F32)
    d = [bx.astype(np.int64) | (by.astype(np.int64) << 1)]
    for _ in range(1, levels):
        r = F32(r * F32(0.5))
        bx, by = x < px, y < py
        d.append(bx.astype(np.int64) | (by.astype(np.int64) << 1))
        x = (x + np.where(bx, r, -r)).astype(F32)
        y = (y + np.where(by, r, -r)).astype(F32)
    return np.stack(d, 1)


def trie(pos, mass):
    """Pre-order cells (level, count, mass, centroid) and in-order body permutation of the
    quadtree whose cells are the path prefixes shared by two or more bodies."""
    minx, maxx = pos[:, 0].min(), pos[:, 0].max()
    miny, maxy = pos[:, 1].min(), pos[:, 1].max()
    radius = F32(max(F32(maxx - minx), F32(maxy - miny)) * F32(0.5))
    rootx, rooty = F32((minx + maxx) * F32(0.5)), F32((miny + maxy) * F32(0.5))
    D = path_digits(pos, rootx, rooty, radius)
    cells, perm = [], []

    def rec(idx, level):
        me = len(cells)
        cells.append([level, len(idx), None, None, None])
        cm, sx, sy = F32(0), F32(0), F32(0)
        for q in range(4):
            sub = idx[D[idx, level] == q]
            if len(sub) == 0:
                continue
            if len(sub) == 1:
                b = sub[0]
                perm.append(b)
                m, cx, cy = F32(mass[b]), F32(pos[b, 0]), F32(pos[b, 1])
            else:
                c = rec(sub, level + 1)
                m, cx, cy = cells[c][2], cells[c][3], cells[c][4]
            sx = F32(sx + F32(cx * m))
            sy = F32(sy + F32(cy * m))
            cm = F32(cm + m)
        inv = F32(F32(1) / cm)
        cells[me][2:] = [cm, F32(sx * inv), F32(sy * inv)]
        return me

    rec(np.arange(len(pos)), 0)
    arr = np.array([[c[0], c[1]] for c in cells], np.int64)
    return {"level": arr[:, 0], "count": arr[:, 1], "mass": np.array([c[2] for c in cells], F32),
            "xy": np.array([[c[3], c[4]] for c in cells], F32), "perm": np.array(perm), "radius": radius}


def rand_case(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:
        pos = np.concatenate([rng.normal(0, 0.01, (n // 2, 2)), rng.normal(5, 2, (n - n // 2, 2))])
    else:
        pos = (rng.random((n, 2)) - 0.5) * 2 * np.sqrt(n)
    mass = rng.integers(1, 30, n).astype(F32)
    return pos.astype(F32), mass


@pytest.mark.parametrize("n,seed,clustered", [(1, 0, False), (2, 1, False), (5, 2, False), (300, 3, False),
                                              (400, 4, True)])
def test_tree_matches_independent_trie(oracle, n, seed, clustered):
    pos, mass = rand_case(n, seed, clustered)
    T = oracle.tree(pos, mass)
    R = trie(pos, mass)
    np.testing.assert_array_equal(T["level"], R["level"])
    np.testing.assert_array_equal(T["count"], R["count"])
    np.testing.assert_array_equal(T["perm"], R["perm"])
    np.testing.assert_array_equal(T["mass"].view(np.uint32), R["mass"].view(np.uint32))
    np.testing.assert_array_equal(T["xy"].view(np.uint32), R["xy"].view(np.uint32))


def test_tree_insertion_order_independent(oracle):
    pos, mass = rand_case(500, 9)
    T = oracle.tree(pos, mass)
    p = np.random.default_rng(1).permutation(len(pos))
    U = oracle.tree(pos[p], mass[p])
    np.testing.assert_array_equal(T["level"], U["level"])
    np.testing.assert_array_equal(T["count"], U["count"])
    np.testing.assert_array_equal(T["mass"], U["mass"])
    np.testing.assert_array_equal(p[U["perm"]], T["perm"])


def test_tree_hand_cases(oracle):
    # four bodies, one per quadrant: the root alone, bodies in quadrant order 0,1,2,3
    pos = np.array([[1, 1], [-1, -1], [1, -1], [-1, 1]], F32)   # quadrants 3, 0, 1, 2
    T = oracle.tree(pos, np.ones(4, F32))
    assert T["level"].tolist() == [0]
    assert T["count"].tolist() == [4]
    assert T["perm"].tolist() == [1, 2, 3, 0]
    # a close pair forces a chain of single-child cells down to where they part
    pos = np.array([[-10, -10], [10, 10], [9.0, 9.0], [9.01, 9.0]], F32)
    T = oracle.tree(pos, np.array([1, 2, 3, 4], F32))
    assert T["count"][0] == 4 and T["mass"][0] == 10
    c = T["count"]
    assert np.all(np.diff(c) <= 0) and c[-1] == 2 and 3 in c.tolist()   # 4 -> 3 ... -> 2 ... (a chain)
    assert np.all(np.diff(T["level"]) == 1)
    # the pair parts where a cell is narrower than its separation (0.01): deeper than
    # log2(20/0.01) - 1 levels, and no deeper than 2 levels past that
    assert np.log2(20 / 0.01) - 1 <= T["level"][-1] + 1 <= np.log2(20 / 0.01) + 2
    assert T["mass"][-1] == 7
    # root mass is the sum of masses, count the number of bodies
    pos, mass = rand_case(1000, 5)
    T = oracle.tree(pos, mass)
    assert T["count"][0] == 1000 and T["mass"][0] == mass.sum()


def test_bbox(oracle):
    pos, _ = rand_case(777, 12)
    b = oracle.bbox(pos)
    mn, mx = pos.min(0), pos.max(0)
    assert b[0] == F32((mn[0] + mx[0]) * F32(0.5)) and b[1] == F32((mn[1] + mx[1]) * F32(0.5))
    assert b[2] == F32(max(mx[0] - mn[0], mx[1] - mn[1]) * F32(0.5))


# ---------------------------------------------------------------- walk
def per_body_walk(pos, mass, k_r, itolsq):
    """Barnes-Hut with a per-body acceptance (group of one), over the independent trie, in
    float64: cells at depth d (children of the root at 0) accepted iff
    |p-c|^2 + eps^2 >= (2 radius)^2 itolsq / 4^d + eps^2."""
    R = trie(pos, mass)
    radius = float(R["radius"])
    D = path_digits(pos, *oracle_root(pos), levels=48)
    n = len(pos)
    out = np.zeros((n, 2))
    eps = float(EPS)
    pos64, m64 = pos.astype(np.float64), mass.astype(np.float64)

    def walk(i, idx, level):
        acc = np.zeros(2)
        for q in range(4):
            sub = idx[D[idx, level] == q]
            if len(sub) == 0:
                continue
            if len(sub) == 1:
                d = pos64[i] - pos64[sub[0]]
                acc += k_r * m64[i] * m64[sub[0]] * d / (d @ d + eps)
                continue
            m = m64[sub].sum()
            c = (pos64[sub] * m64[sub, None]).sum(0) / m
            d = pos64[i] - c
            if d @ d + eps >= (2 * radius) ** 2 * itolsq / 4 ** level + eps:
                acc += k_r * m64[i] * m * d / (d @ d + eps)
            else:
                acc += walk(i, sub, level + 1)
        return acc

    for i in range(n):
        out[i] = walk(i, np.arange(n), 0)
    return out


def oracle_root(pos):
    minx, maxx = pos[:, 0].min(), pos[:, 0].max()
    miny, maxy = pos[:, 1].min(), pos[:, 1].max()
    radius = F32(max(F32(maxx - minx), F32(maxy - miny)) * F32(0.5))
    return F32((minx + maxx) * F32(0.5)), F32((miny + maxy) * F32(0.5)), radius


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


def test_walk_group_of_one_matches_per_body_walk(oracle):
    pos, mass = rand_case(120, 21)
    rep, _ = oracle.bh_repulsion(pos, mass, 3.0, group=1)
    ref = per_body_walk(pos, mass, 3.0, 1.0)
    assert rel(rep, ref) < 1e-5


def test_walk_theta_to_zero_is_exact_pairwise(oracle):
    pos, mass = rand_case(300, 22)
    rep, info = oracle.bh_repulsion(pos, mass, 2.0, theta=1e-6, group=32)
    exact = oracle.exact_repulsion_f64(pos, mass, 2.0, np.arange(300))
    assert rel(rep, exact) < 1e-5
    assert info["interactions"] == 300 * 300    # every body met directly, itself included


def test_walk_group_vote_is_conservative(oracle):
    """A group accepts a cell only if every member would: larger groups open more cells,
    so their interaction count is never smaller and their error against exact pairwise
    is never much larger."""
    pos, mass = rand_case(2000, 23)
    exact = oracle.exact_repulsion_f64(pos, mass, 1.0, np.arange(2000))
    errs, inter = [], []
    for G in (1, 32, 64):
        rep, info = oracle.bh_repulsion(pos, mass, 1.0, group=G)
        errs.append(rel(rep, exact))
        inter.append(info["interactions"])
    assert inter[0] <= inter[1] <= inter[2]
    assert errs[2] <= errs[0] * 1.05
    assert max(errs) < 0.05


def test_walk_of_listed_groups_equals_full_walk(oracle):
    """Walking only some groups (the large-graph check) gives those groups' bodies exactly
    the full walk's forces, and nothing to the others."""
    pos, mass = rand_case(3000, 24)
    full, info = oracle.bh_repulsion(pos, mass, 2.0, group=32)
    starts = np.array([0, 32 * 7, 32 * 50, 32 * 93], np.int32)
    part, pinfo = oracle.bh_repulsion_groups(pos, mass, 2.0, starts, group=32)
    ids = np.concatenate([info["perm"][s:s + 32] for s in starts])
    np.testing.assert_array_equal(pinfo["perm"], info["perm"])
    np.testing.assert_array_equal(part[ids].view(np.uint32), full[ids].view(np.uint32))
    rest = np.setdiff1d(np.arange(3000), ids)
    assert not np.any(part[rest])
    assert 0 < pinfo["interactions"] < info["interactions"]


def test_exact_pairwise_does_not_depend_on_threads(oracle):
    pos, mass = rand_case(500, 25)
    a = oracle.exact_repulsion_f64(pos, mass, 1.0, np.arange(37), threads=1)
    b = oracle.exact_repulsion_f64(pos, mass, 1.0, np.arange(37), threads=5)
    np.testing.assert_array_equal(a, b)


def test_walk_symmetric_square(oracle):
    pos = np.array([[-1, -1], [1, -1], [-1, 1], [1, 1]], F32)
    rep, _ = oracle.bh_repulsion(pos, np.full(4, 2, F32), 1.0, theta=1e-6)
    mag = np.linalg.norm(rep, axis=1)
    assert np.allclose(mag, mag[0], rtol=1e-6)
    assert np.all(np.sign(rep) == np.sign(pos))        # pushed outward


def test_exact_pairwise_closed_form(oracle):
    pos = np.array([[0, 0], [3, 4]], F32)
    mass = np.array([2, 5], F32)
    out = oracle.exact_repulsion_f64(pos, mass, 1.5, [0, 1])
    d = np.array([-3.0, -4.0])
    expect0 = 1.5 * 2 * 5 * d / (25 + float(EPS))
    assert np.allclose(out[0], expect0, rtol=1e-12)
    assert np.allclose(out[1], -expect0, rtol=1e-12)


# ---------------------------------------------------------------- FA2 pieces
def test_gravity_closed_forms(oracle):
    pos = np.array([[3, 4], [0, 0], [-1, 0]], F32)
    mass = np.array([2, 3, 1], F32)
    fs = oracle.gravity(pos, mass, True, 0.5)
    assert np.allclose(fs, [[-3, -4], [0, 0], [0.5, 0]])
    fw = oracle.gravity(pos, mass, False, 0.5)
    assert np.allclose(fw, [[-0.6, -0.8], [0, 0], [0.5, 0]], rtol=1e-6)


def test_attraction_closed_form_and_balance(oracle):
    s, t = graphs.barabasi_albert(200, 3, 4)
    g = graphs.edges_to_graph(s, t)
    pos, _ = rand_case(g.n, 5)
    f = oracle.attraction_coo(pos, g.src, g.dst)
    assert np.abs(f.astype(np.float64).sum(0)).max() < 1e-3
    p64 = pos.astype(np.float64)
    ref = np.zeros((g.n, 2))
    for a, b in zip(g.src, g.dst):
        ref[a] += p64[b] - p64[a]
        ref[b] += p64[a] - p64[b]
    assert rel(f, ref) < 1e-6


def test_speed_known_answers(oracle):
    # n=4, S=T=1: eojt=0.1, jt=sqrt(0.1); S > jt*T so efficiency *= 0.7;
    # target = jt*1*T/S = 0.3162; speed += min(target-1, 0.5) = target
    st = oracle.speed_update(4, 1.0, 1.0, [1.0, 1.0, 1.0, 10.0])
    assert st[1] == pytest.approx(0.7, rel=1e-6)
    assert st[0] == pytest.approx(np.sqrt(0.1), rel=1e-6)
    # S/T = 3 > 2: efficiency halves, jt raised to the jitter tolerance (1):
    # target = 1*0.5*1/3, and S > jt*T halves... then *0.7
    st = oracle.speed_update(4, 3.0, 1.0, [1.0, 1.0, 1.0, 10.0])
    assert st[1] == pytest.approx(0.35, rel=1e-6)
    assert st[0] == pytest.approx(1.0 + (0.5 / 3 - 1.0), rel=1e-6)
    # S small: efficiency grows by 1.3 while speed < 1000; the rise is capped at 50%
    st = oracle.speed_update(4, 0.01, 1.0, [1.0, 1.0, 1.0, 10.0])
    assert st[1] == pytest.approx(1.3, rel=1e-6)
    assert st[0] == pytest.approx(1.5, rel=1e-6)
    # speed >= 1000: efficiency unchanged; a fall is not capped: speed = target = jt*T/S
    st = oracle.speed_update(4, 0.01, 1.0, [2000.0, 1.0, 1.0, 10.0])
    assert st[1] == pytest.approx(1.0) and st[0] == pytest.approx(np.sqrt(0.1) / 0.01, rel=1e-6)


def test_speed_sums(oracle):
    f = np.array([[3, 4], [0, 1]], F32)
    fp = np.array([[0, 0], [0, -1]], F32)
    m = np.array([2, 3], F32)
    _, st = oracle.speed(f, fp, m, [1, 1, 1, 10])
    assert st[0] == pytest.approx(2 * 5 + 3 * 2)        # sum m |F-F'|
    assert st[1] == pytest.approx(2 * 2.5 + 3 * 0)      # sum m |F+F'| / 2


def test_displacement_closed_form(oracle):
    f = np.array([[3, 4]], F32)
    fp = np.array([[0, 0]], F32)
    p, fp2 = oracle.displace(4.0, f, fp, np.array([[1, 1]], F32))
    factor = 4.0 / (1 + np.sqrt(4.0 * 5.0))
    assert np.allclose(p, [[1 + 3 * factor, 1 + 4 * factor]], rtol=1e-6)
    assert np.array_equal(fp2, f)


# ---------------------------------------------------------------- reference's own code
def _golden(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.mark.parametrize("name", ["hand", "chain", "random_a", "random_b"])
def test_loader_matches_reference_ugraph(name):
    gold = _golden(f"ugraph_{name}.npz")
    g = graphs.load_edgelist(os.path.join(GOLDEN, "edgelists", f"{name}.txt"))
    assert g.n == int(gold["n"]) and g.m == int(gold["m"])
    np.testing.assert_array_equal(g.ids, gold["ids"])
    np.testing.assert_array_equal(g.degrees(), gold["deg"])
    np.testing.assert_array_equal(g.src, gold["src"])
    np.testing.assert_array_equal(g.dst, gold["dst"])


@pytest.mark.parametrize("n", [7, 1000])
def test_initial_layout_matches_reference(n):
    gold = _golden(f"init_{n}.npz")
    xy = graphs.glibc_random_positions(n, float(gold["w"]), float(gold["w"]), seed=int(gold["seed"]))
    np.testing.assert_array_equal(xy.view(np.uint32), gold["xy"].view(np.uint32))


@pytest.mark.parametrize("name", ["uniform", "clustered", "coincident"])
def test_cpu_baseline_bh_matches_reference(oracle, name):
    gold = _golden(f"cpu_bh_{name}.npz")
    f = oracle.cpu_bh_forces(gold["pos"], gold["mass"], gold["cx"], gold["cy"], gold["length"], gold["theta"],
                             gold["pos"], gold["mass"])
    np.testing.assert_array_equal(f.view(np.uint32), gold["forces"].view(np.uint32))


def test_cpu_baseline_bh_live_reference(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    pos, mass = rand_case(700, 31, clustered=True)
    span = np.float32(np.ceil(np.ptp(pos, 0).max()))
    cx, cy = [F32(pos[:, k].min() + np.ptp(pos[:, k]) / 2.0) for k in (0, 1)]
    a = oracle.cpu_bh_forces(pos, mass, cx, cy, span + 10, 1.0, pos, mass)
    b = oracle.cpu_bh_forces(pos, mass, cx, cy, span + 10, 1.0, pos, mass, use_ref=True)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_cpu_baseline_two_node_step(oracle):
    """Two nodes, one edge: the CPU Barnes-Hut opens the root (length > distance), so the
    step is exact: gravity (strong), attraction p_t - p_s, repulsion k_r m m d/|d|^2."""
    pos = np.array([[-1.0, 0.5], [2.0, -1.5]], F32)
    cb = oracle.CPUBaseline(2, np.array([0]), np.array([1]), pos, True, 0.5, 2.0)
    cb.step(1)
    p1 = cb.positions()
    st = cb.state()
    m = np.array([2.0, 2.0])
    p = pos.astype(np.float64)
    d = p[0] - p[1]
    F = np.zeros((2, 2))
    F[0] = -0.5 * m[0] * p[0] + (p[1] - p[0]) + 2.0 * m[0] * m[1] * d / (d @ d)
    F[1] = -0.5 * m[1] * p[1] + (p[0] - p[1]) - 2.0 * m[0] * m[1] * d / (d @ d)
    S = (m * np.linalg.norm(F, axis=1)).sum()
    T = (m * np.linalg.norm(F, axis=1) / 2).sum()
    ref_state = oracle.speed_update(2, S, T, [1.0, 1.0, 1.0, 10.0])
    assert st[0] == pytest.approx(float(ref_state[0]), rel=1e-5)
    assert st[4] == pytest.approx(S, rel=1e-5) and st[5] == pytest.approx(T, rel=1e-5)
    gs = float(st[0])
    expect = p + F * (gs / (1 + np.sqrt(gs * np.linalg.norm(F, axis=1))))[:, None]
    assert np.allclose(p1, expect, rtol=1e-5, atol=1e-5)
    cb.close()


# ---------------------------------------------------------------- quality metric
def test_normalized_stress(oracle):
    s, t = graphs.barabasi_albert(60, 2, 3)
    g = graphs.edges_to_graph(s, t)
    rowptr, col = g.csr()
    d0 = oracle.bfs(rowptr, col, 0)
    # BFS against a plain Python BFS
    adj = [[] for _ in range(g.n)]
    for a, b in zip(g.src, g.dst):
        adj[a].append(b)
        adj[b].append(a)
    ref = [-1] * g.n
    ref[0] = 0
    q = [0]
    for v in q:
        for u in adj[v]:
            if ref[u] < 0:
                ref[u] = ref[v] + 1
                q.append(u)
    assert d0.tolist() == ref
    pos, _ = rand_case(g.n, 4)
    a = np.zeros(g.n - 1, np.int32)
    b = np.arange(1, g.n, dtype=np.int32)
    ns, alpha = oracle.normalized_stress(a, b, d0[1:], pos)
    ns2, alpha2 = oracle.normalized_stress(a, b, d0[1:], pos * 7.0)
    assert ns == pytest.approx(ns2, rel=1e-6) and alpha2 == pytest.approx(alpha / 7.0, rel=1e-6)
    D = np.linalg.norm(pos[a].astype(np.float64) - pos[b], axis=1) / d0[1:]
    al = D.sum() / (D ** 2).sum()
    assert ns == pytest.approx(np.mean((al * D - 1) ** 2), rel=1e-9)
    # a layout with distances proportional to hop distance has zero stress
    line = np.stack([np.arange(5.0), np.zeros(5)], 1).astype(F32)
    ns0, _ = oracle.normalized_stress(np.zeros(4, np.int32), np.arange(1, 5, dtype=np.int32),
                                      np.arange(1, 5, dtype=np.int32), line)
    assert ns0 == pytest.approx(0.0, abs=1e-12)


def test_attraction_general_closed_forms(oracle):
    pos = np.array([[0.0, 0.0], [3.0, 0.0]], np.float32)
    s, t = np.array([0]), np.array([1])
    # linear, unweighted: node 0 pulled by (3, 0)
    np.testing.assert_allclose(oracle.attraction_general(pos, s, t), [[3, 0], [-3, 0]], rtol=1e-7)
    # linlog: magnitude log(1 + d)
    f = oracle.attraction_general(pos, s, t, linlog=True)
    np.testing.assert_allclose(f, [[np.log(4.0), 0], [-np.log(4.0), 0]], rtol=1e-6)
    # weights: w^delta
    for delta, k in ((0.0, 1.0), (1.0, 2.0), (2.0, 4.0), (0.5, np.sqrt(2.0))):
        f = oracle.attraction_general(pos, s, t, w=np.array([2.0]), delta=delta)
        np.testing.assert_allclose(f[0], [3 * k, 0], rtol=1e-6)
    f = oracle.attraction_general(pos, s, t, w=np.array([3.0]), delta=1.0, linlog=True)
    np.testing.assert_allclose(f[0], [3 * np.log(4.0), 0], rtol=1e-6)
    # coincident endpoints add nothing under linlog (the reference's 0-vector * FLT_MAX)
    f = oracle.attraction_general(np.zeros((2, 2), np.float32), s, t, linlog=True)
    assert np.all(f == 0)


def test_attraction_general_sums_to_zero_and_matches_linear(oracle):
    rng = np.random.default_rng(3)
    n, m = 300, 1200
    src = rng.integers(0, n, m)
    dst = (src + 1 + rng.integers(0, n - 1, m)) % n
    pos = rng.normal(size=(n, 2)).astype(np.float32) * 10
    for kw in ({}, {"linlog": True}, {"w": rng.uniform(0.5, 2.0, m), "delta": 1.5}):
        f = oracle.attraction_general(pos, src, dst, **kw).astype(np.float64)
        assert np.abs(f.sum(0)).max() < 1e-3 * np.abs(f).max()
    np.testing.assert_array_equal(oracle.attraction_general(pos, src, dst), oracle.attraction_coo(pos, src, dst))

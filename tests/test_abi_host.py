"""CPU checks of the C ABI library and the host-side logic (no device compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

import gpugraphlayout_amd as g2
from gpugraphlayout_amd import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "gfa2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gfa2_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_bound_symbols():
    assert header_symbols() == sorted(g2.ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(g2.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_library_host_only_calls():
    lib = g2.lib()
    assert lib.gfa2_strerror(0) == b"ok"
    assert b"capacity" in lib.gfa2_strerror(-5)
    p = g2.default_params()
    assert p.use_barneshut == 1 and p.theta == 1.0 and p.group_size == 32
    assert p.epssq == np.float32(0.05 * 0.05)
    assert p.k_s_max == 10.0 and p.jitter_tolerance == 1.0 and p.global_speed == 1.0
    # invalid arguments are rejected before any device work
    h = ctypes.c_void_p()
    bad = g2.default_params(group_size=48)
    assert lib.gfa2_create(ctypes.byref(bad), 0, ctypes.byref(h)) != 0
    assert lib.gfa2_step(None, 1) == -3
    overlap = g2.default_params(prevent_overlap=1)
    assert lib.gfa2_create(ctypes.byref(overlap), 0, ctypes.byref(h)) == -8   # unsupported, as the reference
    assert lib.gfa2_exact_repulsion(None, 1, None) == -3
    c = g2.LaunchConfig()
    assert lib.gfa2_tuned_config(0, ctypes.byref(c)) == -3
    assert lib.gfa2_tuned_config(1000, ctypes.byref(c)) == -7   # nothing tuned in this process
    assert lib.gfa2_autotune(None, 3, ctypes.byref(c)) == -3
    assert lib.gfa2_set_launch_config(None, ctypes.byref(c)) == -3


def test_ba_generator_invariants():
    s, t = graphs.barabasi_albert(1000, 5, 0)
    assert len(s) == 5 + (1000 - 6) * 5
    g = graphs.edges_to_graph(s, t)
    assert g.n == 1000 and g.m == len(s)        # no duplicates, no loops
    assert g.degrees().min() >= 1 and g.degrees().sum() == 2 * g.m
    s2, t2 = graphs.barabasi_albert(1000, 5, 0)
    assert np.array_equal(s, s2) and np.array_equal(t, t2)


def test_rmat_generator_invariants():
    s, t = graphs.rmat(12, 30000, seed=1)
    assert len(s) == 30000 and np.all(s != t)
    code = np.minimum(s, t) << 12 | np.maximum(s, t)
    assert len(np.unique(code)) == 30000
    assert s.max() < 4096 and t.max() < 4096
    g = graphs.edges_to_graph(s, t)
    assert g.m == 30000 and g.n <= 4096
    # the skew: node 0 (all-zero bits, the most likely quadrant path) is the top hub
    deg = np.bincount(np.concatenate([s, t]), minlength=4096)
    assert deg.argmax() == 0


def _splitmix_at(seed, k):
    """SplitMix64 (Steele, Lea, Flood 2014) output number k of the stream seeded with seed."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (k.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _rmat_numpy(scale, n_edges, a, b, c, seed, batch=1 << 14):
    """The R-MAT draw restated with numpy: all draws of a batch at once, first occurrences
    of distinct undirected codes kept in draw order."""
    ab, abc = a + b, a + b + c
    ss, tt, cc = [], [], []
    k0 = 0
    while True:
        k = np.arange(k0, k0 + batch, dtype=np.uint64)
        u = np.zeros(batch, np.int64)
        v = np.zeros(batch, np.int64)
        for lvl in range(scale):
            x = (_splitmix_at(seed, k * np.uint64(scale) + np.uint64(lvl)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
            u = (u << 1) | (x >= ab)
            v = (v << 1) | (((x >= a) & (x < ab)) | (x >= abc))
        ok = u != v
        ss.append(u[ok]); tt.append(v[ok]); cc.append(np.minimum(u[ok], v[ok]) << scale | np.maximum(u[ok], v[ok]))
        code = np.concatenate(cc)
        _, first = np.unique(code, return_index=True)
        if len(first) >= n_edges:
            first.sort()
            first = first[:n_edges]
            return np.concatenate(ss)[first], np.concatenate(tt)[first]
        k0 += batch


@pytest.mark.parametrize("scale,n_edges,abc,seed", [(12, 30000, (0.57, 0.19, 0.19), 1), (5, 200, (0.45, 0.15, 0.15), 9),
                                                    (16, 100000, (0.25, 0.25, 0.25), 3)])
def test_rmat_generator_matches_numpy_restatement(scale, n_edges, abc, seed):
    s, t = graphs.rmat(scale, n_edges, *abc, seed=seed)
    rs, rt = _rmat_numpy(scale, n_edges, *abc, seed)
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(t, rt)


def test_rmat_generator_rejects_bad_arguments():
    with pytest.raises(ValueError):
        graphs.rmat(0, 10)
    with pytest.raises(ValueError):
        graphs.rmat(4, 100)      # more than half of the 120 pairs


def test_splitmix_restatement_known_answer():
    # SplitMix64 seeded with 0: the first output is the published 0xE220A8397B1DCDAF.
    assert int(_splitmix_at(0, np.array([0], np.uint64))[0]) == 0xE220A8397B1DCDAF


def test_sbm_generator_invariants():
    s, t, block = graphs.sbm(5000, 32, 2.0, 1000)
    g = graphs.edges_to_graph(s, t)
    assert g.m > 0 and len(block) >= 5000 - 64
    same = block[s] == block[t]
    assert same.mean() > 0.8       # p_in >> p_out


def test_coo_to_csr_symmetric():
    s, t = graphs.barabasi_albert(300, 3, 2)
    g = graphs.edges_to_graph(s, t)
    rowptr, col = g.csr()
    assert rowptr[-1] == 2 * g.m
    deg = np.diff(rowptr)
    assert np.array_equal(deg, g.degrees())
    pairs = set()
    for v in range(g.n):
        for u in col[rowptr[v]:rowptr[v + 1]]:
            pairs.add((v, int(u)))
    assert all((u, v) in pairs for v, u in pairs)

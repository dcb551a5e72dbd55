"""Normalized stress on the device against the oracle (hop distances exact, NS to 1e-9)."""
import numpy as np
import pytest

from gpugraphlayout_amd import Engine, graphs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["ba", "sbm", "rmat"])
def test_stress_matches_oracle(oracle, kind):
    print(f"case stress {kind}", flush=True)
    if kind == "ba":
        s, t = graphs.barabasi_albert(3000, 3, 1)
    elif kind == "sbm":
        s, t, _ = graphs.sbm(3000, 8, 2.0, 3)
    else:
        s, t = graphs.rmat(12, 20000, seed=5)     # not connected: some pairs unreachable
    g = graphs.edges_to_graph(s, t)
    pos = graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 7)
    eng = Engine(g.n, g.src, g.dst)
    eng.set_positions(pos)
    npairs = eng.stress_setup(100, 50, seed=11)
    a, b, d = eng.stress_pairs()
    assert len(a) == 5000 and np.all(a != b)
    assert len(np.unique(a)) == 100
    rowptr, col = g.csr()
    for src in np.unique(a)[:20]:
        ref = oracle.bfs(rowptr, col, int(src))
        sel = a == src
        expect = np.where(ref[b[sel]] < 0, 0, ref[b[sel]])
        np.testing.assert_array_equal(d[sel], expect)
    assert npairs == int((d > 0).sum())
    ok = d > 0
    ns_ref, alpha_ref = oracle.normalized_stress(a[ok], b[ok], d[ok], pos)
    ns, alpha = eng.normalized_stress()
    assert ns == pytest.approx(ns_ref, rel=1e-9, abs=1e-12)
    assert alpha == pytest.approx(alpha_ref, rel=1e-9)
    ns2, _ = eng.normalized_stress(pos * 4.0)      # exact scaling in float32
    assert ns2 == pytest.approx(ns, rel=1e-12)
    eng.step(50)
    eng.sync()
    ns3, _ = eng.normalized_stress()
    ns3_ref, _ = oracle.normalized_stress(a[ok], b[ok], d[ok], eng.positions())
    assert ns3 == pytest.approx(ns3_ref, rel=1e-9)
    assert ns3 < ns       # FA2 improves a random layout
    eng.close()

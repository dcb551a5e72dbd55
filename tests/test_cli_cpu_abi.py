"""The C++ host layer (RPGraph::GPUForceAtlas2, GraphLayout, graph_viewer) driven through a CPU
implementation of the engine ABI (oracle/cpu_abi.cpp, test infrastructure), so that its
staging, iteration bookkeeping, snapshots and writers are checked without a GPU: the BIN
snapshots must equal, bit for bit, the oracle's reference-semantics trajectory from the
reference's glibc start."""
import os
import struct
import subprocess

import numpy as np
import pytest

from gpugraphlayout_amd import graphs
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GV = os.path.join(ROOT, "oracle", "_build", "graph_viewer_cpu")
EDGES = os.path.join(ROOT, "tests", "golden", "edgelists", "random_a.txt")


def run(args):
    if not os.path.exists(GV):
        pytest.skip("oracle/_build/graph_viewer_cpu not built (make -C oracle cpu_abi)")
    return subprocess.run([GV] + args, capture_output=True, text=True, timeout=300)


def read_bin(path):
    raw = open(path, "rb").read()
    recs = [struct.unpack("<Iff", raw[12 * k:12 * k + 12]) for k in range(len(raw) // 12)]
    return np.array([r[0] for r in recs]), np.array([[r[1], r[2]] for r in recs], np.float32)


@pytest.mark.parametrize("gravity", ["wg", "sg"])
def test_graph_viewer_snapshots_follow_the_oracle(tmp_path, gravity):
    r = run(["gpu", "12", "3", gravity, "5", "1", "approximate", EDGES, str(tmp_path), "bin"])
    assert r.returncode == 0, r.stderr
    assert "fetched 300 nodes and 1473 edges." in r.stdout
    g = graphs.load_edgelist(EDGES)
    src, dst = g.src, g.dst          # the COO order the host layer stages (ascending smaller id)
    pos = graphs.glibc_random_positions(g.n, 10000.0, 10000.0, seed=1234)
    fprev = np.zeros_like(pos)
    state = np.array([1.0, 1.0, 1.0, 10.0], np.float32)
    for it in range(1, 13):
        pos, fprev, state, _ = orc.fa2_step(pos, fprev, state, src, dst, g.mass(), gravity == "sg", 1.0, 5.0)
        if it % 4 == 0:
            ids, xy = read_bin(tmp_path / f"random_a.txt_{it}.bin")
            assert np.array_equal(ids, g.ids)
            np.testing.assert_array_equal(xy.view(np.uint32), pos.view(np.uint32))


def test_graph_viewer_csv_snapshot(tmp_path):
    r = run(["gpu", "4", "1", "wg", "2", "1", "approximate", EDGES, str(tmp_path), "csv"])
    assert r.returncode == 0, r.stderr
    lines = open(tmp_path / "random_a.txt_4.csv").read().splitlines()
    g = graphs.load_edgelist(EDGES)
    assert len(lines) == g.n
    ids = [int(l.split(",")[0]) for l in lines]
    assert ids == list(g.ids)


def test_graph_viewer_refuses_what_the_engine_refuses(tmp_path):
    # exact mode is not provided by the CPU ABI: the engine error surfaces as a failed run
    r = run(["gpu", "4", "1", "wg", "2", "1", "exact", EDGES, str(tmp_path), "bin"])
    assert r.returncode != 0
    assert "unsupported" in (r.stderr + r.stdout)

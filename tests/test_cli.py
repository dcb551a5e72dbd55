"""graph_viewer: argument handling on the CPU; an end-to-end GPU run whose BIN output must
equal, bit for bit, the same layout driven through the C ABI from Python."""
import os
import struct
import subprocess

import numpy as np
import pytest

from gpugraphlayout_amd import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GV = os.path.join(ROOT, "gpugraphlayout_amd", "graph_viewer")
EDGES = os.path.join(ROOT, "tests", "golden", "edgelists", "random_a.txt")


def run(args):
    return subprocess.run([GV] + args, capture_output=True, text=True, timeout=300)


def test_usage_and_rejections(tmp_path):
    r = run(["gpu", "10"])
    assert r.returncode != 0 and "Usage" in r.stderr
    r = run(["gpu", "10", "1", "wg", "1", "1", "approximate", str(tmp_path / "none.txt"), str(tmp_path), "csv"])
    assert r.returncode != 0 and "No edgelist" in r.stderr
    r = run(["gpu", "10", "1", "wg", "1", "1", "approximate", EDGES, str(tmp_path / "nodir"), "csv"])
    assert r.returncode != 0 and "No output folder" in r.stderr
    r = run(["cpu", "10", "1", "wg", "1", "1", "approximate", EDGES, str(tmp_path), "csv"])
    assert r.returncode != 0


@pytest.mark.gpu
def test_graph_viewer_matches_engine(tmp_path):
    print("case graph_viewer end to end", flush=True)
    r = run(["gpu", "40", "2", "wg", "5", "1", "approximate", EDGES, str(tmp_path), "bin"])
    assert r.returncode == 0, r.stderr
    assert "fetched 300 nodes and 1473 edges." in r.stdout
    for it in (20, 40):
        assert os.path.exists(tmp_path / f"random_a.txt_{it}.bin")
    raw = open(tmp_path / "random_a.txt_40.bin", "rb").read()
    recs = np.array([struct.unpack("<Iff", raw[12 * k:12 * k + 12]) for k in range(len(raw) // 12)])
    xy_cli = recs[:, 1:].astype(np.float32)

    from gpugraphlayout_amd import Engine
    g = graphs.load_edgelist(EDGES)
    pos = graphs.glibc_random_positions(g.n, 10000.0, 10000.0, seed=1234)
    eng = Engine(g.n, g.src, g.dst, mass=g.mass(), scale=5.0, gravity=1.0)
    eng.set_positions(pos)
    eng.step(20)
    eng.sync()
    eng.step(20)
    eng.sync()
    xy = eng.positions()
    eng.close()
    np.testing.assert_array_equal(xy_cli.view(np.uint32), xy.view(np.uint32))
    assert np.array_equal(recs[:, 0].astype(np.int64), g.ids)
    r = run(["gpu", "10", "1", "sg", "2", "1", "approximate", EDGES, str(tmp_path), "png", "300", "200"])
    assert r.returncode == 0, r.stderr
    assert open(tmp_path / "random_a.txt_10.png", "rb").read()[:4] == b"\x89PNG"


@pytest.mark.gpu
def test_graph_viewer_exact_mode_matches_engine(tmp_path):
    """`exact` on the GPU (the reference's GPU backend refuses it) runs the all-pairs
    repulsion; the CLI's layout equals the engine's in exact mode bit for bit."""
    print("case graph_viewer exact", flush=True)
    r = run(["gpu", "20", "1", "wg", "5", "1", "exact", EDGES, str(tmp_path), "bin"])
    assert r.returncode == 0, r.stderr
    raw = open(tmp_path / "random_a.txt_20.bin", "rb").read()
    recs = np.array([struct.unpack("<Iff", raw[12 * k:12 * k + 12]) for k in range(len(raw) // 12)])
    xy_cli = recs[:, 1:].astype(np.float32)
    from gpugraphlayout_amd import Engine
    g = graphs.load_edgelist(EDGES)
    pos = graphs.glibc_random_positions(g.n, 10000.0, 10000.0, seed=1234)
    eng = Engine(g.n, g.src, g.dst, mass=g.mass(), scale=5.0, gravity=1.0, use_barneshut=0)
    eng.set_positions(pos)
    eng.step(20)
    eng.sync()
    xy = eng.positions()
    eng.close()
    np.testing.assert_array_equal(xy_cli.view(np.uint32), xy.view(np.uint32))

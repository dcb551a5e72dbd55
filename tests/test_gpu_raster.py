"""The device rasteriser (gfa2_render) against the host layer's CPU rasteriser, which the CPU
suite pins against a numpy restatement: the same bytes for the same positions."""
import os
import subprocess

import numpy as np
import pytest

from gpugraphlayout_amd import Engine, graphs
from gpugraphlayout_amd.host import HostGraph, png_frame, render_gray

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n,m,w,h", [(500, 3, 200, 120), (20000, 4, 1000, 800)])
def test_device_raster_equals_cpu(n, m, w, h):
    print(f"case raster n={n} image {w}x{h}", flush=True)
    s, t = graphs.barabasi_albert(n, m, 3)
    hg = HostGraph.from_edges(s, t)
    g = hg.arrays()
    eng = Engine(g.n, g.src, g.dst, scale=2.0)
    eng.set_positions(graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), 4))
    eng.step(30)
    eng.sync()
    xy = eng.positions()
    bb = eng.bbox()
    np.testing.assert_array_equal(bb, [xy[:, 0].min(), xy[:, 0].max(), xy[:, 1].min(), xy[:, 1].max()])
    f = png_frame(bb, w, h, g.n, g.m)
    gpu = eng.render(f)
    cpu = render_gray(hg, xy, f)
    np.testing.assert_array_equal(gpu, cpu)
    assert gpu.max() == 255 and gpu.min() < 255
    eng.close()


def test_graph_viewer_png_equals_cpu_writer(tmp_path):
    print("case graph_viewer png", flush=True)
    gv = os.path.join(ROOT, "gpugraphlayout_amd", "graph_viewer")
    edges = os.path.join(ROOT, "tests", "golden", "edgelists", "random_a.txt")
    r = subprocess.run([gv, "gpu", "30", "1", "wg", "5", "1", "approximate", edges, str(tmp_path), "png", "320", "240"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    png_gpu = open(tmp_path / "random_a.txt_30.png", "rb").read()
    hg = HostGraph.load(edges)
    g = hg.arrays()
    pos = graphs.glibc_random_positions(g.n, 10000.0, 10000.0, seed=1234)
    eng = Engine(g.n, g.src, g.dst, mass=g.mass(), scale=5.0, gravity=1.0)
    eng.set_positions(pos)
    eng.step(30)
    eng.sync()
    cpu_png = str(tmp_path / "cpu.png")
    assert hg.write_layout(eng.positions(), cpu_png, "png", 320, 240) == 0
    eng.close()
    assert open(cpu_png, "rb").read() == png_gpu

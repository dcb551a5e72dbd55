"""The allocation-free entry point gfa2_render_noalloc against gfa2_render: the same bytes on
two calls in a row, the second on a smaller graph and a smaller picture, and the scratch
limit refused with GFA2_ERR_SCRATCH_LIMIT before any device work."""
import numpy as np
import pytest

from gpugraphlayout_amd import Engine, GFA2Error, graphs
from gpugraphlayout_amd.host import HostGraph, png_frame

pytestmark = pytest.mark.gpu


def test_render_noalloc_equals_render():
    for n, m, w, h, seed in ((20000, 4, 1000, 800, 3), (500, 3, 200, 120, 5)):
        print(f"case render_noalloc n={n} image {w}x{h}", flush=True)
        s, t = graphs.barabasi_albert(n, m, seed)
        g = HostGraph.from_edges(s, t).arrays()
        eng = Engine(g.n, g.src, g.dst, scale=2.0)
        eng.set_positions(graphs.uniform_positions(g.n, 2 * np.sqrt(g.n), seed + 1))
        eng.step(20)
        eng.sync()
        f = png_frame(eng.bbox(), w, h, g.n, g.m)
        want = eng.render(f)
        got = eng.render_noalloc(f)
        np.testing.assert_array_equal(got, want)
        assert want.max() == 255 and want.min() < 255
        eng.close()


def test_render_noalloc_refuses_pictures_over_its_scratch():
    print("case render_noalloc over the limit", flush=True)
    s, t = graphs.barabasi_albert(100, 2, 1)
    g = HostGraph.from_edges(s, t).arrays()
    eng = Engine(g.n, g.src, g.dst)
    eng.set_positions(graphs.uniform_positions(g.n, 20.0, 2))
    f = png_frame(eng.bbox(), 1025, 1024, g.n, g.m)
    with pytest.raises(GFA2Error) as ei:
        eng.render_noalloc(f)
    assert ei.value.code == -10
    f = png_frame(eng.bbox(), 1024, 1024, g.n, g.m)
    np.testing.assert_array_equal(eng.render_noalloc(f), eng.render(f))
    eng.close()

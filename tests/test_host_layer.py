"""The C++ host layer on the CPU: loader against the reference's UGraph golden vectors,
binary cache round trip, and the CSV / BIN / PNG writers' formats."""
import os
import struct
import zlib

import numpy as np
import pytest

from gpugraphlayout_amd import graphs
from gpugraphlayout_amd.host import HostGraph

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["hand", "chain", "random_a", "random_b"])
def test_cpp_loader_matches_reference(name):
    gold = np.load(os.path.join(GOLDEN, f"ugraph_{name}.npz"))
    hg = HostGraph.load(os.path.join(GOLDEN, "edgelists", f"{name}.txt"))
    g = hg.arrays()
    assert g.n == int(gold["n"]) and g.m == int(gold["m"])
    np.testing.assert_array_equal(g.ids, gold["ids"])
    np.testing.assert_array_equal(hg.deg, gold["deg"])
    np.testing.assert_array_equal(g.src, gold["src"])
    np.testing.assert_array_equal(g.dst, gold["dst"])


def test_cpp_loader_agrees_with_python_on_rmat():
    s, t = graphs.rmat(14, 60000, seed=4)
    # repeat some edges reversed and add self-loops: the loader must drop them
    s2 = np.concatenate([s, t[:500], s[:50]])
    t2 = np.concatenate([t, s[:500], s[:50]])
    a = HostGraph.from_edges(s2, t2).arrays()
    b = graphs.edges_to_graph(s2, t2)
    assert a.n == b.n and a.m == b.m == 60000
    np.testing.assert_array_equal(a.ids, b.ids)
    np.testing.assert_array_equal(a.src, b.src)
    np.testing.assert_array_equal(a.dst, b.dst)


def test_binary_cache_round_trip(tmp_path):
    hg = HostGraph.load(os.path.join(GOLDEN, "edgelists", "random_a.txt"))
    path = str(tmp_path / "g.csrbin")
    hg.save_binary(path)
    a, b = hg.arrays(), HostGraph.load(path).arrays()
    assert a.n == b.n and a.m == b.m
    for k in ("ids", "src", "dst"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k))


def test_csv_and_bin_writers(tmp_path):
    hg = HostGraph.load(os.path.join(GOLDEN, "edgelists", "hand.txt"))
    g = hg.arrays()
    xy = np.array([[1.5, -2.25], [1234567.0, 0.1], [3.0, 4.0], [-0.000125, 7.0], [1e-7, -1e9]], np.float32)
    csv = str(tmp_path / "l.csv")
    assert hg.write_layout(xy, csv, "csv") == 0
    lines = open(csv).read().splitlines()
    assert lines[0] == "5,1.5,-2.25"                    # default stream formatting
    assert lines[1] == "9,1.23457e+06,0.1"
    assert lines[3] == "100,-0.000125,7"
    assert lines[4] == "3,1e-07,-1e+09"
    assert hg.write_layout(xy, csv, "csv") == -1          # refuses to overwrite
    binp = str(tmp_path / "l.bin")
    assert hg.write_layout(xy, binp, "bin") == 0
    raw = open(binp, "rb").read()
    assert len(raw) == 12 * g.n
    recs = [struct.unpack("<Iff", raw[12 * k:12 * k + 12]) for k in range(g.n)]
    assert [r[0] for r in recs] == g.ids.tolist()
    assert np.array_equal(np.array([r[1:] for r in recs], np.float32), xy)


def test_png_writer(tmp_path):
    s, t = graphs.barabasi_albert(400, 2, 1)
    hg = HostGraph.from_edges(s, t)
    xy = graphs.uniform_positions(hg.n, 40, 2)
    path = str(tmp_path / "l.png")
    assert hg.write_layout(xy, path, "png", 320, 200) == 0
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    w, h, depth, ctype = struct.unpack(">IIBB", data[16:26])
    assert (w, h, depth, ctype) == (320, 200, 8, 2)
    idat = data[data.index(b"IDAT") + 4:data.index(b"IEND") - 8]
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(raw[:, 0] == 0)
    px = raw[:, 1:]
    assert px.max() == 255                 # white background remains
    assert 0.01 < (px < 255).mean() < 0.9   # ink from nodes and edges


def _render_numpy(g, xy, f):
    """The PNG picture restated with numpy float32 scalars: frame mapping, discs of radius 3,
    ceil(max(|dx|,|dy|))+1 rounded samples per edge (smaller id first; sample positions in
    float64), per-pixel hit counts,
    brightness round(255 (1-a_n)^k (1-a_e)^l) with powers by repeated squaring."""
    F = np.float32
    W, H = f.width, f.height
    cn = np.zeros((H, W), np.int64)
    ce = np.zeros((H, W), np.int64)

    def rnd(v):  # round half away from zero, as lround
        return int(np.floor(abs(float(v)) + 0.5)) * (1 if v >= 0 else -1)

    def hit(c, x, y):
        if 0 <= x < W and 0 <= y < H:
            c[H - 1 - y, x] += 1

    px = [(F(F(xy[i, 0]) - F(f.x0)) * F(f.sx), F(F(xy[i, 1]) - F(f.y0)) * F(f.sy)) for i in range(g.n)]
    for i in range(g.n):
        cx, cy = rnd(px[i][0]), rnd(px[i][1])
        for dy in range(-3, 4):
            for dx in range(-3, 4):
                if dx * dx + dy * dy <= 9:
                    hit(cn, cx + dx, cy + dy)
    for a, b in zip(g.src, g.dst):
        i, j = min(a, b), max(a, b)
        x0, y0 = px[i]
        x1, y1 = px[j]
        ddx, ddy = F(x1 - x0), F(y1 - y0)
        steps = int(np.ceil(max(abs(ddx), abs(ddy))))
        if steps == 0:
            hit(ce, rnd(x0), rnd(y0))
            continue
        for k in range(steps + 1):
            t = np.float64(k) / np.float64(steps)
            hit(ce, rnd(F(np.float64(x0) + t * np.float64(ddx))), rnd(F(np.float64(y0) + t * np.float64(ddy))))

    def powi(b, k):
        r = F(1.0)
        b = F(b)
        while k:
            if k & 1:
                r = F(r * b)
            b = F(b * b)
            k >>= 1
        return r

    kn, ke = F(F(1.0) - F(f.node_alpha)), F(F(1.0) - F(f.edge_alpha))
    out = np.zeros((H, W), np.uint8)
    for y in range(H):
        for x in range(W):
            out[y, x] = rnd(F(F(255.0) * F(powi(kn, cn[y, x]) * powi(ke, ce[y, x]))))
    return out


def test_cpu_rasteriser_matches_numpy_restatement():
    from gpugraphlayout_amd.host import png_frame, render_gray
    s, t = graphs.barabasi_albert(120, 2, 4)
    hg = HostGraph.from_edges(s, t)
    g = hg.arrays()
    xy = graphs.uniform_positions(g.n, 30, 8)
    bb = (xy[:, 0].min(), xy[:, 0].max(), xy[:, 1].min(), xy[:, 1].max())
    f = png_frame(bb, 64, 40, g.n, g.m)
    # the frame: the bounding box spans the image
    assert abs(float(f.sx) - 64 / (bb[1] - bb[0])) < 1e-4 * float(f.sx)
    assert f.node_alpha == 1.0 and f.edge_alpha == 1.0   # 10000/V and 100000/E exceed 1 here
    got = render_gray(hg, xy, f)
    np.testing.assert_array_equal(got, _render_numpy(g, xy, f))
    # partial opacity: a frame with alphas below one
    f.node_alpha, f.edge_alpha = 0.3, 0.05
    got = render_gray(hg, xy, f)
    np.testing.assert_array_equal(got, _render_numpy(g, xy, f))
    assert got.max() == 255 and got.min() < 100

"""Writes golden vectors from the reference's own code (oracle/_ref/libref.so, compiled
from the reference's src/RPGraph.cpp, src/RPCommon.cpp and src/RPBarnesHutApproximator.cpp
by oracle/Makefile). Run in a checkout that holds the reference; the fixtures it writes
(edge lists and .npz arrays, data only) let the tests pin the loader, the initial layout
and the CPU Barnes-Hut everywhere else.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

EDGELISTS = {
    # comments, self-loops, repeated edges in both orientations, sparse ids
    "hand.txt": "# a comment line\n5 9\n9 5\n7 7\n9 12\n12 5\n# another\n100 5\n5 100\n12 9\n3 100\n",
    "chain.txt": "1 2\n2 3\n3 4\n4 5\n5 1\n",
}


def random_edgelist(seed, n, m):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m) * 7 + 3
    t = rng.integers(0, n, m) * 7 + 3
    dup = rng.integers(0, m, m // 5)
    s = np.concatenate([s, t[dup]])
    t = np.concatenate([t, s[dup]])
    return "".join(f"{a} {b}\n" for a, b in zip(s.tolist(), t.tolist()))


def main():
    orc.build(ref=True)
    if not orc.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (needs the reference checkout)")
    edge_dir = os.path.join(HERE, "edgelists")
    os.makedirs(edge_dir, exist_ok=True)
    lists = dict(EDGELISTS)
    lists["random_a.txt"] = random_edgelist(1, 300, 1500)
    lists["random_b.txt"] = random_edgelist(2, 60, 400)
    for name, text in lists.items():
        path = os.path.join(edge_dir, name)
        with open(path, "w") as fh:
            fh.write(text)
        n, m, ids, deg, src, dst = orc.ref_load(path)
        np.savez(os.path.join(HERE, f"ugraph_{name[:-4]}.npz"), n=n, m=m, ids=ids, deg=deg, src=src, dst=dst)
        print(f"{name}: n={n} m={m}")

    # initial layout draws of the reference (srandom + get_random)
    R = orc.ref()
    for n, w in ((1000, 63.245552), (7, 10000.0)):
        xy = np.zeros((n, 2), np.float32)
        R.ref_random_positions(1234, n, orc.F(w), orc.F(w), orc._p(xy))
        np.savez(os.path.join(HERE, f"init_{n}.npz"), seed=1234, n=n, w=np.float32(w), xy=xy)
        print(f"init_{n}: first {xy[0]}")

    # the reference's CPU Barnes-Hut forces
    rng = np.random.default_rng(3)
    cases = {
        "uniform": (rng.random((400, 2)) * 100 - 50).astype(np.float32),
        "clustered": np.concatenate([rng.normal(0, 0.5, (150, 2)), rng.normal(20, 3, (150, 2))]).astype(np.float32),
    }
    dup = cases["uniform"].copy()
    dup[10] = dup[11]       # coincident pair (merged by the reference)
    cases["coincident"] = dup
    for name, pos in cases.items():
        mass = rng.integers(1, 20, len(pos)).astype(np.float32)
        span = np.float32(np.ceil(max(pos[:, 0].max() - pos[:, 0].min(), pos[:, 1].max() - pos[:, 1].min())))
        cx = np.float32(pos[:, 0].min() + (pos[:, 0].max() - pos[:, 0].min()) / 2.0)
        cy = np.float32(pos[:, 1].min() + (pos[:, 1].max() - pos[:, 1].min()) / 2.0)
        length = np.float32(span + 10)
        f = orc.cpu_bh_forces(pos, mass, cx, cy, length, 1.0, pos, mass, use_ref=True)
        np.savez(os.path.join(HERE, f"cpu_bh_{name}.npz"), pos=pos, mass=mass, cx=cx, cy=cy, length=length,
                 theta=np.float32(1.0), forces=f)
        print(f"cpu_bh_{name}: |f|max={np.abs(f).max():.3e}")


if __name__ == "__main__":
    main()

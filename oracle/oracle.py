"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liboracle.so (the CPU restatement of the reference's GPU FA2
iteration, the FP64 exact pairwise repulsion, the CPU-baseline restatement and the
layout-quality references) and, where it was built, of oracle/_ref/libref.so (the
reference's own loader and CPU Barnes-Hut compiled from its sources). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

_P = ctypes.c_void_p
_lib = None
_ref = None


def build(ref: bool = False):
    """Compiles the oracle (and, if asked and the reference checkout exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_bh_repulsion.restype = ctypes.c_int
        L.orc_bh_repulsion_groups.restype = ctypes.c_int
        L.orc_tree.restype = ctypes.c_int
        L.orc_fa2_step.restype = ctypes.c_int
        L.orc_cpu_create.restype = ctypes.c_void_p
        L.orc_cpu_create.argtypes = [ctypes.c_int, ctypes.c_longlong, _P, _P, _P, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_int]
        L.orc_cpu_step.argtypes = [_P, ctypes.c_int]
        L.orc_cpu_get_positions.argtypes = [_P, _P]
        L.orc_cpu_get_state.argtypes = [_P, _P]
        L.orc_cpu_destroy.argtypes = [_P]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_LIB)
    return _ref


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


F = ctypes.c_float


def gravity(pos, mass, strong, k_g, f=None):
    n = len(mass)
    f = np.zeros((n, 2), np.float32) if f is None else f32(f).copy()
    lib().orc_gravity(n, _p(f32(pos)), _p(f32(mass)), int(strong), F(k_g), _p(f))
    return f


def attraction_coo(pos, src, dst, f=None):
    n = len(pos)
    f = np.zeros((n, 2), np.float32) if f is None else f32(f).copy()
    lib().orc_attraction_coo(ctypes.c_longlong(len(src)), _p(i32(src)), _p(i32(dst)), _p(f32(pos)), _p(f))
    return f


def attraction_general(pos, src, dst, w=None, delta=0.0, linlog=False, f=None):
    n = len(pos)
    f = np.zeros((n, 2), np.float32) if f is None else f32(f).copy()
    wp = None if w is None else f32(w)
    lib().orc_attraction_general(ctypes.c_longlong(len(src)), _p(i32(src)), _p(i32(dst)), _p(wp),
                                 ctypes.c_float(delta), int(bool(linlog)), _p(f32(pos)), _p(f))
    return f


def bbox(pos):
    out = np.zeros(3, np.float32)
    lib().orc_bbox(len(pos), _p(f32(pos)), _p(out))
    return out


def tree(pos, mass, max_levels=128, cap=None):
    """Reference quadtree: cells in pre-order (level, count, mass, centroid), permutation."""
    n = len(mass)
    pos = f32(pos).copy()
    cap = cap or 4 * n + 64
    perm = np.zeros(n, np.int32)
    lvl = np.zeros(cap, np.int32)
    cnt = np.zeros(cap, np.int32)
    cm = np.zeros(cap, np.float32)
    cp = np.zeros((cap, 2), np.float32)
    md = ctypes.c_int(0)
    root = np.zeros(3, np.float32)
    c = lib().orc_tree(n, _p(pos), _p(f32(mass)), max_levels, cap, _p(perm), _p(lvl), _p(cnt), _p(cm), _p(cp),
                       ctypes.byref(md), _p(root))
    if c < 0:
        raise RuntimeError(f"orc_tree failed: {c}")
    return {"level": lvl[:c], "count": cnt[:c], "mass": cm[:c], "xy": cp[:c], "perm": perm,
            "max_depth": md.value, "root": root, "pos": pos}


def bh_repulsion(pos, mass, k_r, theta=1.0, epssq=None, group=32, max_levels=128):
    n = len(mass)
    pos = f32(pos).copy()
    rep = np.zeros((n, 2), np.float32)
    perm = np.zeros(n, np.int32)
    inter = ctypes.c_longlong(0)
    eps = np.float32(0.05 * 0.05) if epssq is None else np.float32(epssq)
    itolsq = np.float32(1.0) / (np.float32(theta) * np.float32(theta))
    md = lib().orc_bh_repulsion(n, _p(pos), _p(f32(mass)), F(k_r), F(itolsq), F(eps), group, max_levels,
                                _p(rep), _p(perm), ctypes.byref(inter))
    if md < 0:
        raise RuntimeError("oracle tree exceeded max_levels")
    return rep, {"perm": perm, "max_depth": md, "interactions": inter.value}


def bh_repulsion_groups(pos, mass, k_r, starts, theta=1.0, epssq=None, group=32, max_levels=128):
    """bh_repulsion for the groups starting at the given sorted positions only. Returns the
    repulsion (zero outside those groups), the permutation and the groups' interaction count."""
    n = len(mass)
    pos = f32(pos).copy()
    rep = np.zeros((n, 2), np.float32)
    perm = np.zeros(n, np.int32)
    inter = ctypes.c_longlong(0)
    eps = np.float32(0.05 * 0.05) if epssq is None else np.float32(epssq)
    itolsq = np.float32(1.0) / (np.float32(theta) * np.float32(theta))
    starts = i32(starts)
    md = lib().orc_bh_repulsion_groups(n, _p(pos), _p(f32(mass)), F(k_r), F(itolsq), F(eps), group, max_levels,
                                       len(starts), _p(starts), _p(rep), _p(perm), ctypes.byref(inter))
    if md < 0:
        raise RuntimeError("oracle tree exceeded max_levels")
    return rep, {"perm": perm, "max_depth": md, "interactions": inter.value}


def speed(f, fprev, mass, state):
    st = f32(state).copy()
    out = np.zeros(2, np.float64)
    lib().orc_speed(len(mass), _p(f32(f)), _p(f32(fprev)), _p(f32(mass)), _p(st), _p(out))
    return st, out


def speed_update(n, S, T, state):
    st = f32(state).copy()
    lib().orc_speed_update(n, ctypes.c_double(S), ctypes.c_double(T), _p(st))
    return st


def displace(gs, f, fprev, pos):
    fp = f32(fprev).copy()
    p = f32(pos).copy()
    lib().orc_displace(len(p), F(gs), _p(f32(f)), _p(fp), _p(p))
    return p, fp


def fa2_step(pos, fprev, state, src, dst, mass, strong, k_g, k_r, theta=1.0, epssq=None, group=32,
             max_levels=128):
    """One reference-semantics iteration. Returns (pos, fprev, state, info)."""
    n = len(mass)
    pos = f32(pos).copy()
    fprev = f32(fprev).copy()
    st = f32(state).copy()
    out_f = np.zeros((n, 2), np.float32)
    out_rep = np.zeros((n, 2), np.float32)
    out_st = np.zeros(2, np.float64)
    inter = ctypes.c_longlong(0)
    eps = np.float32(0.05 * 0.05) if epssq is None else np.float32(epssq)
    itolsq = np.float32(1.0) / (np.float32(theta) * np.float32(theta))
    md = lib().orc_fa2_step(n, ctypes.c_longlong(len(src)), _p(i32(src)), _p(i32(dst)), _p(f32(mass)),
                            int(strong), F(k_g), F(k_r), F(itolsq), F(eps), group, max_levels, _p(pos),
                            _p(fprev), _p(st), _p(out_f), _p(out_rep), _p(out_st), ctypes.byref(inter))
    if md < 0:
        raise RuntimeError("oracle tree exceeded max_levels")
    return pos, fprev, st, {"f": out_f, "rep": out_rep, "S": out_st[0], "T": out_st[1],
                            "interactions": inter.value, "max_depth": md}


def exact_repulsion_f64(pos, mass, k_r, sample, epssq=None, threads=None):
    eps = float(np.float32(0.05 * 0.05)) if epssq is None else float(epssq)
    sample = i32(sample)
    out = np.zeros((len(sample), 2), np.float64)
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    lib().orc_exact_repulsion_f64(len(mass), _p(f32(pos)), _p(f32(mass)), ctypes.c_double(k_r), ctypes.c_double(eps),
                                  len(sample), _p(sample), _p(out), int(threads))
    return out


def bfs(rowptr, col, src):
    n = len(rowptr) - 1
    d = np.zeros(n, np.int32)
    lib().orc_bfs(n, _p(i32(rowptr)), _p(i32(col)), int(src), _p(d))
    return d


def normalized_stress(a, b, d, pos):
    out = np.zeros(2, np.float64)
    lib().orc_normalized_stress(ctypes.c_longlong(len(a)), _p(i32(a)), _p(i32(b)), _p(i32(d)), _p(f32(pos)),
                                _p(out))
    return float(out[0]), float(out[1])


class CPUBaseline:
    """Restated serial CPUForceAtlas2 (the timed CPU baseline)."""

    def __init__(self, n, src, dst, pos, strong, k_g, k_r, use_bh=True):
        self._s, self._d, self._p0 = i32(src), i32(dst), f32(pos)
        self.n = n
        self.h = lib().orc_cpu_create(n, len(self._s), _p(self._s), _p(self._d), _p(self._p0), int(strong),
                                      F(k_g), F(k_r), int(use_bh))

    def step(self, iters=1):
        lib().orc_cpu_step(self.h, int(iters))

    def positions(self):
        out = np.zeros((self.n, 2), np.float32)
        lib().orc_cpu_get_positions(self.h, _p(out))
        return out

    def state(self):
        """{global speed, speed efficiency, jitter tolerance, k_s_max, S, T of the last iteration}."""
        out = np.zeros(6, np.float32)
        lib().orc_cpu_get_state(self.h, _p(out))
        return out

    def close(self):
        if self.h:
            lib().orc_cpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def cpu_bh_forces(pos, mass, cx, cy, length, theta, qpos, qmass, use_ref=False):
    L = ref() if use_ref else lib()
    fn = L.ref_bh_forces if use_ref else L.orc_cpu_bh_forces
    out = np.zeros((len(qmass), 2), np.float32)
    fn(len(mass), _p(f32(pos)), _p(f32(mass)), F(cx), F(cy), F(length), F(theta), len(qmass), _p(f32(qpos)),
       _p(f32(qmass)), _p(out))
    return out


def ref_load(path):
    """The reference's UGraph on an edge-list file: (n, m, file ids, degrees, COO src, dst)."""
    R = ref()
    n = ctypes.c_int(0)
    m = ctypes.c_longlong(0)
    R.ref_load(path.encode(), ctypes.byref(n), ctypes.byref(m))
    ids = np.zeros(n.value, np.int64)
    deg = np.zeros(n.value, np.int32)
    src = np.zeros(m.value, np.int32)
    dst = np.zeros(m.value, np.int32)
    R.ref_node_ids(_p(ids))
    R.ref_degrees(_p(deg))
    R.ref_coo(_p(src), _p(dst))
    return n.value, m.value, ids, deg, src, dst

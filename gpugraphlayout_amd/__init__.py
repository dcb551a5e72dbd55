"""MI355X-native ForceAtlas2 engine: Python access to libgfa2.so through its C ABI.

The engine itself is HIP (``csrc/*.hip``) behind ``include/gfa2.h``; this module only
binds that ABI with ctypes for tests, benchmarks and scripts. There is no fallback: if
the shared library is missing or no HIP device is present, calls raise.

Reference interface mirrored (govertb/GPUGraphLayout): ``RPGraph::ForceAtlas2``
(src/RPForceAtlas2.hpp:32-60) and ``CUDAForceAtlas2`` (src/RPGPUForceAtlas2.hpp:30-68):
``do_step``/``do_steps``, ``sync_layout`` (positions out), ``set_scale``/``set_gravity``
via parameters.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgfa2.so")

GFA2_FORCE_TOTAL = 0
GFA2_FORCE_REPULSION = 1
GFA2_FORCE_ATTRACTION_GRAVITY = 2

# Every symbol include/gfa2.h declares.
ABI_SYMBOLS = (
    "gfa2_default_params", "gfa2_create", "gfa2_set_graph", "gfa2_set_graph_coo",
    "gfa2_set_positions", "gfa2_get_positions", "gfa2_reset", "gfa2_step", "gfa2_sync",
    "gfa2_get_forces", "gfa2_set_capture", "gfa2_get_tree", "gfa2_get_stats",
    "gfa2_set_timing", "gfa2_get_timing", "gfa2_set_graph_capture", "gfa2_destroy",
    "gfa2_strerror", "gfa2_device_info", "gfa2_set_params",
    "gfa2_stress_setup", "gfa2_stress_pairs", "gfa2_normalized_stress",
    "gfa2_set_graph_batch", "gfa2_get_batch_stats",
    "gfa2_autotune", "gfa2_get_launch_config", "gfa2_set_launch_config", "gfa2_tuned_config",
    "gfa2_set_edge_weights", "gfa2_set_graph_coo_weighted", "gfa2_exact_repulsion",
    "gfa2_get_bbox", "gfa2_render", "gfa2_render_noalloc",
)


class Params(ctypes.Structure):
    _fields_ = [
        ("use_barneshut", ctypes.c_int32),
        ("strong_gravity", ctypes.c_int32),
        ("gravity", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("theta", ctypes.c_float),
        ("epssq", ctypes.c_float),
        ("jitter_tolerance", ctypes.c_float),
        ("k_s_max", ctypes.c_float),
        ("global_speed", ctypes.c_float),
        ("speed_efficiency", ctypes.c_float),
        ("group_size", ctypes.c_int32),
        ("cell_factor", ctypes.c_int32),
        ("linlog", ctypes.c_int32),
        ("edge_weight_influence", ctypes.c_float),
        ("prevent_overlap", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("iteration", ctypes.c_int32),
        ("nodes", ctypes.c_int32),
        ("edges", ctypes.c_int64),
        ("global_speed", ctypes.c_float),
        ("speed_efficiency", ctypes.c_float),
        ("swing", ctypes.c_double),
        ("traction", ctypes.c_double),
        ("interactions", ctypes.c_int64),
        ("interactions_total", ctypes.c_int64),
        ("cells", ctypes.c_int32),
        ("max_level", ctypes.c_int32),
        ("root_x", ctypes.c_float),
        ("root_y", ctypes.c_float),
        ("radius", ctypes.c_float),
        ("error_flags", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class LaunchConfig(ctypes.Structure):
    _fields_ = [
        ("walk_block", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("attraction_stream", ctypes.c_int32),
        ("attraction_blocks_per_cu", ctypes.c_int32),
        ("ms_per_iteration", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Raster(ctypes.Structure):
    """Frame of the PNG picture (gfa2_raster)."""
    _fields_ = [
        ("x0", ctypes.c_float), ("y0", ctypes.c_float), ("sx", ctypes.c_float), ("sy", ctypes.c_float),
        ("node_alpha", ctypes.c_float), ("edge_alpha", ctypes.c_float),
        ("width", ctypes.c_int32), ("height", ctypes.c_int32),
    ]


class GFA2Error(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {lib().gfa2_strerror(code).decode()} ({code})")


_lib: Optional[ctypes.CDLL] = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64


def lib() -> ctypes.CDLL:
    """Loads libgfa2.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    L.gfa2_default_params.argtypes = [ctypes.POINTER(Params)]
    L.gfa2_default_params.restype = None
    L.gfa2_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(_P)]
    L.gfa2_set_params.argtypes = [_P, ctypes.POINTER(Params)]
    L.gfa2_set_graph.argtypes = [_P, _I32, _P, _P, _P]
    L.gfa2_set_graph_coo.argtypes = [_P, _I32, _I64, _P, _P, _P]
    L.gfa2_set_positions.argtypes = [_P, _P]
    L.gfa2_get_positions.argtypes = [_P, _P]
    L.gfa2_reset.argtypes = [_P]
    L.gfa2_step.argtypes = [_P, _I32]
    L.gfa2_sync.argtypes = [_P]
    L.gfa2_get_forces.argtypes = [_P, _I32, _P]
    L.gfa2_set_capture.argtypes = [_P, _I32]
    L.gfa2_get_tree.argtypes = [_P, _I32, _P, _P, _P, _P, _P, _P]
    L.gfa2_get_stats.argtypes = [_P, ctypes.POINTER(Stats)]
    L.gfa2_set_timing.argtypes = [_P, _I32]
    L.gfa2_get_timing.argtypes = [_P, _I32, _P, ctypes.POINTER(ctypes.c_char_p), _P]
    L.gfa2_set_graph_capture.argtypes = [_P, _I32]
    L.gfa2_destroy.argtypes = [_P]
    L.gfa2_strerror.argtypes = [ctypes.c_int]
    L.gfa2_strerror.restype = ctypes.c_char_p
    L.gfa2_device_info.argtypes = [ctypes.c_int, _P, _P, _P]
    L.gfa2_stress_setup.argtypes = [_P, _I32, _I32, ctypes.c_uint64, _P]
    L.gfa2_stress_pairs.argtypes = [_P, _I64, _P, _P, _P]
    L.gfa2_normalized_stress.argtypes = [_P, _P, _P, _P]
    L.gfa2_set_graph_batch.argtypes = [_P, _I32, _P, _P, _P, _P]
    L.gfa2_get_batch_stats.argtypes = [_P, _I32, _P, _P, _P, _P, _P]
    L.gfa2_autotune.argtypes = [_P, _I32, ctypes.POINTER(LaunchConfig)]
    L.gfa2_get_launch_config.argtypes = [_P, ctypes.POINTER(LaunchConfig)]
    L.gfa2_set_launch_config.argtypes = [_P, ctypes.POINTER(LaunchConfig)]
    L.gfa2_tuned_config.argtypes = [_I32, ctypes.POINTER(LaunchConfig)]
    L.gfa2_set_edge_weights.argtypes = [_P, _I64, _P]
    L.gfa2_set_graph_coo_weighted.argtypes = [_P, _I32, _I64, _P, _P, _P, _P]
    L.gfa2_exact_repulsion.argtypes = [_P, _I32, _P]
    L.gfa2_get_bbox.argtypes = [_P, _P]
    L.gfa2_render.argtypes = [_P, ctypes.POINTER(Raster), _P]
    L.gfa2_render_noalloc.argtypes = [_P, ctypes.POINTER(Raster), _P]
    for name in ABI_SYMBOLS:
        fn = getattr(L, name)
        if name not in ("gfa2_default_params", "gfa2_strerror"):
            fn.restype = ctypes.c_int
    _lib = L
    return L


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int, what: str):
    if rc != 0:
        raise GFA2Error(rc, what)


def default_params(**kw) -> Params:
    p = Params()
    lib().gfa2_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def coo_to_csr(n: int, src: np.ndarray, dst: np.ndarray):
    """Symmetric CSR (int64 rowptr, int32 col) from an undirected COO list, each edge once.
    Row order follows the COO order, as gfa2_set_graph_coo builds it."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    rows = np.concatenate([np.stack([src, dst], 1), np.stack([dst, src], 1)], 0)
    # within a row, entries keep the COO order of their edge
    key = np.concatenate([np.arange(len(src)) * 2, np.arange(len(src)) * 2 + 1])
    order = np.lexsort((key, rows[:, 0]))
    r = rows[order]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowptr, r[:, 0] + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int64)
    return rowptr, r[:, 1].astype(np.int32)


class Engine:
    """One layout instance on one GPU (gfa2_handle)."""

    def __init__(self, n: int, src=None, dst=None, rowptr=None, col=None, mass=None, device: int = 0,
                 weights=None, **params):
        self.params = default_params(**params)
        h = ctypes.c_void_p()
        _check(lib().gfa2_create(ctypes.byref(self.params), device, ctypes.byref(h)), "gfa2_create")
        self.h = h
        self.n = int(n)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float32)
        if rowptr is not None:
            self._rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
            self._col = np.ascontiguousarray(col, dtype=np.int32)
            _check(lib().gfa2_set_graph(self.h, self.n, _ptr(self._rowptr), _ptr(self._col), _ptr(m)),
                   "gfa2_set_graph")
        else:
            s = np.ascontiguousarray(src, dtype=np.int32)
            d = np.ascontiguousarray(dst, dtype=np.int32)
            if weights is not None:
                w = np.ascontiguousarray(weights, dtype=np.float32)
                _check(lib().gfa2_set_graph_coo_weighted(self.h, self.n, len(s), _ptr(s), _ptr(d), _ptr(w), _ptr(m)),
                       "gfa2_set_graph_coo_weighted")
            else:
                _check(lib().gfa2_set_graph_coo(self.h, self.n, len(s), _ptr(s), _ptr(d), _ptr(m)),
                       "gfa2_set_graph_coo")

    def close(self):
        if getattr(self, "h", None):
            lib().gfa2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_positions(self, xy: np.ndarray):
        a = np.ascontiguousarray(xy, dtype=np.float32).reshape(self.n, 2)
        _check(lib().gfa2_set_positions(self.h, _ptr(a)), "gfa2_set_positions")

    def positions(self) -> np.ndarray:
        out = np.empty((self.n, 2), dtype=np.float32)
        _check(lib().gfa2_get_positions(self.h, _ptr(out)), "gfa2_get_positions")
        return out

    sync_layout = positions

    def reset(self):
        _check(lib().gfa2_reset(self.h), "gfa2_reset")

    def step(self, iters: int = 1):
        _check(lib().gfa2_step(self.h, int(iters)), "gfa2_step")

    do_steps = step

    def do_step(self):
        self.step(1)

    def sync(self):
        _check(lib().gfa2_sync(self.h), "gfa2_sync")

    def forces(self, kind: int = GFA2_FORCE_TOTAL) -> np.ndarray:
        out = np.empty((self.n, 2), dtype=np.float32)
        _check(lib().gfa2_get_forces(self.h, int(kind), _ptr(out)), "gfa2_get_forces")
        return out

    def set_capture(self, on: bool = True):
        _check(lib().gfa2_set_capture(self.h, int(bool(on))), "gfa2_set_capture")

    def set_graph_capture(self, on: bool = True):
        _check(lib().gfa2_set_graph_capture(self.h, int(bool(on))), "gfa2_set_graph_capture")

    def tree(self):
        """Cells in pre-order: dict of level, count, mass, xy; plus the in-order permutation."""
        nc = ctypes.c_int32(0)
        _check(lib().gfa2_get_tree(self.h, 0, ctypes.byref(nc), None, None, None, None, None), "gfa2_get_tree")
        c = nc.value
        level = np.empty(c, np.int32)
        count = np.empty(c, np.int32)
        mass = np.empty(c, np.float32)
        xy = np.empty((c, 2), np.float32)
        perm = np.empty(self.n, np.int32)
        _check(lib().gfa2_get_tree(self.h, c, ctypes.byref(nc), _ptr(level), _ptr(count), _ptr(mass), _ptr(xy),
                                   _ptr(perm)), "gfa2_get_tree")
        return {"level": level, "count": count, "mass": mass, "xy": xy, "perm": perm}

    def stats(self) -> dict:
        s = Stats()
        _check(lib().gfa2_get_stats(self.h, ctypes.byref(s)), "gfa2_get_stats")
        return s.as_dict()

    def stress_setup(self, nsources: int = 1024, targets_per_source: int = 977, seed: int = 0) -> int:
        """Samples the pairs normalized stress is measured on (10^6 by default); returns the
        number of reachable pairs."""
        npairs = ctypes.c_int64(0)
        _check(lib().gfa2_stress_setup(self.h, int(nsources), int(targets_per_source), int(seed),
                                       ctypes.byref(npairs)), "gfa2_stress_setup")
        self._stress_np = int(min(nsources, self.n)) * int(targets_per_source)
        return npairs.value

    def stress_pairs(self):
        n = self._stress_np
        a, b, d = (np.empty(n, np.int32) for _ in range(3))
        _check(lib().gfa2_stress_pairs(self.h, n, _ptr(a), _ptr(b), _ptr(d)), "gfa2_stress_pairs")
        return a, b, d

    def normalized_stress(self, xy: Optional[np.ndarray] = None):
        """(NS, alpha) of the device layout, or of the given positions."""
        ns = ctypes.c_double(0)
        alpha = ctypes.c_double(0)
        p = None if xy is None else np.ascontiguousarray(xy, dtype=np.float32)
        _check(lib().gfa2_normalized_stress(self.h, _ptr(p), ctypes.byref(ns), ctypes.byref(alpha)),
               "gfa2_normalized_stress")
        return ns.value, alpha.value

    def set_params(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.params, k):
                raise KeyError(k)
            setattr(self.params, k, v)
        _check(lib().gfa2_set_params(self.h, ctypes.byref(self.params)), "gfa2_set_params")

    def set_edge_weights(self, w: Optional[np.ndarray]):
        """Weights per CSR entry (None clears them)."""
        if w is None:
            _check(lib().gfa2_set_edge_weights(self.h, 0, None), "gfa2_set_edge_weights")
            return
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().gfa2_set_edge_weights(self.h, len(w), _ptr(w)), "gfa2_set_edge_weights")

    def exact_repulsion(self, fp64: bool = True) -> np.ndarray:
        """All-pairs repulsion of the current positions (V x 2, float64)."""
        out = np.empty((self.n, 2), np.float64)
        _check(lib().gfa2_exact_repulsion(self.h, int(bool(fp64)), _ptr(out)), "gfa2_exact_repulsion")
        return out

    def bbox(self) -> np.ndarray:
        """{min x, max x, min y, max y} of the device positions."""
        out = np.zeros(4, np.float32)
        _check(lib().gfa2_get_bbox(self.h, _ptr(out)), "gfa2_get_bbox")
        return out

    def render(self, frame: "Raster") -> np.ndarray:
        """The PNG picture rasterised on the device: height x width grey bytes, top row first."""
        out = np.zeros((frame.height, frame.width), np.uint8)
        _check(lib().gfa2_render(self.h, ctypes.byref(frame), _ptr(out)), "gfa2_render")
        return out

    def render_noalloc(self, frame: "Raster") -> np.ndarray:
        """gfa2_render through the allocation-free entry point (fixed device scratch)."""
        out = np.zeros((frame.height, frame.width), np.uint8)
        _check(lib().gfa2_render_noalloc(self.h, ctypes.byref(frame), _ptr(out)), "gfa2_render_noalloc")
        return out

    def autotune(self, iters: int = 10) -> dict:
        """Times every launch shape over `iters` iterations and keeps the fastest; the
        layout state is restored afterwards and no shape changes a result bit."""
        c = LaunchConfig()
        _check(lib().gfa2_autotune(self.h, int(iters), ctypes.byref(c)), "gfa2_autotune")
        return c.as_dict()

    def launch_config(self) -> dict:
        c = LaunchConfig()
        _check(lib().gfa2_get_launch_config(self.h, ctypes.byref(c)), "gfa2_get_launch_config")
        return c.as_dict()

    def set_launch_config(self, walk_block: int, blocks_per_cu: int, attraction_stream: int = 0,
                          attraction_blocks_per_cu: int = 0):
        """attraction_stream 0: gravity + attraction on the instance's stream; 1: on a second
        stream beside the walk; 2: on a second stream beside keys, sort and tree, with
        attraction_blocks_per_cu blocks per compute unit (0: one per work item). No setting
        changes a result bit."""
        c = LaunchConfig(int(walk_block), int(blocks_per_cu), int(attraction_stream), int(attraction_blocks_per_cu),
                         0.0)
        _check(lib().gfa2_set_launch_config(self.h, ctypes.byref(c)), "gfa2_set_launch_config")

    def set_timing(self, on=True):
        """True (or 1): events around every kernel group; "walk" (or 2): only around the
        force walk, the least added work; False: off."""
        mode = 2 if on == "walk" else int(on)
        _check(lib().gfa2_set_timing(self.h, mode), "gfa2_set_timing")

    def timing(self) -> dict:
        names = (ctypes.c_char_p * 32)()
        ms = np.zeros(32, dtype=np.float64)
        cnt = ctypes.c_int32(0)
        _check(lib().gfa2_get_timing(self.h, 32, ctypes.byref(cnt), names, _ptr(ms)), "gfa2_get_timing")
        return {names[i].decode(): float(ms[i]) for i in range(cnt.value)}


def device_info(device: int = 0) -> dict:
    cus = ctypes.c_int32(0)
    wave = ctypes.c_int32(0)
    mem = ctypes.c_int64(0)
    _check(lib().gfa2_device_info(device, ctypes.byref(cus), ctypes.byref(wave), ctypes.byref(mem)),
           "gfa2_device_info")
    return {"cus": cus.value, "wave": wave.value, "mem_bytes": mem.value}


class BatchEngine(Engine):
    """Many independent graphs laid out together (gfa2_set_graph_batch). graphs: a list of
    (n, src, dst) with local ids; positions are concatenated in list order."""

    def __init__(self, graphs, masses=None, device: int = 0, **params):
        self.params = default_params(**params)
        h = ctypes.c_void_p()
        _check(lib().gfa2_create(ctypes.byref(self.params), device, ctypes.byref(h)), "gfa2_create")
        self.h = h
        sizes = np.array([int(g[0]) for g in graphs], np.int64)
        self.node_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        self.n = int(self.node_off[-1])
        self.ngraphs = len(graphs)
        rows, cols = [np.zeros(1, np.int64)], []
        base_e = 0
        for k, (n, s, d) in enumerate(graphs):
            rp, c = coo_to_csr(int(n), s, d)
            rows.append(rp[1:] + base_e)
            cols.append(c.astype(np.int64) + self.node_off[k])
            base_e += int(rp[-1])
        self._rowptr = np.ascontiguousarray(np.concatenate(rows), dtype=np.int64)
        self._col = np.ascontiguousarray(np.concatenate(cols) if cols else np.zeros(0), dtype=np.int32)
        m = None if masses is None else np.ascontiguousarray(np.concatenate(masses), dtype=np.float32)
        _check(lib().gfa2_set_graph_batch(self.h, self.ngraphs, _ptr(self.node_off), _ptr(self._rowptr),
                                          _ptr(self._col), _ptr(m)), "gfa2_set_graph_batch")

    def batch_stats(self) -> dict:
        G = self.ngraphs
        speed = np.zeros(G, np.float32)
        eff = np.zeros(G, np.float32)
        swing = np.zeros(G, np.float64)
        trac = np.zeros(G, np.float64)
        inter = np.zeros(G, np.int64)
        _check(lib().gfa2_get_batch_stats(self.h, G, _ptr(speed), _ptr(eff), _ptr(swing), _ptr(trac), _ptr(inter)),
               "gfa2_get_batch_stats")
        return {"speed": speed, "efficiency": eff, "swing": swing, "traction": trac, "interactions": inter}

    def split(self, xy):
        return [xy[self.node_off[k]:self.node_off[k + 1]] for k in range(self.ngraphs)]


MPNN_SYMBOLS = (
    "gfa2_mpnn_num_params", "gfa2_mpnn_create", "gfa2_mpnn_get_params", "gfa2_mpnn_set_params",
    "gfa2_mpnn_add_graph", "gfa2_mpnn_add_engine_graph", "gfa2_mpnn_release_graph", "gfa2_mpnn_features", "gfa2_mpnn_loss_grad", "gfa2_mpnn_train",
    "gfa2_mpnn_target_scale", "gfa2_mpnn_infer", "gfa2_mpnn_infer_to_engine", "gfa2_mpnn_destroy",
)
ABI_SYMBOLS = ABI_SYMBOLS + MPNN_SYMBOLS


def _bind_mpnn(L):
    if getattr(L, "_mpnn_bound", False):
        return L
    L.gfa2_mpnn_num_params.argtypes = []
    L.gfa2_mpnn_num_params.restype = ctypes.c_int32
    L.gfa2_mpnn_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_P)]
    L.gfa2_mpnn_get_params.argtypes = [_P, _P]
    L.gfa2_mpnn_set_params.argtypes = [_P, _P]
    L.gfa2_mpnn_add_graph.argtypes = [_P, _I32, _P, _P, _P, ctypes.c_uint64, _I32, _P]
    L.gfa2_mpnn_add_engine_graph.argtypes = [_P, _P, ctypes.c_uint64, _I32, _P]
    L.gfa2_mpnn_release_graph.argtypes = [_P, _I32]
    L.gfa2_mpnn_features.argtypes = [_P, _I32, _P]
    L.gfa2_mpnn_loss_grad.argtypes = [_P, _I32, _P, _I32, _P, _P]
    L.gfa2_mpnn_train.argtypes = [_P, _I32, ctypes.c_float, _I32, ctypes.c_uint64, _P]
    L.gfa2_mpnn_target_scale.argtypes = [_P]
    L.gfa2_mpnn_target_scale.restype = ctypes.c_float
    L.gfa2_mpnn_infer.argtypes = [_P, _I32, ctypes.c_float, _P]
    L.gfa2_mpnn_infer_to_engine.argtypes = [_P, _I32, ctypes.c_float, _P]
    L.gfa2_mpnn_destroy.argtypes = [_P]
    for name in MPNN_SYMBOLS:
        if name not in ("gfa2_mpnn_num_params", "gfa2_mpnn_target_scale"):
            getattr(L, name).restype = ctypes.c_int
    L._mpnn_bound = True
    return L


class Mpnn:
    """The learned initialiser (gfa2_mpnn_*): 2 message-passing layers, hidden 64, BF16
    weights, FP32 accumulation, trained on the device."""

    IN, HID = 16, 64

    def __init__(self, seed: int = 0, device: int = 0):
        L = _bind_mpnn(lib())
        h = ctypes.c_void_p()
        _check(L.gfa2_mpnn_create(device, int(seed), ctypes.byref(h)), "gfa2_mpnn_create")
        self.h = h
        self.nparams = L.gfa2_mpnn_num_params()
        self.sizes = {}

    def close(self):
        if getattr(self, "h", None):
            lib().gfa2_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        _check(lib().gfa2_mpnn_get_params(self.h, _ptr(out)), "gfa2_mpnn_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().gfa2_mpnn_set_params(self.h, _ptr(p)), "gfa2_mpnn_set_params")

    def add_graph(self, n, rowptr, col, target=None, feature_seed: int = 0, steps: int = 2) -> int:
        rp = np.ascontiguousarray(rowptr, np.int64)
        c = np.ascontiguousarray(col, np.int32)
        t = None if target is None else np.ascontiguousarray(target, np.float32)
        gid = ctypes.c_int32(0)
        _check(lib().gfa2_mpnn_add_graph(self.h, int(n), _ptr(rp), _ptr(c), _ptr(t), int(feature_seed), int(steps),
                                         ctypes.byref(gid)), "gfa2_mpnn_add_graph")
        self.sizes[gid.value] = int(n)
        return gid.value

    def add_engine_graph(self, engine: "Engine", feature_seed: int = 0, steps: int = 2) -> int:
        """Registers the graph `engine` holds (its device CSR, no host copy); no target."""
        gid = ctypes.c_int32(0)
        _check(lib().gfa2_mpnn_add_engine_graph(self.h, engine.h, int(feature_seed), int(steps), ctypes.byref(gid)),
               "gfa2_mpnn_add_engine_graph")
        self.sizes[gid.value] = engine.n
        return gid.value

    def _size(self, gid) -> int:
        if gid not in self.sizes:
            raise GFA2Error(-3, f"graph id {gid} is not registered")
        return self.sizes[gid]

    def release_graph(self, gid):
        """Frees a registered graph's device memory; its id is invalid afterwards."""
        _check(lib().gfa2_mpnn_release_graph(self.h, int(gid)), "gfa2_mpnn_release_graph")
        self.sizes.pop(gid, None)

    def features(self, gid) -> np.ndarray:
        out = np.zeros((self._size(gid), self.IN), np.float32)
        _check(lib().gfa2_mpnn_features(self.h, gid, _ptr(out)), "gfa2_mpnn_features")
        return out

    def loss_grad(self, gid, offsets):
        o = np.ascontiguousarray(offsets, np.int32)
        loss = ctypes.c_double(0)
        g = np.zeros(self.nparams, np.float32)
        _check(lib().gfa2_mpnn_loss_grad(self.h, gid, _ptr(o), len(o), ctypes.byref(loss), _ptr(g)),
               "gfa2_mpnn_loss_grad")
        return loss.value, g

    def train(self, steps, lr=1e-3, R=8, seed=0) -> float:
        loss = ctypes.c_double(0)
        _check(lib().gfa2_mpnn_train(self.h, int(steps), float(lr), int(R), int(seed), ctypes.byref(loss)),
               "gfa2_mpnn_train")
        return loss.value

    def target_scale(self) -> float:
        return float(lib().gfa2_mpnn_target_scale(self.h))

    def infer(self, gid, scale) -> np.ndarray:
        out = np.zeros((self._size(gid), 2), np.float32)
        _check(lib().gfa2_mpnn_infer(self.h, gid, float(scale), _ptr(out)), "gfa2_mpnn_infer")
        return out

    def infer_to_engine(self, gid, scale, engine: "Engine"):
        _check(lib().gfa2_mpnn_infer_to_engine(self.h, gid, float(scale), engine.h), "gfa2_mpnn_infer_to_engine")

"""ctypes binding of libgfa2host.so (include/gfa2host.h): the C++ edge-list loader and
layout writers. No device work."""
import ctypes
import os

import numpy as np

from .graphs import Graph

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgfa2host.so")
_lib = None
_P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.gfa2host_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(_P), _P, _P]
        L.gfa2host_from_edges.argtypes = [ctypes.c_int64, _P, _P, ctypes.POINTER(_P), _P, _P]
        L.gfa2host_graph_arrays.argtypes = [_P, _P, _P, _P, _P]
        L.gfa2host_graph_arrays.restype = None
        L.gfa2host_save_binary.argtypes = [_P, ctypes.c_char_p]
        L.gfa2host_write_layout.argtypes = [_P, _P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32]
        L.gfa2host_free.argtypes = [_P]
        L.gfa2host_free.restype = None
        L.gfa2host_barabasi_albert.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, _P, _P]
        L.gfa2host_barabasi_albert.restype = ctypes.c_int64
        L.gfa2host_rmat.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, _P, _P]
        L.gfa2host_rmat.restype = ctypes.c_int64
        from . import Raster
        L.gfa2host_png_frame.argtypes = [ctypes.c_float] * 4 + [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                                                  ctypes.c_int64, ctypes.POINTER(Raster)]
        L.gfa2host_png_frame.restype = None
        L.gfa2host_render_gray.argtypes = [_P, _P, ctypes.POINTER(Raster), _P]
        L.gfa2host_write_png_gray.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, _P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


class HostGraph:
    """A graph held by the C++ UGraph."""

    def __init__(self, handle, n, m):
        self.h, self.n, self.m = handle, n, m

    @classmethod
    def load(cls, path):
        h = _P()
        n = ctypes.c_int32(0)
        m = ctypes.c_int64(0)
        if lib().gfa2host_load(path.encode(), ctypes.byref(h), ctypes.byref(n), ctypes.byref(m)) != 0:
            raise OSError(f"cannot load {path}")
        return cls(h, n.value, m.value)

    @classmethod
    def from_edges(cls, s, t):
        s = np.ascontiguousarray(s, np.uint32)
        t = np.ascontiguousarray(t, np.uint32)
        h = _P()
        n = ctypes.c_int32(0)
        m = ctypes.c_int64(0)
        lib().gfa2host_from_edges(len(s), _p(s), _p(t), ctypes.byref(h), ctypes.byref(n), ctypes.byref(m))
        return cls(h, n.value, m.value)

    def arrays(self) -> Graph:
        ids = np.zeros(self.n, np.int64)
        deg = np.zeros(self.n, np.int32)
        src = np.zeros(self.m, np.int32)
        dst = np.zeros(self.m, np.int32)
        lib().gfa2host_graph_arrays(self.h, _p(ids), _p(deg), _p(src), _p(dst))
        self.deg = deg
        return Graph(self.n, src, dst, ids)

    def save_binary(self, path):
        if lib().gfa2host_save_binary(self.h, path.encode()) != 0:
            raise OSError(path)

    def write_layout(self, xy, path, fmt, w=1250, h=1250):
        xy = np.ascontiguousarray(xy, np.float32)
        return lib().gfa2host_write_layout(self.h, _p(xy), path.encode(), fmt.encode(), w, h)

    def close(self):
        if self.h:
            lib().gfa2host_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def png_frame(bbox, width: int, height: int, nodes: int, edges: int):
    """The frame fitting bbox = (min x, max x, min y, max y) to a width x height image."""
    from . import Raster
    f = Raster()
    lib().gfa2host_png_frame(float(bbox[0]), float(bbox[1]), float(bbox[2]), float(bbox[3]), int(width), int(height),
                             int(nodes), int(edges), ctypes.byref(f))
    return f


def render_gray(hg: "HostGraph", xy, frame) -> np.ndarray:
    """The CPU rasteriser of the PNG picture: height x width grey bytes, top row first."""
    xy = np.ascontiguousarray(xy, np.float32)
    out = np.zeros((frame.height, frame.width), np.uint8)
    if lib().gfa2host_render_gray(hg.h, _p(xy), ctypes.byref(frame), _p(out)) != 0:
        raise ValueError("render_gray")
    return out


def write_png_gray(path: str, gray: np.ndarray):
    gray = np.ascontiguousarray(gray, np.uint8)
    if lib().gfa2host_write_png_gray(path.encode(), gray.shape[1], gray.shape[0], _p(gray)) != 0:
        raise OSError(path)

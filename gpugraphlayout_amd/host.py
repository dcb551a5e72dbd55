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

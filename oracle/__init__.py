"""CPU oracle for the Gemini edge-processing engine (test infrastructure).

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only to check or time the reference
algorithm; the product path never calls it.  The C restatement lives in
``oracle.c`` / ``rmat.c`` (see ``oracle.h`` for the reference anchors and the
pinning status of every function) and is built by ``oracle/Makefile`` into
``oracle/liboracle.so``.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

RMAT_A, RMAT_B, RMAT_C = 0.57, 0.19, 0.19


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _RmatParams(ctypes.Structure):
    _fields_ = [
        ("scale", ctypes.c_int),
        ("num_edges", ctypes.c_uint64),
        ("a", ctypes.c_double),
        ("b", ctypes.c_double),
        ("c", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("permute", ctypes.c_int),
        ("weighted", ctypes.c_int),
    ]


class _OgGraph(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_uint32),
        ("edges", ctypes.c_uint64),
        ("arcs", ctypes.c_uint64),
        ("weighted", ctypes.c_int),
        ("symmetric", ctypes.c_int),
        ("out_degree", ctypes.POINTER(ctypes.c_uint32)),
        ("in_degree", ctypes.POINTER(ctypes.c_uint32)),
        ("out_index", ctypes.POINTER(ctypes.c_uint64)),
        ("out_nbr", ctypes.POINTER(ctypes.c_uint32)),
        ("out_w", ctypes.POINTER(ctypes.c_float)),
        ("in_index", ctypes.POINTER(ctypes.c_uint64)),
        ("in_nbr", ctypes.POINTER(ctypes.c_uint32)),
        ("in_w", ctypes.POINTER(ctypes.c_float)),
    ]


class _OgTrace(ctypes.Structure):
    _fields_ = [
        ("modes", ctypes.POINTER(ctypes.c_int32)),
        ("active", ctypes.POINTER(ctypes.c_uint32)),
        ("cap", ctypes.c_int),
        ("n", ctypes.c_int),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        G = ctypes.POINTER(_OgGraph)
        T = ctypes.POINTER(_OgTrace)
        L.og_rmat_generate.argtypes = [ctypes.POINTER(_RmatParams), vp]
        L.og_rmat_generate.restype = None
        L.og_rmat_permute.argtypes = [u32, i32, u64]
        L.og_rmat_permute.restype = u32
        L.og_graph_build.argtypes = [vp, u64, u32, i32, i32]
        L.og_graph_build.restype = G
        L.og_graph_free.argtypes = [G]
        L.og_graph_transpose.argtypes = [G]
        L.og_set_threads.argtypes = [i32]
        L.og_get_threads.restype = i32
        L.og_pagerank.argtypes = [G, i32, vp, vp, T]
        L.og_bfs.argtypes = [G, u32, vp, vp, vp, i32, T]
        L.og_bfs.restype = i32
        L.og_cc.argtypes = [G, vp, vp, i32, T]
        L.og_cc.restype = i32
        L.og_sssp.argtypes = [G, u32, vp, vp, i32, T]
        L.og_sssp.restype = i32
        L.og_sssp_fixed_point_violations.argtypes = [vp, vp, vp, u32, vp, u32]
        L.og_sssp_fixed_point_violations.restype = u64
        L.og_bc.argtypes = [G, u32, vp, vp, vp, i32, T]
        L.og_bc.restype = i32
        L.og_time.restype = ctypes.c_double
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


EDGE_DTYPE = np.dtype([("src", "<u4"), ("dst", "<u4")])
WEDGE_DTYPE = np.dtype([("src", "<u4"), ("dst", "<u4"), ("w", "<f4")])


def sssp_fixed_point_violations(index, src, w, dist, root):
    """Vertices where `dist` is not a fixed point of the reference's f32
    relaxation over the in-arcs (CSC arrays index/src/w); 0 for a correct
    SSSP result (og_sssp_fixed_point_violations)."""
    index = np.ascontiguousarray(index, dtype=np.uint64)
    src = np.ascontiguousarray(src, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    dist = np.ascontiguousarray(dist, dtype=np.float32)
    V = len(index) - 1
    assert len(dist) == V and len(src) == len(w) == int(index[-1])
    return int(lib().og_sssp_fixed_point_violations(_ptr(index), _ptr(src), _ptr(w), V,
                                                    _ptr(dist), int(root)))


def rmat_edges(scale, edge_factor=16, seed=1, a=RMAT_A, b=RMAT_B, c=RMAT_C,
               permute=True, weighted=False, num_edges=None):
    """Synthetic RMAT edge list as a structured array (packed 8/12-byte records)."""
    ne = (1 << scale) * edge_factor if num_edges is None else int(num_edges)
    p = _RmatParams(scale, ne, a, b, c, seed, int(permute), int(weighted))
    out = np.empty(ne, dtype=WEDGE_DTYPE if weighted else EDGE_DTYPE)
    lib().og_rmat_generate(ctypes.byref(p), _ptr(out))
    return out


def permute(x, scale, seed):
    return lib().og_rmat_permute(int(x), int(scale), int(seed))


def set_threads(n):
    lib().og_set_threads(int(n))


def get_threads():
    return lib().og_get_threads()


def wall_time():
    return lib().og_time()


def edges_from_pairs(pairs, weights=None):
    pairs = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    if weights is None:
        out = np.empty(len(pairs), dtype=EDGE_DTYPE)
    else:
        out = np.empty(len(pairs), dtype=WEDGE_DTYPE)
        out["w"] = np.asarray(weights, dtype=np.float32)
    out["src"] = pairs[:, 0]
    out["dst"] = pairs[:, 1]
    return out


class Trace:
    def __init__(self, cap=4096):
        self.modes = np.zeros(cap, dtype=np.int32)
        self.active = np.zeros(cap, dtype=np.uint32)
        self.t = _OgTrace(self.modes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                          self.active.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap, 0)

    def result(self):
        n = min(self.t.n, len(self.modes))
        return self.modes[:n].copy(), self.active[:n].copy()


class Graph:
    """Oracle graph: load_directed / load_undirected_from_directed semantics."""

    def __init__(self, edges, vertices, undirected=False):
        edges = np.ascontiguousarray(edges)
        weighted = edges.dtype.itemsize == 12
        self._edges = edges
        self.vertices = int(vertices)
        self.weighted = weighted
        self.undirected = bool(undirected)
        self._g = lib().og_graph_build(_ptr(edges), len(edges), self.vertices, int(weighted),
                                       int(undirected))
        g = self._g.contents
        self.edges = g.edges
        self.arcs = g.arcs

    def __del__(self):
        if getattr(self, "_g", None) is not None and _lib is not None:
            _lib.og_graph_free(self._g)
            self._g = None

    def _arr(self, p, n, dt):
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

    def out_degree(self):
        return self._arr(self._g.contents.out_degree, self.vertices, np.uint32)

    def in_degree(self):
        return self._arr(self._g.contents.in_degree, self.vertices, np.uint32)

    def csr(self):
        g = self._g.contents
        idx = self._arr(g.out_index, self.vertices + 1, np.uint64)
        nbr = self._arr(g.out_nbr, int(g.arcs), np.uint32)
        w = self._arr(g.out_w, int(g.arcs), np.float32) if self.weighted else None
        return idx, nbr, w

    def csc(self):
        g = self._g.contents
        idx = self._arr(g.in_index, self.vertices + 1, np.uint64)
        nbr = self._arr(g.in_nbr, int(g.arcs), np.uint32)
        w = self._arr(g.in_w, int(g.arcs), np.float32) if self.weighted else None
        return idx, nbr, w

    def transpose(self):
        lib().og_graph_transpose(self._g)

    def pagerank(self, iterations=20, trace=None):
        curr = np.empty(self.vertices, dtype=np.float64)
        deltas = np.empty(iterations + 1, dtype=np.float64)
        lib().og_pagerank(self._g, iterations, _ptr(curr), _ptr(deltas),
                          ctypes.byref(trace.t) if trace else None)
        return curr, deltas

    def bfs(self, root, trace=None):
        V = self.vertices
        parent = np.empty(V, dtype=np.uint32)
        depth = np.empty(V, dtype=np.uint32)
        fr = np.zeros(V + 2, dtype=np.uint32)
        n = lib().og_bfs(self._g, int(root), _ptr(parent), _ptr(depth), _ptr(fr), len(fr),
                         ctypes.byref(trace.t) if trace else None)
        return parent, depth, fr[:n].copy()

    def cc(self, trace=None):
        label = np.empty(self.vertices, dtype=np.uint32)
        fr = np.zeros(self.vertices + 2, dtype=np.uint32)
        n = lib().og_cc(self._g, _ptr(label), _ptr(fr), len(fr),
                        ctypes.byref(trace.t) if trace else None)
        return label, fr[:n].copy()

    def sssp(self, root, trace=None):
        dist = np.empty(self.vertices, dtype=np.float32)
        fr = np.zeros(self.vertices + 2, dtype=np.uint32)
        n = lib().og_sssp(self._g, int(root), _ptr(dist), _ptr(fr), len(fr),
                          ctypes.byref(trace.t) if trace else None)
        return dist, fr[:n].copy()

    def bc(self, root, trace=None):
        dep = np.empty(self.vertices, dtype=np.float64)
        inv = np.empty(self.vertices, dtype=np.float64)
        fr = np.zeros(self.vertices + 2, dtype=np.uint32)
        n = lib().og_bc(self._g, int(root), _ptr(dep), _ptr(inv), _ptr(fr), len(fr),
                        ctypes.byref(trace.t) if trace else None)
        return dep, inv, fr[:n].copy()

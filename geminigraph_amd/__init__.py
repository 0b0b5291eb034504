"""MI355X-native Gemini edge-processing engine: Python host side.

This package mirrors the reference's host interface for one partition on one
GPU (``Graph<EdgeData>`` of core/graph.hpp and the toolkits'
``compute()``), calling the HIP kernels in ``libgg.so`` through the C ABI
declared in ``include/gg.h``.  Nothing here computes on the CPU: if the
library is missing or no GPU is present, the calls raise.

    import geminigraph_amd as gg
    g = gg.Graph.load_directed("graph.bin", vertices)       # core/graph.hpp:772
    parent, depth, frontier, stats = g.bfs(root)             # toolkits/bfs.cpp
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgg.so")
_lib = None

GG_OK = 0
RMAT_A, RMAT_B, RMAT_C = 0.57, 0.19, 0.19


class GGError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"gg error {code}: {msg}")
        self.code = code


def build():
    """Compile libgg.so and the facade toolkits for gfx950 (hipcc; runs
    without a GPU)."""
    subprocess.run(["make", "-s", "-B", "-j4", "-C", _HERE], check=True)
    subprocess.run(["make", "-s", "-B", "-j4", "-C", _HERE, "toolkits"], check=True)


STATS_RECORDS = 256  # GG_STATS_RECORDS in gg.h


class RunStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("push_iterations", ctypes.c_int),
        ("edges_examined", ctypes.c_uint64),
        ("m_reach", ctypes.c_uint64),
        ("reached", ctypes.c_uint64),
        ("device_ms", ctypes.c_double),
        ("modes", ctypes.c_uint32 * STATS_RECORDS),
        ("it_active", ctypes.c_uint32 * STATS_RECORDS),
        ("it_new", ctypes.c_uint32 * STATS_RECORDS),
        ("it_examined", ctypes.c_uint64 * STATS_RECORDS),
        ("it_unvisited", ctypes.c_uint64 * STATS_RECORDS),
    ]

    def as_dict(self):
        n = min(self.iterations, STATS_RECORDS)
        return {
            "iterations": self.iterations,
            "push_iterations": self.push_iterations,
            "edges_examined": int(self.edges_examined),
            "m_reach": int(self.m_reach),
            "reached": int(self.reached),
            "device_ms": float(self.device_ms),
            "modes": [int(x) for x in self.modes[:n]],
            "it_active": [int(x) for x in self.it_active[:n]],
            "it_new": [int(x) for x in self.it_new[:n]],
            "it_examined": [int(x) for x in self.it_examined[:n]],
            "it_unvisited": [int(x) for x in self.it_unvisited[:n]],
        }


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_uint32),
        ("edges", ctypes.c_uint64),
        ("arcs", ctypes.c_uint64),
        ("weighted", ctypes.c_int),
        ("symmetric", ctypes.c_int),
        ("transposed", ctypes.c_int),
        ("device_bytes", ctypes.c_size_t),
        ("d_out_degree", ctypes.c_void_p),
        ("d_in_degree", ctypes.c_void_p),
        ("d_out_index", ctypes.c_void_p),
        ("d_out_nbr", ctypes.c_void_p),
        ("d_out_w", ctypes.c_void_p),
        ("d_in_index", ctypes.c_void_p),
        ("d_in_nbr", ctypes.c_void_p),
        ("d_in_w", ctypes.c_void_p),
        ("reordered", ctypes.c_int),
        ("d_perm", ctypes.c_void_p),
        ("d_iperm", ctypes.c_void_p),
    ]


BUILD_REORDER = 1
BUILD_HUB_FIRST = 2


def _flags(reorder, hub_first):
    return (BUILD_REORDER if reorder else 0) | (BUILD_HUB_FIRST if hub_first else 0)


class TuneResult(ctypes.Structure):
    """gg_tune_result of include/gg.h."""
    _fields_ = [("kind", ctypes.c_int), ("n_options", ctypes.c_int),
                ("names", (ctypes.c_char * 24) * 8), ("values", ctypes.c_int64 * 8),
                ("baseline_ms", ctypes.c_double), ("tuned_ms", ctypes.c_double),
                ("trials", ctypes.c_int)]

    def as_dict(self):
        n = self.n_options
        return {"options": {bytes(self.names[i].value).decode(): int(self.values[i]) for i in range(n)},
                "baseline_ms": self.baseline_ms, "tuned_ms": self.tuned_ms, "trials": self.trials}


TUNE_BFS, TUNE_PAGERANK = 0, 1


class ServerInfo(ctypes.Structure):
    """gg_server_info of include/gg.h."""
    _fields_ = [("budget_bytes", ctypes.c_size_t), ("used_bytes", ctypes.c_size_t),
                ("graphs", ctypes.c_int), ("workers", ctypes.c_int),
                ("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("max_batch", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


QUERY_BFS, QUERY_SSSP, QUERY_BC = 0, 1, 2
_QUERY_DTYPE = {QUERY_BFS: np.uint32, QUERY_SSSP: np.float32, QUERY_BC: np.float64}


# C ABI signatures: name -> (restype, argtypes)
_vp, _u32, _u64, _i32, _i64 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                               ctypes.c_int64)
_dbl, _sz, _cp = ctypes.c_double, ctypes.c_size_t, ctypes.c_char_p
_pp = ctypes.POINTER(ctypes.c_void_p)
_stats = ctypes.POINTER(RunStats)
SIGNATURES = {
    "gg_last_error": (_cp, []),
    "gg_init": (_i32, [_i32]),
    "gg_synchronize": (_i32, []),
    "gg_malloc": (_i32, [_pp, _sz]),
    "gg_free": (_i32, [_vp]),
    "gg_memcpy_d2h": (_i32, [_vp, _vp, _sz]),
    "gg_memcpy_h2d": (_i32, [_vp, _vp, _sz]),
    "gg_memcpy_d2d": (_i32, [_vp, _vp, _sz]),
    "gg_mem_info": (_i32, [ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "gg_device_copy": (_i32, [_vp, _vp, _sz]),
    "gg_set_option": (_i32, [_cp, _i64]),
    "gg_get_option": (_i32, [_cp, ctypes.POINTER(_i64)]),
    "gg_rmat_generate": (_i32, [_vp, _u64, _i32, _dbl, _dbl, _dbl, _u64, _i32, _i32]),
    "gg_graph_build": (_i32, [_vp, _i32, _u64, _u32, _i32, _i32, _pp]),
    "gg_graph_build_ex": (_i32, [_vp, _i32, _u64, _u32, _i32, _i32, ctypes.c_uint, _pp]),
    "gg_graph_load": (_i32, [_cp, _u32, _i32, _i32, _pp]),
    "gg_graph_load_ex": (_i32, [_cp, _u32, _i32, _i32, ctypes.c_uint, _pp]),
    "gg_graph_destroy": (_i32, [_vp]),
    "gg_graph_transpose": (_i32, [_vp]),
    "gg_graph_get_info": (_i32, [_vp, ctypes.POINTER(GraphInfo)]),
    "gg_subset_create": (_i32, [_vp, _pp]),
    "gg_subset_destroy": (_i32, [_vp]),
    "gg_subset_clear": (_i32, [_vp]),
    "gg_subset_fill": (_i32, [_vp]),
    "gg_subset_set_bit": (_i32, [_vp, _u32]),
    "gg_subset_get_bit": (_i32, [_vp, _u32, ctypes.POINTER(ctypes.c_int)]),
    "gg_subset_count": (_i32, [_vp, ctypes.POINTER(_u64)]),
    "gg_subset_device_words": (_vp, [_vp]),
    "gg_pagerank": (_i32, [_vp, _i32, _vp, _vp, _stats]),
    "gg_bfs": (_i32, [_vp, _u32, _vp, _vp, _vp, _i32, _stats]),
    "gg_bfs_noalloc": (_i32, [_vp, _u32, _vp, _vp, _vp, _i32, _stats]),
    "gg_cc": (_i32, [_vp, _vp, _vp, _i32, _stats]),
    "gg_sssp": (_i32, [_vp, _u32, _vp, _vp, _i32, _stats]),
    "gg_bc": (_i32, [_vp, _u32, _vp, _vp, _vp, _i32, _stats]),
    "gg_bfs_batch": (_i32, [_vp, _vp, _i32, _vp, _stats]),
    "gg_sssp_batch": (_i32, [_vp, _vp, _i32, _vp, _stats]),
    "gg_bc_batch": (_i32, [_vp, _vp, _i32, _vp, _vp, _stats]),
    "gg_server_create": (_i32, [_sz, _i32, _pp]),
    "gg_server_destroy": (_i32, [_vp]),
    "gg_server_add_graph": (_i32, [_vp, _vp, ctypes.POINTER(ctypes.c_int)]),
    "gg_server_submit": (_i32, [_vp, _i32, _i32, _u32, _i32, ctypes.POINTER(_u64)]),
    "gg_server_wait": (_i32, [_vp, _u64, _vp, _sz, ctypes.POINTER(_dbl)]),
    "gg_server_replay": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_dbl)]),
    "gg_server_get_info": (_i32, [_vp, ctypes.POINTER(ServerInfo)]),
    "gg_tune": (_i32, [_vp, _i32, _i32, _i32, ctypes.POINTER(TuneResult)]),
    "gg_dump_vertex_array": (_i32, [_vp, _vp, _sz, _cp]),
    "gg_restore_vertex_array": (_i32, [_vp, _vp, _sz, _cp]),
    "gg_gather_vertex_array": (_i32, [_vp, _vp, _sz, _vp]),
    "gg_graph_save": (_i32, [_vp, _cp]),
    "gg_graph_open": (_i32, [_cp, _pp]),
    "gg_prof_enable": (_i32, [_i32]),
    "gg_prof_reset": (_i32, []),
    "gg_prof_read": (_i32, [_cp, ctypes.POINTER(_dbl), ctypes.POINTER(_u64)]),
}


def lib():
    """Load libgg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run geminigraph_amd.build() "
                               "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(code):
    if code != GG_OK:
        raise GGError(code, lib().gg_last_error().decode(errors="replace"))


def init(device=0):
    _check(lib().gg_init(int(device)))


def synchronize():
    _check(lib().gg_synchronize())


def set_option(name, value):
    _check(lib().gg_set_option(name.encode(), int(value)))


def get_option(name):
    v = ctypes.c_int64()
    _check(lib().gg_get_option(name.encode(), ctypes.byref(v)))
    return v.value


def device_copy(dst, src, nbytes):
    """Plain streaming copy kernel between two DeviceArrays (HBM calibration)."""
    _check(lib().gg_device_copy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), int(nbytes)))


def mem_info():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib().gg_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


class options:
    """Context manager that sets options and restores them on exit."""

    def __init__(self, **kw):
        self.kw = kw
        self.old = {}

    def __enter__(self):
        for k, v in self.kw.items():
            self.old[k] = get_option(k)
            set_option(k, v)
        return self

    def __exit__(self, *exc):
        for k, v in self.old.items():
            set_option(k, v)


class DeviceArray:
    """A device buffer (alloc_vertex_array / dealloc_vertex_array analogue)."""

    def __init__(self, n, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().gg_malloc(ctypes.byref(p), max(self.nbytes, 8)))
        self.ptr = p.value

    def numpy(self):
        """gather_vertex_array: copy to host."""
        out = np.empty(self.n, dtype=self.dtype)
        if self.nbytes:
            _check(lib().gg_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.size == self.n
        if self.nbytes:
            _check(lib().gg_memcpy_h2d(self.ptr, arr.ctypes.data, self.nbytes))
        return self

    def free(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.gg_free(ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rmat_device(scale, num_edges, seed, a=RMAT_A, b=RMAT_B, c=RMAT_C, permute=True,
                weighted=False):
    """Generate an RMAT edge list in device memory (packed 8/12-byte records)."""
    rec = 12 if weighted else 8
    buf = DeviceArray(num_edges * rec, np.uint8)
    _check(lib().gg_rmat_generate(ctypes.c_void_p(buf.ptr), int(num_edges), int(scale),
                                  float(a), float(b), float(c), int(seed), int(permute),
                                  int(weighted)))
    return buf


class Graph:
    """Graph<EdgeData> for one partition on one GPU (core/graph.hpp:91-1869)."""

    def __init__(self, handle):
        self._h = handle
        info = GraphInfo()
        _check(lib().gg_graph_get_info(ctypes.c_void_p(handle), ctypes.byref(info)))
        self.vertices = info.vertices
        self.edges = info.edges
        self.arcs = info.arcs
        self.weighted = bool(info.weighted)
        self.symmetric = bool(info.symmetric)

    # -- construction -----------------------------------------------------
    @classmethod
    def build(cls, edges, vertices, undirected=False, reorder=False, hub_first=False):
        """From a host structured array of packed records (oracle.EDGE_DTYPE / WEDGE_DTYPE).
        reorder=True relabels vertices internally by descending out-degree;
        hub_first=True keeps file ids and orders each list hubs first."""
        edges = np.ascontiguousarray(edges)
        weighted = edges.dtype.itemsize == 12
        h = ctypes.c_void_p()
        _check(lib().gg_graph_build_ex(ctypes.c_void_p(edges.ctypes.data), 0, len(edges),
                                       int(vertices), int(weighted), int(undirected),
                                       _flags(reorder, hub_first), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def build_device(cls, dev_edges, num_edges, vertices, weighted=False, undirected=False,
                     reorder=False, hub_first=False):
        h = ctypes.c_void_p()
        _check(lib().gg_graph_build_ex(ctypes.c_void_p(dev_edges.ptr), 1, int(num_edges),
                                       int(vertices), int(weighted), int(undirected),
                                       _flags(reorder, hub_first), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def rmat(cls, scale, edge_factor=16, seed=1, a=RMAT_A, b=RMAT_B, c=RMAT_C, permute=True,
             weighted=False, undirected=False, reorder=False, hub_first=False):
        ne = (1 << scale) * edge_factor
        buf = rmat_device(scale, ne, seed, a, b, c, permute, weighted)
        try:
            return cls.build_device(buf, ne, 1 << scale, weighted, undirected, reorder, hub_first)
        finally:
            buf.free()

    @classmethod
    def load_directed(cls, path, vertices, weighted=False, reorder=False, hub_first=False):
        """core/graph.hpp:772 -- |E| = file size / record size."""
        h = ctypes.c_void_p()
        _check(lib().gg_graph_load_ex(str(path).encode(), int(vertices), int(weighted), 0,
                                      _flags(reorder, hub_first), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def load_undirected_from_directed(cls, path, vertices, weighted=False, reorder=False,
                                      hub_first=False):
        """core/graph.hpp:350 -- symmetric CSR, CSC aliased, |E| kept."""
        h = ctypes.c_void_p()
        _check(lib().gg_graph_load_ex(str(path).encode(), int(vertices), int(weighted), 1,
                                      _flags(reorder, hub_first), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def open(cls, path):
        """A graph written by save(): no edge sort on load."""
        h = ctypes.c_void_p()
        _check(lib().gg_graph_open(str(path).encode(), ctypes.byref(h)))
        return cls(h.value)

    def save(self, path):
        """Both directions, weights and the relabeling, to a binary file."""
        _check(lib().gg_graph_save(self.handle, str(path).encode()))

    # -- vertex arrays (core/graph.hpp:253-324, one partition) -------------
    def dump_vertex_array(self, array, path):
        """A DeviceArray of `vertices` elements to a file of its bytes."""
        self._check_vertex_array(array)
        _check(lib().gg_dump_vertex_array(self.handle, ctypes.c_void_p(array.ptr),
                                          array.dtype.itemsize, str(path).encode()))

    def restore_vertex_array(self, array, path):
        """A file written by dump_vertex_array back into a DeviceArray."""
        self._check_vertex_array(array)
        _check(lib().gg_restore_vertex_array(self.handle, ctypes.c_void_p(array.ptr),
                                             array.dtype.itemsize, str(path).encode()))

    def gather_vertex_array(self, array):
        """The whole array on the host (one partition holds every vertex)."""
        self._check_vertex_array(array)
        out = np.empty(self.vertices, dtype=array.dtype)
        _check(lib().gg_gather_vertex_array(self.handle, ctypes.c_void_p(array.ptr),
                                            array.dtype.itemsize, out.ctypes.data))
        return out

    def _check_vertex_array(self, array, dtype=None, per_vertex=1):
        """A DeviceArray of at least per_vertex * V elements (of `dtype`):
        the kernels write that many, so a shorter buffer would be overrun."""
        if not isinstance(array, DeviceArray):
            raise TypeError("expected a DeviceArray")
        if array.n < self.vertices * per_vertex:
            raise ValueError(f"vertex array of {array.n} elements; the call writes "
                             f"{self.vertices * per_vertex}")
        if dtype is not None and array.dtype != np.dtype(dtype):
            raise ValueError(f"vertex array of {array.dtype}; the call writes {np.dtype(dtype)}")

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gg_graph_destroy(ctypes.c_void_p(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return ctypes.c_void_p(self._h)

    def info(self):
        info = GraphInfo()
        _check(lib().gg_graph_get_info(self.handle, ctypes.byref(info)))
        return info

    def transpose(self):
        _check(lib().gg_graph_transpose(self.handle))

    def _copy(self, ptr, n, dtype):
        out = np.empty(n, dtype=dtype)
        if n:
            _check(lib().gg_memcpy_d2h(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes))
        return out

    @property
    def reordered(self):
        return bool(self.info().reordered)

    def perm(self):
        """perm[original id] = internal id (identity when not reordered)."""
        i = self.info()
        if not i.reordered:
            return np.arange(self.vertices, dtype=np.uint32)
        return self._copy(i.d_perm, self.vertices, np.uint32)

    def out_degree(self):
        d = self._copy(self.info().d_out_degree, self.vertices, np.uint32)
        return d[self.perm()] if self.reordered else d

    def in_degree(self):
        d = self._copy(self.info().d_in_degree, self.vertices, np.uint32)
        return d[self.perm()] if self.reordered else d

    def _adjacency(self, idx_p, nbr_p, w_p):
        idx = self._copy(idx_p, self.vertices + 1, np.uint64)
        nbr = self._copy(nbr_p, self.arcs, np.uint32)
        w = self._copy(w_p, self.arcs, np.float32) if self.weighted else None
        if not self.reordered:
            return idx, nbr, w
        # back to original ids: each list sorted by original neighbour, ties
        # in stored (input) order
        iperm = np.argsort(self.perm()).astype(np.uint32)
        deg = np.diff(idx).astype(np.int64)
        key = iperm[np.repeat(np.arange(self.vertices, dtype=np.int64), deg)]
        onbr = iperm[nbr]
        order = np.lexsort((np.arange(len(onbr)), onbr, key))
        odeg = np.bincount(key, minlength=self.vertices) if len(key) else \
            np.zeros(self.vertices, dtype=np.int64)
        oidx = np.zeros(self.vertices + 1, dtype=np.uint64)
        oidx[1:] = np.cumsum(odeg)
        return oidx, onbr[order], (w[order] if w is not None else None)

    def csr(self):
        i = self.info()
        return self._adjacency(i.d_out_index, i.d_out_nbr, i.d_out_w)

    def csc(self):
        i = self.info()
        return self._adjacency(i.d_in_index, i.d_in_nbr, i.d_in_w)

    # -- toolkits (device outputs; *_into variants keep results on the GPU) --
    def pagerank_into(self, out, iterations=20, deltas=None):
        self._check_vertex_array(out, np.float64)
        if deltas is not None and len(deltas) < iterations + 1:
            raise ValueError("deltas needs iterations + 1 entries")
        st = RunStats()
        dptr = deltas.ctypes.data if deltas is not None else None
        _check(lib().gg_pagerank(self.handle, int(iterations), ctypes.c_void_p(out.ptr),
                                 ctypes.c_void_p(dptr) if dptr else None, ctypes.byref(st)))
        return st

    def pagerank(self, iterations=20):
        out = DeviceArray(self.vertices, np.float64)
        deltas = np.zeros(iterations + 1, dtype=np.float64)
        st = self.pagerank_into(out, iterations, deltas)
        return out.numpy(), deltas, st.as_dict()

    def bfs_into(self, root, parent, depth=None, frontier=None):
        """BFS into device arrays; depth None keeps parents only (the
        reference's output, bfs.cpp:26-94)."""
        self._check_vertex_array(parent, np.uint32)
        if depth is not None:
            self._check_vertex_array(depth, np.uint32)
        st = RunStats()
        fr_ptr = frontier.ctypes.data if frontier is not None else None
        cap = len(frontier) if frontier is not None else 0
        dp = ctypes.c_void_p(depth.ptr) if depth is not None else None
        _check(lib().gg_bfs(self.handle, int(root), ctypes.c_void_p(parent.ptr),
                            dp, ctypes.c_void_p(fr_ptr) if fr_ptr else None,
                            cap, ctypes.byref(st)))
        return st

    def bfs_noalloc_into(self, root, parent, depth=None, frontier=None):
        """bfs_into through gg_bfs_noalloc: no allocation inside the call
        (graphs of at most 2^20 vertices; overlapping calls on a GPU race)."""
        self._check_vertex_array(parent, np.uint32)
        if depth is not None:
            self._check_vertex_array(depth, np.uint32)
        st = RunStats()
        fr_ptr = frontier.ctypes.data if frontier is not None else None
        cap = len(frontier) if frontier is not None else 0
        dp = ctypes.c_void_p(depth.ptr) if depth is not None else None
        _check(lib().gg_bfs_noalloc(self.handle, int(root), ctypes.c_void_p(parent.ptr),
                                    dp, ctypes.c_void_p(fr_ptr) if fr_ptr else None,
                                    cap, ctypes.byref(st)))
        return st

    def bfs(self, root):
        parent = DeviceArray(self.vertices, np.uint32)
        depth = DeviceArray(self.vertices, np.uint32)
        fr = np.zeros(min(self.vertices + 2, 1 << 16), dtype=np.uint32)
        st = self.bfs_into(root, parent, depth, fr)
        levels = int(np.max(np.nonzero(fr)[0]) + 1) if fr.any() else 0
        return parent.numpy(), depth.numpy(), fr[:levels].copy(), st.as_dict()

    def cc_into(self, label, frontier=None):
        self._check_vertex_array(label, np.uint32)
        st = RunStats()
        fr_ptr = frontier.ctypes.data if frontier is not None else None
        cap = len(frontier) if frontier is not None else 0
        _check(lib().gg_cc(self.handle, ctypes.c_void_p(label.ptr),
                           ctypes.c_void_p(fr_ptr) if fr_ptr else None, cap, ctypes.byref(st)))
        return st

    def cc(self):
        label = DeviceArray(self.vertices, np.uint32)
        fr = np.zeros(min(self.vertices + 2, 1 << 16), dtype=np.uint32)
        st = self.cc_into(label, fr)
        n = int(np.max(np.nonzero(fr)[0]) + 1) if fr.any() else 0
        return label.numpy(), fr[:n].copy(), st.as_dict()

    def sssp_into(self, root, dist, frontier=None):
        self._check_vertex_array(dist, np.float32)
        st = RunStats()
        fr_ptr = frontier.ctypes.data if frontier is not None else None
        cap = len(frontier) if frontier is not None else 0
        _check(lib().gg_sssp(self.handle, int(root), ctypes.c_void_p(dist.ptr),
                             ctypes.c_void_p(fr_ptr) if fr_ptr else None, cap, ctypes.byref(st)))
        return st

    def sssp(self, root):
        dist = DeviceArray(self.vertices, np.float32)
        fr = np.zeros(min(self.vertices + 2, 1 << 16), dtype=np.uint32)
        st = self.sssp_into(root, dist, fr)
        n = int(np.max(np.nonzero(fr)[0]) + 1) if fr.any() else 0
        return dist.numpy(), fr[:n].copy(), st.as_dict()

    def bc_into(self, root, dep, inv, frontier=None):
        self._check_vertex_array(dep, np.float64)
        self._check_vertex_array(inv, np.float64)
        st = RunStats()
        fr_ptr = frontier.ctypes.data if frontier is not None else None
        cap = len(frontier) if frontier is not None else 0
        _check(lib().gg_bc(self.handle, int(root), ctypes.c_void_p(dep.ptr),
                           ctypes.c_void_p(inv.ptr), ctypes.c_void_p(fr_ptr) if fr_ptr else None,
                           cap, ctypes.byref(st)))
        return st

    def bc(self, root):
        dep = DeviceArray(self.vertices, np.float64)
        inv = DeviceArray(self.vertices, np.float64)
        fr = np.zeros(min(self.vertices + 2, 1 << 16), dtype=np.uint32)
        st = self.bc_into(root, dep, inv, fr)
        n = int(np.max(np.nonzero(fr)[0]) + 1) if fr.any() else 0
        return dep.numpy(), inv.numpy(), fr[:n].copy(), st.as_dict()


def tune(graph, kind="bfs", samples=8, apply=True):
    """Times candidate option values on a sample of this graph's work (BFS
    from `samples` roots, or `samples` PageRank iterations) and keeps the
    fastest per option; with apply=True the winners stay set (options are
    process-wide).  Returns {"options": {...}, "baseline_ms", "tuned_ms",
    "trials"}."""
    k = {"bfs": TUNE_BFS, "pagerank": TUNE_PAGERANK}.get(kind, kind)
    r = TuneResult()
    _check(lib().gg_tune(graph.handle, int(k), int(samples), int(bool(apply)), ctypes.byref(r)))
    return r.as_dict()


def _batch_roots(roots):
    roots = np.ascontiguousarray(roots, dtype=np.uint32)
    if not 1 <= len(roots) <= 64:
        raise ValueError("1 to 64 roots per batch")
    return roots


def bfs_batch_into(graph, roots, depth):
    """Up to 64 BFS roots in one traversal; depth is a DeviceArray of V*64
    uint32, vertex-major (depth[v*64 + i] = level of v from roots[i])."""
    roots = _batch_roots(roots)
    graph._check_vertex_array(depth, np.uint32, 64)
    st = RunStats()
    _check(lib().gg_bfs_batch(graph.handle, ctypes.c_void_p(roots.ctypes.data), len(roots),
                              ctypes.c_void_p(depth.ptr), ctypes.byref(st)))
    return st


def bfs_batch(graph, roots):
    """Returns depths as a (len(roots), V) array (one row per root) and stats."""
    roots = _batch_roots(roots)
    depth = DeviceArray(graph.vertices * 64, np.uint32)
    st = bfs_batch_into(graph, roots, depth)
    d = depth.numpy().reshape(graph.vertices, 64)[:, :len(roots)].T.copy()
    return d, st.as_dict()


def sssp_batch_into(graph, roots, dist):
    roots = _batch_roots(roots)
    graph._check_vertex_array(dist, np.float32, 64)
    st = RunStats()
    _check(lib().gg_sssp_batch(graph.handle, ctypes.c_void_p(roots.ctypes.data), len(roots),
                               ctypes.c_void_p(dist.ptr), ctypes.byref(st)))
    return st


def sssp_batch(graph, roots):
    """Distances as a (len(roots), V) float32 array and stats."""
    roots = _batch_roots(roots)
    dist = DeviceArray(graph.vertices * 64, np.float32)
    st = sssp_batch_into(graph, roots, dist)
    d = dist.numpy().reshape(graph.vertices, 64)[:, :len(roots)].T.copy()
    return d, st.as_dict()


def bc_batch_into(graph, roots, dep, inv):
    roots = _batch_roots(roots)
    graph._check_vertex_array(dep, np.float64, 64)
    graph._check_vertex_array(inv, np.float64, 64)
    st = RunStats()
    _check(lib().gg_bc_batch(graph.handle, ctypes.c_void_p(roots.ctypes.data), len(roots),
                             ctypes.c_void_p(dep.ptr), ctypes.c_void_p(inv.ptr), ctypes.byref(st)))
    return st


def bc_batch(graph, roots):
    """(dependencies, inv_num_paths) as (len(roots), V) float64 arrays and stats."""
    roots = _batch_roots(roots)
    dep = DeviceArray(graph.vertices * 64, np.float64)
    inv = DeviceArray(graph.vertices * 64, np.float64)
    st = bc_batch_into(graph, roots, dep, inv)
    k = len(roots)
    d = dep.numpy().reshape(graph.vertices, 64)[:, :k].T.copy()
    i = inv.numpy().reshape(graph.vertices, 64)[:, :k].T.copy()
    return d, i, st.as_dict()


class VertexSubset:
    """Bitmap / VertexSubset (core/bitmap.hpp:20-61) in device memory."""

    def __init__(self, graph):
        h = ctypes.c_void_p()
        _check(lib().gg_subset_create(graph.handle, ctypes.byref(h)))
        self._h = h.value
        self.size = graph.vertices

    def clear(self):
        _check(lib().gg_subset_clear(ctypes.c_void_p(self._h)))

    def fill(self):
        _check(lib().gg_subset_fill(ctypes.c_void_p(self._h)))

    def set_bit(self, v):
        _check(lib().gg_subset_set_bit(ctypes.c_void_p(self._h), int(v)))

    def get_bit(self, v):
        b = ctypes.c_int()
        _check(lib().gg_subset_get_bit(ctypes.c_void_p(self._h), int(v), ctypes.byref(b)))
        return b.value

    def count(self):
        c = ctypes.c_uint64()
        _check(lib().gg_subset_count(ctypes.c_void_p(self._h), ctypes.byref(c)))
        return c.value

    def words(self):
        n = (self.size >> 6) + 1
        out = np.empty(n, dtype=np.uint64)
        p = lib().gg_subset_device_words(ctypes.c_void_p(self._h))
        _check(lib().gg_memcpy_d2h(out.ctypes.data, ctypes.c_void_p(p), out.nbytes))
        return out

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gg_subset_destroy(ctypes.c_void_p(self._h))
            self._h = None


class Server:
    """Query server over resident graphs (native, in libgg): continuous
    batching of same-graph, same-kind queries into batched traversals of up
    to 64 roots, on `workers` streams, with device memory admitted against
    `budget_bytes`.  Graphs must outlive the server."""

    def __init__(self, budget_bytes, workers=4):
        h = ctypes.c_void_p()
        _check(lib().gg_server_create(int(budget_bytes), int(workers), ctypes.byref(h)))
        self._h = h.value
        self._graphs = []

    def add_graph(self, graph):
        gid = ctypes.c_int()
        _check(lib().gg_server_add_graph(ctypes.c_void_p(self._h), graph.handle,
                                         ctypes.byref(gid)))
        self._graphs.append(graph)
        return gid.value

    def submit(self, graph_id, kind, root, want_result=True):
        qid = ctypes.c_uint64()
        _check(lib().gg_server_submit(ctypes.c_void_p(self._h), int(graph_id), int(kind),
                                      int(root), int(bool(want_result)), ctypes.byref(qid)))
        return qid.value, graph_id, kind

    def wait(self, ticket):
        """Blocks until the query is done; returns (result or None, latency_ms)."""
        qid, graph_id, kind = ticket
        lat = ctypes.c_double()
        out = np.empty(self._graphs[graph_id].vertices, dtype=_QUERY_DTYPE[kind])
        _check(lib().gg_server_wait(ctypes.c_void_p(self._h), qid, out.ctypes.data, out.nbytes,
                                    ctypes.byref(lat)))
        return out, lat.value

    def discard(self, ticket):
        """Waits for a query submitted with want_result=False; returns latency_ms."""
        lat = ctypes.c_double()
        _check(lib().gg_server_wait(ctypes.c_void_p(self._h), ticket[0], None, 0,
                                    ctypes.byref(lat)))
        return lat.value

    def replay(self, arrival_s, graph_ids, kinds, roots):
        """Open-loop replay of a query trace; returns (latencies_ms, wall_s)."""
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        gid = np.ascontiguousarray(graph_ids, dtype=np.int32)
        knd = np.ascontiguousarray(kinds, dtype=np.int32)
        rts = np.ascontiguousarray(roots, dtype=np.uint32)
        n = len(arr)
        if not (len(gid) == len(knd) == len(rts) == n):
            raise ValueError("replay: trace columns differ in length")
        lat = np.empty(n, dtype=np.float64)
        wall = ctypes.c_double()
        _check(lib().gg_server_replay(ctypes.c_void_p(self._h), n, arr.ctypes.data, gid.ctypes.data,
                                      knd.ctypes.data, rts.ctypes.data, lat.ctypes.data,
                                      ctypes.byref(wall)))
        return lat, wall.value

    def info(self):
        i = ServerInfo()
        _check(lib().gg_server_get_info(ctypes.c_void_p(self._h), ctypes.byref(i)))
        return i.as_dict()

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gg_server_destroy(ctypes.c_void_p(self._h))
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def prof_enable(on=True):
    _check(lib().gg_prof_enable(int(on)))


def prof_reset():
    _check(lib().gg_prof_reset())


def prof_read(kernel):
    ms, n = ctypes.c_double(), ctypes.c_uint64()
    _check(lib().gg_prof_read(kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value

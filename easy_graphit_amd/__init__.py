"""MI355X-native GraphIt GPU runtime — Python binding of the C ABI.

``libgraphit_amd.so`` (built from ``csrc/`` by ``build.py``) exports the
``gi_*`` entry points declared in ``include/graphit_amd.h``. This module binds
them with ctypes and mirrors the reference's runtime vocabulary
(``GraphT``/``load_graph``, ``VertexFrontier``, ``SimpleGPUSchedule``,
``HybridGPUSchedule``, ``edgeset_apply``, ``vertexset_apply``,
``VertexFrontierList``, ``GPUPriorityQueue``) plus the fused app drivers.

There is no CPU fallback: every call goes to the HIP library, and importing a
compute entry point without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgraphit_amd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "graphit_amd.h")

# --------------------------------------------------------------- constants
GI_OK = 0
ERRORS = {-1: "INVALID", -2: "HIP", -3: "OOM", -4: "UNSUPPORTED", -5: "RANGE", -6: "IO",
          -7: "UNDERFLOW"}

PUSH, PULL = 0, 1
PULL_BITMAP, PULL_BOOLMAP = 0, 1
FC_SPARSE, FC_BITMAP, FC_BOOLMAP = 0, 1, 2
FMT_SPARSE, FMT_BITMAP, FMT_BOOLMAP = 0, 1, 2
LB_VERTEX_BASED, LB_TWC, LB_TWCE, LB_WM, LB_CM, LB_STRICT, LB_EDGE_ONLY = range(7)
LB_NAMES = ["VERTEX_BASED", "TWC", "TWCE", "WM", "CM", "STRICT", "EDGE_ONLY"]
BLOCKED, UNBLOCKED = 0, 1
SELECT_HEURISTIC, SELECT_LEARNED, SELECT_S1, SELECT_S2 = 0, 1, 2, 3
LAYOUT_NATIVE, LAYOUT_DEGREE = 0, 1
(OP_BFS_PARENT, OP_BFS_DEPTH, OP_SSSP_RELAX, OP_CC_HOOK, OP_PR_PULL, OP_BC_FORWARD,
 OP_BC_BACKWARD) = range(7)
(VOP_FILL_I32, VOP_IOTA_I32, VOP_CC_PJUMP, VOP_PR_CONTRIB, VOP_PR_UPDATE, VOP_PR_RESET,
 VOP_BC_RESET, VOP_BC_MARK, VOP_BC_BACKWARD_V, VOP_BC_FINAL) = range(10)

PR_DAMP = float(np.float32(0.85))  # apps/pr.cpp:23 (float constant)


class GIError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gi error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


# --------------------------------------------------------------- structs
class Schedule(C.Structure):
    """SimpleGPUSchedule (include/graphit/schedule.h:26-118)."""
    _fields_ = [(n, C.c_int32) for n in (
        "direction", "pull_frontier_rep", "frontier_creation", "deduplication",
        "load_balancing", "edge_blocking", "block_size", "kernel_fusion", "max_cta",
        "cta_size")]

    def __init__(self, **kw):
        super().__init__()
        lib().gi_schedule_init(C.byref(self))
        for k, v in kw.items():
            setattr(self, k, v)

    def __repr__(self):
        d = "PULL" if self.direction == PULL else "PUSH"
        return f"Schedule({d},{LB_NAMES[self.load_balancing]},fc={self.frontier_creation},cta={self.cta_size})"


class HybridSchedule(C.Structure):
    """HybridGPUSchedule (include/graphit/schedule.h:122-139) + selector."""
    _fields_ = [("s1", Schedule), ("s2", Schedule), ("threshold", C.c_float),
                ("selector", C.c_int32)]

    def __init__(self, s1: Schedule, s2: Schedule, threshold: float = 0.0,
                 selector: int = SELECT_HEURISTIC):
        super().__init__()
        self.s1 = s1
        self.s2 = s2
        self.threshold = threshold
        self.selector = selector


class OpArgs(C.Structure):
    _fields_ = [("data0", C.c_void_p), ("data1", C.c_void_p), ("data2", C.c_void_p),
                ("i0", C.c_int32), ("i1", C.c_int32), ("f0", C.c_float), ("f1", C.c_float)]


class LevelRecord(C.Structure):
    _fields_ = [("level", C.c_int32), ("direction", C.c_int32), ("load_balancing", C.c_int32),
                ("frontier_size", C.c_int32), ("next_size", C.c_int32), ("reserved", C.c_int32),
                ("kernel_ms", C.c_double), ("alg_bytes", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class HostCSR(C.Structure):
    _fields_ = [("num_vertices", C.c_int32), ("num_edges", C.c_int64), ("directed", C.c_int32),
                ("offsets", C.POINTER(C.c_int32)), ("nbrs", C.POINTER(C.c_int32)),
                ("weights", C.POINTER(C.c_int32))]


LOAD_SYMMETRIZE = 1
LOAD_CACHE = 2


class RunStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("push_iterations", C.c_int32),
                ("pull_iterations", C.c_int32), ("reached", C.c_int32),
                ("edges_traversed", C.c_int64), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# --------------------------------------------------------------- loading
_LIB = None


def header_symbols() -> list[str]:
    """Every function the C ABI header declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gi_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libgraphit_amd.so. Raises when it has not been built: there is no
    fallback path."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing; run python -m easy_graphit_amd.build")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
    pp = C.POINTER(C.c_void_p)
    sig = {
        "gi_last_error": (C.c_char_p, []),
        "gi_version": (i32, []),
        "gi_device_count": (i32, []),
        "gi_malloc": (C.c_int, [pp, u64]),
        "gi_free": (C.c_int, [vp]),
        "gi_memcpy_h2d": (C.c_int, [vp, vp, u64, vp]),
        "gi_memcpy_d2h": (C.c_int, [vp, vp, u64, vp]),
        "gi_memcpy_d2d": (C.c_int, [vp, vp, u64, vp]),
        "gi_fill_i32": (C.c_int, [vp, i64, i32, vp]),
        "gi_fill_f32": (C.c_int, [vp, i64, C.c_float, vp]),
        "gi_fill_u8": (C.c_int, [vp, i64, C.c_uint8, vp]),
        "gi_stream_create": (C.c_int, [pp]),
        "gi_stream_destroy": (C.c_int, [vp]),
        "gi_stream_synchronize": (C.c_int, [vp]),
        "gi_timer_start": (C.c_int, [vp]),
        "gi_timer_stop": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "gi_schedule_init": (None, [C.POINTER(Schedule)]),
        "gi_graph_from_csr": (C.c_int, [i32, i64, vp, vp, vp, i32, pp]),
        "gi_graph_from_edges": (C.c_int, [i32, i64, vp, vp, i32, pp]),
        "gi_read_graph_file": (C.c_int, [C.c_char_p, i32, C.POINTER(HostCSR)]),
        "gi_host_csr_free": (None, [C.POINTER(HostCSR)]),
        "gi_write_graphit_bin": (C.c_int, [C.POINTER(HostCSR), C.c_char_p]),
        "gi_graph_load": (C.c_int, [C.c_char_p, i32, pp]),
        "gi_graph_rmat": (C.c_int, [i32, i32, u64, i32, pp]),
        "gi_graph_uniform": (C.c_int, [i32, i64, u64, i32, pp]),
        "gi_graph_grid": (C.c_int, [i32, i64, u64, pp]),
        "gi_graph_set_weights": (C.c_int, [vp, u64, i32, i32]),
        "gi_graph_num_vertices": (C.c_int, [vp, C.POINTER(i32)]),
        "gi_graph_num_edges": (C.c_int, [vp, C.POINTER(i64)]),
        "gi_graph_read_csr": (C.c_int, [vp, vp, vp, vp]),
        "gi_graph_device_csr": (C.c_int, [vp, pp, pp, pp]),
        "gi_graph_transpose": (C.c_int, [vp, pp]),
        "gi_graph_blocked": (C.c_int, [vp, i32, pp]),
        "gi_graph_destroy": (C.c_int, [vp]),
        "gi_graph_set_layout": (C.c_int, [vp, i32]),
        "gi_graph_layout": (i32, [vp]),
        "gi_frontier_create": (C.c_int, [i32, i32, pp]),
        "gi_frontier_destroy": (C.c_int, [vp]),
        "gi_frontier_add_vertex": (C.c_int, [vp, i32, vp]),
        "gi_frontier_clear": (C.c_int, [vp, vp]),
        "gi_frontier_size": (C.c_int, [vp, C.POINTER(i32), vp]),
        "gi_frontier_convert": (C.c_int, [vp, i32, vp]),
        "gi_frontier_format": (C.c_int, [vp, C.POINTER(i32)]),
        "gi_frontier_read": (C.c_int, [vp, vp, i32, C.POINTER(i32), vp]),
        "gi_frontier_swap": (C.c_int, [vp, i32, vp]),
        "gi_frontier_dedup": (C.c_int, [vp, vp]),
        "gi_edgeset_apply": (C.c_int, [vp, C.POINTER(Schedule), vp, vp, i32, C.POINTER(OpArgs), vp]),
        "gi_edgeset_apply_hybrid": (C.c_int, [vp, C.POINTER(HybridSchedule), vp, vp, i32,
                                              C.POINTER(OpArgs), vp]),
        "gi_vertexset_apply": (C.c_int, [vp, vp, i32, C.POINTER(OpArgs), vp]),
        "gi_frontier_list_create": (C.c_int, [i32, pp]),
        "gi_frontier_list_insert": (C.c_int, [vp, vp, vp]),
        "gi_frontier_list_retrieve": (C.c_int, [vp, vp, vp]),
        "gi_frontier_list_levels": (C.c_int, [vp, C.POINTER(i32)]),
        "gi_frontier_list_destroy": (C.c_int, [vp]),
        "gi_bfs": (C.c_int, [vp, i32, C.POINTER(HybridSchedule), vp, vp, C.POINTER(RunStats), vp]),
        "gi_bfs_batched": (C.c_int, [vp, vp, i32, vp, C.POINTER(RunStats), vp]),
        "gi_sssp_batched": (C.c_int, [vp, vp, i32, i32, i32, vp, C.POINTER(RunStats), vp]),
        "gi_bfs_static": (C.c_int, [vp, i32, C.POINTER(Schedule), vp, vp, vp]),
        "gi_bfs_profile": (C.c_int, [vp, i32, C.POINTER(HybridSchedule), vp,
                                     C.POINTER(LevelRecord), i32, C.POINTER(i32), vp]),
        "gi_graph_pick_sources": (C.c_int, [vp, u64, i32, vp]),
        "gi_reached_edges": (C.c_int, [vp, vp, C.POINTER(i64), vp]),
        "gi_selector_param_count": (i32, [i32]),
        "gi_selector_load": (C.c_int, [vp, i32, i32, vp, vp]),
        "gi_selector_clear": (C.c_int, []),
        "gi_selector_loaded": (i32, []),
        "gi_selector_loss_grad": (C.c_int, [vp, vp, vp, i32, i32, vp, vp, vp]),
        "gi_selector_train": (C.c_int, [vp, vp, vp, i32, i32, i32, C.c_float, C.c_float, vp, vp]),
        "gi_bfs_set_bins": (C.c_int, [vp, i32, i32, i32]),
        "gi_bfs_multi": (C.c_int, [vp, vp, i32, C.POINTER(HybridSchedule), vp, vp,
                                   C.POINTER(i32), vp]),
        "gi_graph_uniform_batch": (C.c_int, [i32, i32, i32, i32, u64, vp, pp]),
        "gi_bfs_sweep": (C.c_int, [vp, i32, i32, i32, i32, vp, vp, i32, C.POINTER(i32), vp]),
        "gi_pq_create": (C.c_int, [vp, vp, i32, i32, i32, pp]),
        "gi_pq_finished": (C.c_int, [vp, C.POINTER(i32), vp]),
        "gi_pq_dequeue": (C.c_int, [vp, pp, vp]),
        "gi_pq_cutoff": (C.c_int, [vp, C.POINTER(i32)]),
        "gi_pq_destroy": (C.c_int, [vp]),
        "gi_sssp": (C.c_int, [vp, i32, i32, C.POINTER(Schedule), vp, C.POINTER(RunStats), vp]),
        "gi_sssp_trace": (C.c_int, [vp, i32, i32, i32, vp, i32, vp, vp, C.POINTER(i32), vp]),
        "gi_cc": (C.c_int, [vp, C.POINTER(Schedule), vp, C.POINTER(RunStats), vp]),
        "gi_pagerank": (C.c_int, [vp, i32, C.c_float, C.POINTER(Schedule), vp,
                                  C.POINTER(RunStats), vp]),
        "gi_bc": (C.c_int, [vp, vp, i32, i32, vp, C.POINTER(RunStats), vp]),
        "gi_bc_schedule": (C.c_int, [vp, vp, i32, vp, vp, C.POINTER(RunStats), vp]),
    }
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
    _LIB = L
    return L


def check(rc: int) -> None:
    if rc != GI_OK:
        raise GIError(rc, lib().gi_last_error().decode(errors="replace"))


def device_count() -> int:
    return int(lib().gi_device_count())


# --------------------------------------------------------------- device memory
class DeviceArray:
    """A device buffer allocated through gi_malloc (typed, 1-D)."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        check(lib().gi_malloc(C.byref(p), max(1, self.n * self.dtype.itemsize)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        check(lib().gi_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes, stream))
        return d

    def numpy(self, stream=None):
        out = np.empty(self.n, self.dtype)
        check(lib().gi_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, stream))
        return out

    def fill(self, value, stream=None):
        if self.dtype == np.int32:
            check(lib().gi_fill_i32(self.ptr, self.n, int(value), stream))
        elif self.dtype == np.float32:
            check(lib().gi_fill_f32(self.ptr, self.n, float(value), stream))
        elif self.dtype == np.uint8:
            check(lib().gi_fill_u8(self.ptr, self.n, int(value), stream))
        else:
            self.copy_from(np.full(self.n, value, self.dtype), stream)
        return self

    def copy_from(self, a, stream=None):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        check(lib().gi_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream))

    def free(self):
        if getattr(self, "ptr", None):
            lib().gi_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(stream=None):
    check(lib().gi_stream_synchronize(stream))


def copy_device(dst: DeviceArray, src: DeviceArray, stream=None):
    """Device-to-device copy of src into dst (equal byte sizes)."""
    nbytes = src.n * src.dtype.itemsize
    if dst.n * dst.dtype.itemsize != nbytes:
        raise ValueError("copy_device: sizes differ")
    check(lib().gi_memcpy_d2d(dst.ptr, src.ptr, nbytes, stream))


# --------------------------------------------------------------- graph
class Graph:
    """Device CSR graph (GraphT<int> + load_graph, runtime/infra_gpu/graph.h)."""

    def __init__(self, handle, owner=True, parent=None):
        self.h = handle
        self.owner = owner
        # a derived view (transpose, blocked) lives inside its base graph:
        # hold the base so it outlives the view, and let the base invalidate
        # its views when it drops them (set_weights)
        self._parent = parent
        self._views = []

    @property
    def num_vertices(self) -> int:
        v = C.c_int32()
        check(lib().gi_graph_num_vertices(self.h, C.byref(v)))
        return v.value

    @property
    def num_edges(self) -> int:
        e = C.c_int64()
        check(lib().gi_graph_num_edges(self.h, C.byref(e)))
        return e.value

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, seed: int = 27491095,
             symmetrize: bool = True):
        h = C.c_void_p()
        check(lib().gi_graph_rmat(scale, edge_factor, seed, int(symmetrize), C.byref(h)))
        return cls(h.value)

    @classmethod
    def uniform(cls, num_vertices: int, num_pairs: int, seed: int, symmetrize: bool = True):
        h = C.c_void_p()
        check(lib().gi_graph_uniform(num_vertices, num_pairs, seed, int(symmetrize), C.byref(h)))
        return cls(h.value)

    @classmethod
    def uniform_batch(cls, num_graphs: int, vmin: int, vmax: int, avg_degree: int, seed: int):
        """Many small uniform graphs packed in one symmetric CSR; returns
        (graph, vertex_offsets[num_graphs+1])."""
        offs = np.empty(num_graphs + 1, np.int32)
        h = C.c_void_p()
        check(lib().gi_graph_uniform_batch(num_graphs, vmin, vmax, avg_degree, seed,
                                           offs.ctypes.data, C.byref(h)))
        return cls(h.value), offs

    def set_bins(self, lane_max: int, hub_min: int, blocks_per_cu: int = 8):
        check(lib().gi_bfs_set_bins(self.h, lane_max, hub_min, blocks_per_cu))

    @classmethod
    def grid(cls, width: int, num_random: int, seed: int):
        h = C.c_void_p()
        check(lib().gi_graph_grid(width, num_random, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr(cls, offsets, nbrs, weights=None, directed: bool = False):
        off = np.ascontiguousarray(offsets, np.int32)
        nb = np.ascontiguousarray(nbrs, np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, np.int32)
        h = C.c_void_p()
        check(lib().gi_graph_from_csr(len(off) - 1, len(nb), off.ctypes.data, nb.ctypes.data,
                                      None if w is None else w.ctypes.data, int(directed),
                                      C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path: str, symmetrize: bool = False, cache: bool = False):
        """A graph file (.el .wel .gr .graph .mtx .sg .wsg .graphit_bin) on
        the device; `cache` reads / writes the .graphit_bin cache beside it."""
        h = C.c_void_p()
        flags = (LOAD_SYMMETRIZE if symmetrize else 0) | (LOAD_CACHE if cache else 0)
        check(lib().gi_graph_load(os.fsencode(path), flags, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, num_vertices: int, src, dst, symmetrize: bool):
        s = np.ascontiguousarray(src, np.int32)
        d = np.ascontiguousarray(dst, np.int32)
        h = C.c_void_p()
        check(lib().gi_graph_from_edges(num_vertices, len(s), s.ctypes.data, d.ctypes.data,
                                        int(symmetrize), C.byref(h)))
        return cls(h.value)

    def pick_sources(self, seed: int, count: int) -> np.ndarray:
        """SourcePicker semantics (benchmark.h:48-71): skip degree-0 vertices."""
        out = np.empty(count, np.int32)
        check(lib().gi_graph_pick_sources(self.h, seed, count, out.ctypes.data))
        return out

    def reached_edges(self, depth: "DeviceArray", stream=None) -> int:
        e = C.c_int64()
        check(lib().gi_reached_edges(self.h, depth.ptr, C.byref(e), stream))
        return e.value

    def set_weights(self, seed: int, wmin: int, wmax: int):
        check(lib().gi_graph_set_weights(self._handle(), seed, wmin, wmax))
        # the library rebuilt the derived views: the old handles are gone
        for v in self._views:
            v.h = None
        self._views = []
        return self

    def _handle(self):
        if self.h is None:
            raise GIError(-1, "graph handle is no longer valid (destroyed, or a view dropped "
                              "by its base graph's set_weights)")
        return self.h

    def _view(self, h) -> "Graph":
        if h == self.h:  # a symmetric graph is its own transpose
            return self
        for v in self._views:
            if v.h == h:
                return v
        v = Graph(h, owner=False, parent=self)
        self._views.append(v)
        return v

    def read_csr(self, weights: bool = False):
        n, m = self.num_vertices, self.num_edges
        off = np.empty(n + 1, np.int32)
        nb = np.empty(max(m, 1), np.int32)
        w = np.empty(max(m, 1), np.int32) if weights else None
        check(lib().gi_graph_read_csr(self.h, off.ctypes.data, nb.ctypes.data,
                                      None if w is None else w.ctypes.data))
        return off, nb[:m], (None if w is None else w[:m])

    def transpose(self) -> "Graph":
        h = C.c_void_p()
        check(lib().gi_graph_transpose(self._handle(), C.byref(h)))
        return self._view(h.value)

    def blocked(self, block_size: int) -> "Graph":
        h = C.c_void_p()
        check(lib().gi_graph_blocked(self._handle(), block_size, C.byref(h)))
        return self._view(h.value)

    def set_layout(self, layout: int):
        """LAYOUT_DEGREE: fused BFS and CC run on a degree-ordered copy
        (results mapped back to the original ids); LAYOUT_NATIVE drops it."""
        check(lib().gi_graph_set_layout(self.h, layout))
        return self

    @property
    def layout(self) -> int:
        return int(lib().gi_graph_layout(self.h))

    def destroy(self):
        if self.owner and self.h:
            for v in self._views:
                v.h = None
            self._views = []
            lib().gi_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


load_graph = Graph  # name of the reference entry point (graph.h:236)


# --------------------------------------------------------------- frontier
class Frontier:
    """VertexFrontier (runtime/infra_gpu/vertex_frontier.h:26-166)."""

    def __init__(self, num_vertices: int, init_all: bool = False):
        h = C.c_void_p()
        check(lib().gi_frontier_create(num_vertices, int(init_all), C.byref(h)))
        self.h = h.value
        self.num_vertices = num_vertices

    def add_vertex(self, v: int, stream=None):
        check(lib().gi_frontier_add_vertex(self.h, v, stream))

    addVertex = add_vertex

    def clear(self, stream=None):
        check(lib().gi_frontier_clear(self.h, stream))

    def size(self, stream=None) -> int:
        n = C.c_int32()
        check(lib().gi_frontier_size(self.h, C.byref(n), stream))
        return n.value

    def convert(self, fmt: int, stream=None):
        check(lib().gi_frontier_convert(self.h, fmt, stream))

    @property
    def format(self) -> int:
        f = C.c_int32()
        check(lib().gi_frontier_format(self.h, C.byref(f)))
        return f.value

    def read(self, stream=None) -> np.ndarray:
        cap = self.num_vertices
        out = np.empty(cap, np.int32)
        n = C.c_int32()
        check(lib().gi_frontier_read(self.h, out.ctypes.data, cap, C.byref(n), stream))
        return out[: n.value]

    def swap(self, fmt: int, stream=None):
        check(lib().gi_frontier_swap(self.h, fmt, stream))

    def dedup(self, stream=None):
        check(lib().gi_frontier_dedup(self.h, stream))

    def destroy(self):
        if self.h and not getattr(self, "borrowed", False):
            lib().gi_frontier_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


new_vertex_subset = Frontier  # runtime.cpp name for create_new_vertex_set


class FrontierList:
    """VertexFrontierList (runtime/infra_gpu/list.h:18-116)."""

    def __init__(self, max_elems: int):
        h = C.c_void_p()
        check(lib().gi_frontier_list_create(max_elems, C.byref(h)))
        self.h = h.value

    def insert(self, f: Frontier, stream=None):
        check(lib().gi_frontier_list_insert(self.h, f.h, stream))

    def retrieve(self, f: Frontier, stream=None):
        check(lib().gi_frontier_list_retrieve(self.h, f.h, stream))

    @property
    def levels(self) -> int:
        n = C.c_int32()
        check(lib().gi_frontier_list_levels(self.h, C.byref(n)))
        return n.value

    def __del__(self):
        try:
            if self.h:
                lib().gi_frontier_list_destroy(self.h)
        except Exception:
            pass


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    return x


def op_args(data0=None, data1=None, data2=None, i0=0, i1=0, f0=0.0, f1=0.0) -> OpArgs:
    a = OpArgs()
    a.data0, a.data1, a.data2 = _ptr(data0), _ptr(data1), _ptr(data2)
    a.i0, a.i1, a.f0, a.f1 = i0, i1, f0, f1
    return a


def edgeset_apply(g: Graph, s, frm: Frontier | None, to: Frontier | None, op: int,
                  args: OpArgs, stream=None):
    """edgeset_apply(s).from(frm).apply / apply_modified (operators.cpp:513-650)."""
    if isinstance(s, HybridSchedule):
        check(lib().gi_edgeset_apply_hybrid(g.h, C.byref(s), frm.h if frm else None,
                                            to.h if to else None, op, C.byref(args), stream))
    else:
        check(lib().gi_edgeset_apply(g.h, C.byref(s), frm.h if frm else None,
                                     to.h if to else None, op, C.byref(args), stream))


def vertexset_apply(g: Graph, f: Frontier | None, vop: int, args: OpArgs, stream=None):
    check(lib().gi_vertexset_apply(g.h, f.h if f else None, vop, C.byref(args), stream))


# --------------------------------------------------------------- app drivers
def bfs(g: Graph, source: int, schedule: HybridSchedule | None = None,
        want_parent: bool = False, stream=None):
    """Depths (and optionally CAS parents) from source; returns numpy arrays."""
    n = g.num_vertices
    depth = DeviceArray(n, np.int32)
    parent = DeviceArray(n, np.int32) if want_parent else None
    st = RunStats()
    check(lib().gi_bfs(g.h, source, C.byref(schedule) if schedule is not None else None,
                       depth.ptr, parent.ptr if parent else None, C.byref(st), stream))
    return depth.numpy(stream), (parent.numpy(stream) if parent else None), st.as_dict()


def bfs_device(g: Graph, source: int, schedule, depth: DeviceArray, parent=None, stream=None,
               stats: bool = False):
    """BFS into caller-owned device arrays; with stats=False nothing but the
    traversal runs (the form timed by the benchmark)."""
    st = RunStats() if stats else None
    check(lib().gi_bfs(g.h, source, C.byref(schedule) if schedule is not None else None,
                       depth.ptr, parent.ptr if parent else None,
                       C.byref(st) if st is not None else None, stream))
    return st.as_dict() if st is not None else None


def bfs_batched(g: Graph, sources, depth: DeviceArray | None = None, stream=None):
    """Per-source depths [len(sources), V] (gi_bfs_batched: 64 sources per
    pass). With `depth` (a device array of len(sources) * V int32) the result
    stays on the device and only the statistics are returned."""
    src = np.ascontiguousarray(sources, np.int32)
    n, V = len(src), g.num_vertices
    own = depth is None
    if own:
        depth = DeviceArray(max(n, 1) * V, np.int32)
    st = RunStats()
    check(lib().gi_bfs_batched(g.h, src.ctypes.data, n, depth.ptr, C.byref(st), stream))
    if own:
        return depth.numpy(stream)[: n * V].reshape(n, V), st.as_dict()
    return None, st.as_dict()


def sssp_batched(g: Graph, sources, delta: int, batch: int = 0, dist: DeviceArray | None = None,
                 stream=None):
    """Per-source SSSP distances [len(sources), V] (gi_sssp_batched). With
    `dist` (len(sources) * V int32 on the device) only the statistics are
    returned."""
    src = np.ascontiguousarray(sources, np.int32)
    n, V = len(src), g.num_vertices
    own = dist is None
    if own:
        dist = DeviceArray(max(n, 1) * V, np.int32)
    st = RunStats()
    check(lib().gi_sssp_batched(g.h, src.ctypes.data, n, delta, batch, dist.ptr, C.byref(st), stream))
    if own:
        return dist.numpy(stream)[: n * V].reshape(n, V), st.as_dict()
    return None, st.as_dict()


BFS_STATIC_MAX_VERTICES = 1 << 21  # GI_BFS_STATIC_MAX_VERTICES


def bfs_static(g: Graph, source: int, schedule: Schedule | None, depth: DeviceArray,
               parent: DeviceArray | None = None, stream=None):
    """BFS that allocates nothing (gi_bfs_static): the unfused push loop with
    its frontier and visited bitmap in the library's fixed device array;
    results go into the caller's device arrays."""
    check(lib().gi_bfs_static(g.h, source, C.byref(schedule) if schedule is not None else None,
                              depth.ptr, parent.ptr if parent else None, stream))


def bfs_multi(g: Graph, sources, schedule: HybridSchedule, depth: DeviceArray | None = None,
              parent: DeviceArray | None = None, stream=None):
    """Multi-source fused BFS (distance to the nearest source)."""
    src = np.ascontiguousarray(sources, np.int32)
    own = depth is None
    if own:
        depth = DeviceArray(g.num_vertices, np.int32)
    levels = C.c_int32()
    check(lib().gi_bfs_multi(g.h, src.ctypes.data, len(src), C.byref(schedule), depth.ptr,
                             parent.ptr if parent else None, C.byref(levels), stream))
    return (depth.numpy(stream) if own else None), levels.value


def bfs_profile(g: Graph, source: int, schedule, depth: DeviceArray, stream=None):
    recs = (LevelRecord * 4096)()
    n = C.c_int32()
    check(lib().gi_bfs_profile(g.h, source, C.byref(schedule) if schedule is not None else None,
                               depth.ptr, recs, 4096, C.byref(n), stream))
    return [recs[i].as_dict() for i in range(min(n.value, 4096))]


BFS_MODES = 7
BFS_MODE_NAMES = ["push_thread", "push_wave", "push_twce", "push_cm", "push_twc",
                  "pull_thread", "pull_group"]
SEL_INPUTS, SEL_HIDDEN = 16, 64


def bfs_sweep(g: Graph, source: int, mode_mask: int = (1 << BFS_MODES) - 1, follow: int = -1,
              reps: int = 2, max_levels: int = 512, stream=None):
    """Time every fused-BFS step mode at every level (dataset recorder)."""
    feats = np.zeros((max_levels, SEL_INPUTS), np.float32)
    times = np.zeros((max_levels, BFS_MODES), np.float32)
    n = C.c_int32()
    check(lib().gi_bfs_sweep(g.h, source, mode_mask, follow, reps, feats.ctypes.data,
                             times.ctypes.data, max_levels, C.byref(n), stream))
    return feats[: n.value].copy(), times[: n.value].copy()


def selector_param_count(n_out: int = BFS_MODES) -> int:
    return int(lib().gi_selector_param_count(n_out))


def selector_load(params, n_out: int, allowed_mask: int, feat_mean, feat_scale):
    p = np.ascontiguousarray(params, np.float32)
    assert p.size == selector_param_count(n_out)
    mu = np.ascontiguousarray(feat_mean, np.float32)
    sc = np.ascontiguousarray(feat_scale, np.float32)
    check(lib().gi_selector_load(p.ctypes.data, n_out, allowed_mask, mu.ctypes.data,
                                 sc.ctypes.data))


def selector_clear():
    check(lib().gi_selector_clear())


def selector_loss_grad(X, T, M, params, n_out: int):
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    M = np.ascontiguousarray(M, np.float32)
    p = np.ascontiguousarray(params, np.float32)
    g = np.zeros_like(p)
    loss = np.zeros(1, np.float32)
    check(lib().gi_selector_loss_grad(X.ctypes.data, T.ctypes.data, M.ctypes.data, X.shape[0],
                                      n_out, p.ctypes.data, g.ctypes.data, loss.ctypes.data))
    return float(loss[0]), g


def selector_train(X, T, M, params, n_out: int, epochs: int, lr: float = 3e-3,
                   weight_decay: float = 0.0):
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    M = np.ascontiguousarray(M, np.float32)
    p = np.ascontiguousarray(params, np.float32).copy()
    hist = np.zeros(max(epochs, 1), np.float32)
    check(lib().gi_selector_train(X.ctypes.data, T.ctypes.data, M.ctypes.data, X.shape[0], n_out,
                                  epochs, lr, weight_decay, p.ctypes.data, hist.ctypes.data))
    return p, hist[:epochs]


class PriorityQueue:
    """GPUPriorityQueue<int> (runtime/infra_gpu/gpu_priority_queue.h:21-247)."""

    def __init__(self, g: Graph, prio: DeviceArray, initial: int, delta: int, source: int):
        h = C.c_void_p()
        check(lib().gi_pq_create(g.h, prio.ptr, initial, delta, source, C.byref(h)))
        self.h = h.value
        self.g = g

    def finished(self, stream=None) -> bool:
        f = C.c_int32()
        check(lib().gi_pq_finished(self.h, C.byref(f), stream))
        return bool(f.value)

    def dequeue(self, stream=None) -> "Frontier":
        h = C.c_void_p()
        check(lib().gi_pq_dequeue(self.h, C.byref(h), stream))
        f = Frontier.__new__(Frontier)
        f.h = h.value
        f.num_vertices = self.g.num_vertices
        f.borrowed = True
        return f

    def prio_cutoff(self) -> int:
        c = C.c_int32()
        check(lib().gi_pq_cutoff(self.h, C.byref(c)))
        return c.value

    def __del__(self):
        try:
            if self.h:
                lib().gi_pq_destroy(self.h)
        except Exception:
            pass


def read_graph_file(path: str, symmetrize: bool = False):
    """Host CSR of a graph file, read without a device:
    (offsets, nbrs, weights or None, directed)."""
    c = HostCSR()
    check(lib().gi_read_graph_file(os.fsencode(path), LOAD_SYMMETRIZE if symmetrize else 0,
                                   C.byref(c)))
    try:
        n, m = c.num_vertices, c.num_edges
        off = np.ctypeslib.as_array(c.offsets, (n + 1,)).copy()
        nb = np.ctypeslib.as_array(c.nbrs, (max(m, 1),))[:m].copy()
        w = None
        if c.weights:
            w = np.ctypeslib.as_array(c.weights, (max(m, 1),))[:m].copy()
        return off, nb, w, bool(c.directed)
    finally:
        lib().gi_host_csr_free(C.byref(c))


def write_graphit_bin(path: str, offsets, nbrs, weights=None) -> None:
    """The reference's .graphit_bin cache of a host CSR."""
    off = np.ascontiguousarray(offsets, np.int32)
    nb = np.ascontiguousarray(nbrs, np.int32)
    w = None if weights is None else np.ascontiguousarray(weights, np.int32)
    i32p = C.POINTER(C.c_int32)
    c = HostCSR(len(off) - 1, len(nb), 1, off.ctypes.data_as(i32p), nb.ctypes.data_as(i32p),
                w.ctypes.data_as(i32p) if w is not None else None)
    check(lib().gi_write_graphit_bin(C.byref(c), os.fsencode(path)))


def sssp(g: Graph, source: int, delta: int, schedule: Schedule | None = None, stream=None):
    n = g.num_vertices
    dist = DeviceArray(n, np.int32)
    st = RunStats()
    check(lib().gi_sssp(g.h, source, delta, C.byref(schedule) if schedule is not None else None,
                        dist.ptr, C.byref(st), stream))
    return dist.numpy(stream), st.as_dict()


def sssp_trace(g: Graph, source: int, delta: int, pull_min: int = -1, max_slots: int = 4096,
               stream=None):
    """Fused SSSP run slot by slot: (dist, kernel_ms[slots, 6], info[slots, 5])."""
    dist = DeviceArray(g.num_vertices, np.int32)
    ms = np.zeros((max_slots, 6), np.float32)
    info = np.zeros((max_slots, 5), np.int64)
    n = C.c_int32()
    check(lib().gi_sssp_trace(g.h, source, delta, pull_min, dist.ptr, max_slots, ms.ctypes.data,
                              info.ctypes.data, C.byref(n), stream))
    return dist.numpy(stream), ms[: n.value].copy(), info[: n.value].copy()


def cc(g: Graph, schedule: Schedule | None = None, stream=None):
    n = g.num_vertices
    lab = DeviceArray(n, np.int32)
    st = RunStats()
    check(lib().gi_cc(g.h, C.byref(schedule) if schedule is not None else None, lab.ptr,
                      C.byref(st), stream))
    return lab.numpy(stream), st.as_dict()


def pagerank(g: Graph, iters: int = 20, damping: float = PR_DAMP,
             schedule: Schedule | None = None, stream=None):
    n = g.num_vertices
    rank = DeviceArray(n, np.float32)
    st = RunStats()
    check(lib().gi_pagerank(g.h, iters, damping,
                            C.byref(schedule) if schedule is not None else None, rank.ptr,
                            C.byref(st), stream))
    return rank.numpy(stream), st.as_dict()


def bc(g: Graph, sources, batch: int = 0, schedule: Schedule | None = None, stream=None):
    """Brandes BC summed over `sources`. batch > 1: batches of sources on the
    device-driven engine; otherwise one source at a time, by `schedule` when
    given (kernel_fusion=1: the fused single-source mode), else the app's
    unfused push TWCE loop."""
    n = g.num_vertices
    src = np.ascontiguousarray(sources, np.int32)
    scores = DeviceArray(n, np.float32)
    st = RunStats()
    if schedule is not None:
        if batch > 1:
            raise ValueError("pass either batch > 1 or a schedule")
        check(lib().gi_bc_schedule(g.h, src.ctypes.data, len(src), C.byref(schedule), scores.ptr,
                                   C.byref(st), stream))
    else:
        check(lib().gi_bc(g.h, src.ctypes.data, len(src), batch, scores.ptr, C.byref(st), stream))
    return scores.numpy(stream), st.as_dict()

"""CPU oracle for the GraphIt traversal path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, and only as the checker (or the timed CPU baseline).
The product library never imports it.

numpy wrappers over ``liboracle.so`` (serial C, ``oracle.c``). Each C function
cites the reference file:line it restates; how each is pinned is listed at the
top of ``oracle.c``. ``ref()`` loads ``_ref/libgapbs_ref.so``, the reference's
own GAPBS builder and reader compiled from its sources, when it was built;
``files.py`` restates the graph file readers in plain Python.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(ref_root: str | None = None) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    root = ref_root or "/root/reference"
    if os.path.isdir(os.path.join(root, "runtime", "infra_gapbs")):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref", f"REF_ROOT={root}"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        u64, i64, i32 = C.c_uint64, C.c_int64, C.c_int32
        sig = {
            "or_splitmix64_mix": (u64, [u64]),
            "or_hash64": (u64, [u64, u64, u64]),
            "or_permute_id": (C.c_uint32, [C.c_uint32, C.c_int, u64]),
            "or_rmat_edges": (None, [C.c_int, i64, u64, C.c_int, i32p, i32p]),
            "or_grid_edges": (i64, [C.c_int, i64, u64, i32p, i32p]),
            "or_uniform_edges": (None, [i32, i64, u64, i32p, i32p]),
            "or_edge_weight": (i32, [u64, i32, i32, i32, i32]),
            "or_uniform_batch_edges": (i64, [i32, i32, i32, i32, u64, i32p, C.c_void_p,
                                             C.c_void_p]),
            "or_build_csr": (i64, [i32, i64, i32p, i32p, C.c_int, i32p, i32p]),
            "or_fill_weights": (None, [i32, i32p, i32p, u64, i32, i32, i32p]),
            "or_transpose": (None, [i32, i32p, i32p, C.c_void_p, i32p, i32p, C.c_void_p]),
            "or_pick_sources": (i32, [i32, i32p, u64, i32, i32p]),
            "or_bfs_depth": (i32, [i32, i32p, i32p, i32, i32p]),
            "or_bfs_depth_omp": (i32, [i32, i32p, i32p, i32, i32, i32, i32p, C.POINTER(C.c_int32)]),
            "or_bfs_parent_valid": (i32, [i32, i32p, i32p, i32, i32p, i32p]),
            "or_bfs_multi_depth": (i32, [i32, i32p, i32p, i32p, i32, i32p]),
            "or_bfs_pull_examined": (i64, [i32, i32p, i32p, i32p, i32]),
            "or_sssp_dijkstra": (None, [i32, i32p, i32p, i32p, i32, i32p]),
            "or_sssp_delta_stepping": (i64, [i32, i32p, i32p, i32p, i32, i32, i32p]),
            "or_cc_labels": (None, [i32, i32p, i32p, i32p]),
            "or_sssp_delta_stepping_omp": (i64, [i32, i32p, i32p, i32p, i32, i32, i32, i32p,
                                                 C.POINTER(C.c_int32)]),
            "or_cc_omp": (i32, [i32, i32p, i32p, i32, i32p, C.POINTER(C.c_int32)]),
            "or_cc_label_propagation": (i32, [i32, i32p, i32p, i32p]),
            "or_set_threads": (None, [i32]),
            "or_max_threads": (i32, []),
            "or_pagerank": (None, [i32, i32p, i32p, i32, C.c_double, f64p]),
            "or_bc": (None, [i32, i32p, i32p, i32p, i32, f64p]),
            "or_mlp_forward": (None, [i32, i32, i32, i32] + [f64p] * 10),
            "or_mlp_loss_backward": (C.c_double, [i32, i32, i32, i32] + [f64p] * 15),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ref():
    """The reference's GAPBS builder compiled from its sources, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libgapbs_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_build_csr.restype = C.c_int64
        R.ref_build_csr.argtypes = [C.c_int32, C.c_int64, i32p, i32p, C.c_int, i32p, i32p]
        R.ref_rmat_edges.restype = C.c_int64
        R.ref_rmat_edges.argtypes = [C.c_int, C.c_int, i32p, i32p]
        R.ref_read_graph.restype = C.c_int64
        R.ref_read_graph.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        R.ref_read_graph_copy.restype = None
        R.ref_read_graph_copy.argtypes = [i32p, i32p, i32p]
        R.ref_write_serialized.restype = None
        R.ref_write_serialized.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_char_p]
        _REF = R
    return _REF


def ref_read_graph(path: str, symmetrize: bool = False):
    """A graph file read by the reference's own GAPBS reader and builder
    (oracle/_ref), as load_graph reads it: (offsets, nbrs, weights or None,
    directed), or None when _ref was not built. Only for valid files: the
    reference exits the process on errors."""
    R = ref()
    if R is None:
        return None
    e, d = C.c_int64(), C.c_int()
    n = R.ref_read_graph(os.fsencode(path), int(symmetrize), C.byref(e), C.byref(d))
    m = e.value
    off = np.empty(n + 1, np.int32)
    nb = np.empty(max(m, 1), np.int32)
    w = np.empty(max(m, 1), np.int32)
    R.ref_read_graph_copy(off, nb, w)
    weighted = not path.endswith(".sg")
    return off, nb[:m], (w[:m] if weighted else None), bool(d.value)


def ref_write_serialized(text_path: str, symmetrize: bool, weighted: bool, out_path: str) -> bool:
    """GAPBS builds a text graph file and writes it serialized (.sg/.wsg)."""
    R = ref()
    if R is None:
        return False
    R.ref_write_serialized(os.fsencode(text_path), int(symmetrize), int(weighted), os.fsencode(out_path))
    return True


from .files import graphit_bin_bytes, read_graph_file  # noqa: E402


# ---------------------------------------------------------------- generators
def hash64(seed: int, stream: int, counter: int) -> int:
    return lib().or_hash64(seed, stream, counter)


def rmat_edges(scale: int, edge_factor: int, seed: int, permute: bool = True):
    m = (1 << scale) * edge_factor
    s = np.empty(m, np.int32)
    d = np.empty(m, np.int32)
    lib().or_rmat_edges(scale, m, seed, int(permute), s, d)
    return s, d


def grid_edges(width: int, num_random: int, seed: int):
    m = 2 * width * (width - 1) + num_random
    s = np.empty(m, np.int32)
    d = np.empty(m, np.int32)
    k = lib().or_grid_edges(width, num_random, seed, s, d)
    assert k == m
    return s, d


def uniform_edges(num_vertices: int, num_edges: int, seed: int):
    s = np.empty(num_edges, np.int32)
    d = np.empty(num_edges, np.int32)
    lib().or_uniform_edges(num_vertices, num_edges, seed, s, d)
    return s, d


def uniform_batch_edges(num_graphs: int, vmin: int, vmax: int, avg_degree: int, seed: int):
    """Packed small uniform graphs: (vertex_offsets[G+1], src, dst)."""
    offs = np.empty(num_graphs + 1, np.int32)
    m = lib().or_uniform_batch_edges(num_graphs, vmin, vmax, avg_degree, seed, offs, None, None)
    s = np.empty(m, np.int32)
    d = np.empty(m, np.int32)
    lib().or_uniform_batch_edges(num_graphs, vmin, vmax, avg_degree, seed, offs,
                                 s.ctypes.data, d.ctypes.data)
    return offs, s, d


class CSR:
    """Host CSR: offsets int32[V+1], nbrs int32[E], optional weights int32[E]."""

    def __init__(self, offsets, nbrs, weights=None, directed=False):
        self.offsets = np.ascontiguousarray(offsets, np.int32)
        self.nbrs = np.ascontiguousarray(nbrs, np.int32)
        self.weights = None if weights is None else np.ascontiguousarray(weights, np.int32)
        self.directed = directed

    @property
    def num_vertices(self) -> int:
        return len(self.offsets) - 1

    @property
    def num_edges(self) -> int:
        return int(self.offsets[-1])

    def degrees(self):
        return np.diff(self.offsets).astype(np.int32)


def build_csr(num_vertices: int, src, dst, symmetrize: bool) -> CSR:
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    off = np.empty(num_vertices + 1, np.int32)
    cap = len(src) * (2 if symmetrize else 1)
    nb = np.empty(max(cap, 1), np.int32)
    e = lib().or_build_csr(num_vertices, len(src), src, dst, int(symmetrize), off, nb)
    return CSR(off, nb[:e].copy(), directed=not symmetrize)


def ref_build_csr(num_vertices: int, src, dst, symmetrize: bool):
    R = ref()
    if R is None:
        return None
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    off = np.empty(num_vertices + 1, np.int32)
    nb = np.empty(max(len(src) * 2, 1), np.int32)
    e = R.ref_build_csr(num_vertices, len(src), src, dst, int(symmetrize), off, nb)
    return CSR(off, nb[:e].copy(), directed=not symmetrize)


def fill_weights(g: CSR, seed: int, wmin: int, wmax: int) -> CSR:
    w = np.empty(max(g.num_edges, 1), np.int32)
    lib().or_fill_weights(g.num_vertices, g.offsets, g.nbrs, seed, wmin, wmax, w)
    g.weights = w[: g.num_edges].copy()
    return g


def transpose(g: CSR) -> CSR:
    n, m = g.num_vertices, g.num_edges
    to = np.empty(n + 1, np.int32)
    tn = np.empty(max(m, 1), np.int32)
    tw = np.empty(max(m, 1), np.int32) if g.weights is not None else None
    lib().or_transpose(n, g.offsets, g.nbrs,
                       None if g.weights is None else g.weights.ctypes.data,
                       to, tn, None if tw is None else tw.ctypes.data)
    return CSR(to, tn[:m].copy(), None if tw is None else tw[:m].copy(), g.directed)


def pick_sources(g: CSR, seed: int, count: int):
    out = np.empty(count, np.int32)
    got = lib().or_pick_sources(g.num_vertices, g.offsets, seed, count, out)
    if got < 0:
        raise ValueError("graph has no edges")
    return out


# ---------------------------------------------------------------- algorithms
def bfs_depth_omp(g: CSR, source: int, threads: int = 0):
    """Direction-optimizing OpenMP BFS (the bench's CPU baseline): (depth,
    threads used). Bottom-up steps only on symmetric graphs."""
    d = np.empty(g.num_vertices, np.int32)
    used = C.c_int32(0)
    lib().or_bfs_depth_omp(g.num_vertices, g.offsets, g.nbrs, source, int(not g.directed), threads, d,
                           C.byref(used))
    return d, used.value


def bfs_depth(g: CSR, source: int):
    d = np.empty(g.num_vertices, np.int32)
    lib().or_bfs_depth(g.num_vertices, g.offsets, g.nbrs, source, d)
    return d


def bfs_multi_depth(g: CSR, sources):
    d = np.empty(g.num_vertices, np.int32)
    src = np.ascontiguousarray(sources, np.int32)
    lib().or_bfs_multi_depth(g.num_vertices, g.offsets, g.nbrs, src, len(src), d)
    return d


def bfs_parent_violations(g: CSR, source: int, depth, parent) -> int:
    return lib().or_bfs_parent_valid(g.num_vertices, g.offsets, g.nbrs, source,
                                     np.ascontiguousarray(depth, np.int32),
                                     np.ascontiguousarray(parent, np.int32))


def bfs_pull_examined(g: CSR, depth, level: int) -> int:
    return lib().or_bfs_pull_examined(g.num_vertices, g.offsets, g.nbrs,
                                      np.ascontiguousarray(depth, np.int32), level)


def sssp(g: CSR, source: int):
    d = np.empty(g.num_vertices, np.int32)
    lib().or_sssp_dijkstra(g.num_vertices, g.offsets, g.nbrs, g.weights, source, d)
    return d


def sssp_delta_stepping(g: CSR, source: int, delta: int):
    d = np.empty(g.num_vertices, np.int32)
    nb = lib().or_sssp_delta_stepping(g.num_vertices, g.offsets, g.nbrs, g.weights,
                                      source, delta, d)
    return d, nb


def sssp_omp(g: CSR, source: int, delta: int, threads: int = 0):
    """OpenMP delta-stepping (CPU baseline): (dist, threads used)."""
    out = np.empty(g.num_vertices, np.int32)
    used = C.c_int32()
    lib().or_sssp_delta_stepping_omp(g.num_vertices, g.offsets, g.nbrs, g.weights, source, delta, threads, out,
                                     C.byref(used))
    return out, used.value


def cc_omp(g: CSR, threads: int = 0):
    """OpenMP label propagation with shortcutting (CPU baseline): (labels, threads used)."""
    out = np.empty(g.num_vertices, np.int32)
    used = C.c_int32()
    lib().or_cc_omp(g.num_vertices, g.offsets, g.nbrs, threads, out, C.byref(used))
    return out, used.value


def cc(g: CSR):
    lab = np.empty(g.num_vertices, np.int32)
    lib().or_cc_labels(g.num_vertices, g.offsets, g.nbrs, lab)
    return lab


def cc_label_propagation(g: CSR):
    lab = np.empty(g.num_vertices, np.int32)
    sweeps = lib().or_cc_label_propagation(g.num_vertices, g.offsets, g.nbrs, lab)
    return lab, sweeps


PR_DAMP = float(np.float32(0.85))  # apps/pr.cpp:23 declares damp as a float


def pagerank(g: CSR, iters: int = 20, damping: float = PR_DAMP):
    r = np.empty(g.num_vertices, np.float64)
    lib().or_pagerank(g.num_vertices, g.offsets, g.nbrs, iters, damping, r)
    return r


def set_threads(n: int) -> int:
    """Threads of the OpenMP oracle functions (pagerank, bc); returns the
    previous count."""
    prev = int(lib().or_max_threads())
    lib().or_set_threads(int(n))
    return prev


def bc(g: CSR, sources):
    s = np.ascontiguousarray(sources, np.int32)
    out = np.empty(g.num_vertices, np.float64)
    lib().or_bc(g.num_vertices, g.offsets, g.nbrs, s, len(s), out)
    return out


# ---------------------------------------------------------------- selector MLP
def mlp_forward(x, w1, b1, w2, b2, w3, b3):
    x = np.ascontiguousarray(x, np.float64)
    B, nin = x.shape
    nh, nout = w1.shape[1], w3.shape[1]
    h1 = np.empty((B, nh)); h2 = np.empty((B, nh)); y = np.empty((B, nout))
    args = [np.ascontiguousarray(a, np.float64) for a in (w1, b1, w2, b2, w3, b3)]
    lib().or_mlp_forward(B, nin, nh, nout, x, *args, h1, h2, y)
    return y


def mlp_loss_backward(x, target, mask, w1, b1, w2, b2, w3, b3):
    x = np.ascontiguousarray(x, np.float64)
    B, nin = x.shape
    nh, nout = w1.shape[1], w3.shape[1]
    params = [np.ascontiguousarray(a, np.float64) for a in (w1, b1, w2, b2, w3, b3)]
    grads = [np.empty_like(p) for p in params]
    loss = lib().or_mlp_loss_backward(
        B, nin, nh, nout, x, np.ascontiguousarray(target, np.float64),
        np.ascontiguousarray(mask, np.float64), *params, *grads)
    return loss, grads

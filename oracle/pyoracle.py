"""ctypes view of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product package
``artea_amd`` never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

L2SQR, DOT, COSINE, L2 = 0, 1, 2, 3
INVALID_VID = 0x7FFFFFFF
INVALID_LEVEL = 0xFFFFFFFF
NBR_NEW = 0x80000000
EARLY_STOP, L0_ONLY = 1, 2

NBR_DTYPE = np.dtype([("vid_status", "<u4"), ("dist", "<f4")])

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
u32, u64, f32, i32 = C.c_uint32, C.c_uint64, C.c_float, C.c_int


def _declare(l):
    def d(name, res, *args):
        f = getattr(l, name)
        f.restype = res
        f.argtypes = list(args)

    d("orc_distance", f32, i32, P, P, u32)
    d("orc_distance_f64", C.c_double, i32, P, P, u32)
    d("orc_distance_matrix", None, i32, P, u32, P, u32, u32, P, i32)
    d("orc_compact_graph", P, u32, u32, u32, u32, P, P, P, P, u32, i32, C.c_char_p, C.c_size_t)
    d("orc_compact_graph_ordered", P, u32, u32, u32, u32, P, P, P, P, P, u32, i32, C.c_char_p, C.c_size_t)
    d("orc_graph_from_csr", P, u32, u32, u32, u32, u32, P, P, P, u64, C.c_char_p, C.c_size_t)
    d("orc_graph_snapshot", i32, P, C.c_char_p, C.c_char_p, C.c_size_t)
    d("orc_graph_restore", P, C.c_char_p, C.c_char_p, C.c_size_t)
    d("orc_graph_free", None, P)
    for n in ("num_vertices", "top", "entry", "ul_max", "bl_max"):
        d("orc_graph_" + n, u32, P)
    d("orc_graph_highest", u32, P, u32)
    d("orc_graph_local", u32, P, u32)
    d("orc_graph_group_count", u64, P, u32)
    d("orc_graph_group_offsets", C.POINTER(u64), P, u32)
    d("orc_graph_nbrs", C.POINTER(u32), P, C.POINTER(u64))
    d("orc_graph_level_nbrs", C.POINTER(u32), P, u32, u32, C.POINTER(u32))
    d("orc_batch_query", i32, P, P, u32, i32, P, u32, u32, u32, u32, P, P, P, P, i32, C.c_char_p, C.c_size_t)
    d("orc_query_trace", i32, P, P, u32, i32, P, u32, u32, u32, P, P, P, P, P, P)
    d("orc_queue_create", P, u32)
    d("orc_queue_free", None, P)
    d("orc_queue_try_push", i32, P, u32, f32)
    d("orc_queue_pop_best", i32, P, C.POINTER(u32), C.POINTER(f32))
    d("orc_queue_should_terminate", i32, P)
    d("orc_queue_lower_bound", f32, P)
    d("orc_queue_result_size", u32, P)
    d("orc_queue_unexplored_size", u32, P)
    d("orc_queue_reset_exploration", None, P)
    d("orc_queue_extract", u32, P, u32, P, P)
    d("orc_layer_create", P, P, u32, u32, i32, u32, P, u32)
    d("orc_layer_free", None, P)
    d("orc_layer_num_rows", u32, P)
    d("orc_layer_fill_row", None, P, u32, P, u32)
    d("orc_layer_set_row_raw", None, P, u32, P, u32)
    d("orc_layer_row_size", u32, P, u32)
    d("orc_layer_get_row", u32, P, u32, P, u32)
    d("orc_layer_export", u64, P, P, P, u64)
    d("orc_next_prefill", None, P, u32, u32, u32, u32, u32, u64, u64, i32)
    d("orc_next_triangle", None, P, f32, f32, i32)
    d("orc_next_truncate", None, P, i32)
    d("orc_next_reverse", None, P, i32)
    d("orc_next_triangle_range", None, P, f32, f32, u32, u32, i32)
    d("orc_recall_matches", u64, P, u32, P, u32, u32, u32)
    d("orc_adr_at_1", C.c_double, i32, P, u32, P, u32, u32, P, u32, P, u32, P)
    d("orc_probe_knn", None, i32, P, u32, u32, P, u32, P, i32)
    d("orc_probe_farthest", None, i32, P, u32, u32, P, u32, P, i32)
    d("orc_next_reverse_range", None, P, u32, u32, i32)
    d("orc_refine_schedule", None, P, u32, u32, f32, f32, i32)
    d("orc_layer_last_logs", u64, P)
    d("orc_layer_pair_distances", u64, P)
    d("orc_rand_draw", u32, u64, u64, u32, u32, u32, u32)
    d("orc_bruteforce_knn", None, i32, P, u32, P, u32, u32, u32, P, P, i32)
    d("orc_build_full", P, P, u32, u32, i32, u32, u32, u32, u32, u32, f32, f32, f32, f32, u64, f32, i32, P)
    d("orc_build_full_routing", P, P, u32, u32, i32, u32, u32, u32, u32, u32, f32, f32, f32, f32, u64, f32, u32,
      u32, u32, i32, P)
    d("orc_next_routing", None, P, u32, u32, i32)
    d("orc_rnet_radius", f32, f32, f32, f32, u32)
    d("orc_rnet_wave", u32, u32, u32, u32)
    d("orc_rnet_build", P, P, u32, u32, i32, P, i32)
    d("orc_rnet_free", None, P)
    d("orc_rnet_max_level", u32, P)
    d("orc_rnet_top", u32, P)
    d("orc_rnet_waves", u32, P)
    d("orc_rnet_highest", C.POINTER(C.c_uint32), P)
    d("orc_rnet_export_level", u64, P, u32, P, P, u64)
    d("orc_hier_prune", u32, P, u32, i32, P, u32, u32, f32)
    d("orc_build_full_rnet", P, P, u32, u32, i32, P, u32, u32, f32, f32, u32, u32, u32, i32, P)
    d("orc_next_routing_range", None, P, u32, u32, u32, u32, i32)
    d("orc_layer_set_rows_raw", None, P, P, P)
    d("orc_next_pruning", None, P, f32, i32)
    d("orc_routing_loops", None, P, u32, u32, u32, f32, i32)
    d("orc_build_level_of", u32, u64, u32, u32, f32)
    d("orc_max_level", u32, u32)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ------------------------------------------------------------------ distance
def distance(metric: int, a, b) -> np.float32:
    a, b = _f32(a), _f32(b)
    return np.float32(lib().orc_distance(metric, _ptr(a), _ptr(b), a.shape[-1]))


def distance_f64(metric: int, a, b) -> float:
    a, b = _f32(a), _f32(b)
    return lib().orc_distance_f64(metric, _ptr(a), _ptr(b), a.shape[-1])


def distance_matrix(metric: int, a, b, nthreads: int = 0) -> np.ndarray:
    a, b = _f32(a), _f32(b)
    out = np.empty((a.shape[0], b.shape[0]), np.float32)
    lib().orc_distance_matrix(metric, _ptr(a), a.shape[0], _ptr(b), b.shape[0], a.shape[1], _ptr(out), nthreads)
    return out


# --------------------------------------------------------------------- graph
class Graph:
    """Compact hierarchical graph held by the oracle."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("null graph handle")
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_graph_free(self.h)
            self.h = None

    @classmethod
    def compact(cls, highest, levels, vecs, metric, ul_max, bl_max, order=None):
        """levels: list over level l of list over vid of neighbour-id lists;
        order: the order the vertices were assigned to their buckets."""
        n = len(highest)
        max_level = len(levels) - 1
        hl = np.ascontiguousarray(highest, dtype=np.uint32)
        offs, ids = [], []
        for l in range(max_level + 1):
            rows = levels[l]
            cnt = np.array([len(rows[v]) if v < len(rows) and rows[v] is not None else 0 for v in range(n)],
                           dtype=np.uint64)
            off = np.zeros(n + 1, np.uint64)
            np.cumsum(cnt, out=off[1:])
            flat = np.zeros(max(int(off[-1]), 1), np.uint32)
            for v in range(n):
                if cnt[v]:
                    flat[off[v]:off[v + 1]] = rows[v]
            offs.append(off)
            ids.append(flat)
        off_ptrs = (C.c_void_p * (max_level + 1))(*[o.ctypes.data for o in offs])
        id_ptrs = (C.c_void_p * (max_level + 1))(*[i.ctypes.data for i in ids])
        v = _f32(vecs)
        err = C.create_string_buffer(512)
        od = None if order is None else np.ascontiguousarray(order, np.uint32)
        h = lib().orc_compact_graph_ordered(n, max_level, ul_max, bl_max, _ptr(hl), off_ptrs, id_ptrs,
                                            _ptr(od) if od is not None else None, _ptr(v),
                                            v.shape[1] if v.ndim == 2 else 16, metric, err, 512)
        if not h:
            raise RuntimeError(err.value.decode())
        return cls(h)

    @classmethod
    def from_csr(cls, n, top, ul_max, bl_max, entry, vertex_info, group_offsets, nbrs):
        vi = np.ascontiguousarray(vertex_info, np.uint32).reshape(-1)
        gos = [np.ascontiguousarray(g, np.uint64) for g in group_offsets]
        ptrs = (C.c_void_p * len(gos))(*[g.ctypes.data for g in gos])
        nb = np.ascontiguousarray(nbrs, np.uint32)
        if nb.size == 0:
            nb = np.zeros(1, np.uint32)
            nn = 0
        else:
            nn = nb.size
        err = C.create_string_buffer(512)
        h = lib().orc_graph_from_csr(n, top, ul_max, bl_max, entry, _ptr(vi), ptrs, _ptr(nb), nn, err, 512)
        if not h:
            raise RuntimeError(err.value.decode())
        g = cls(h)
        g._keep = (vi, gos, nb)
        return g

    @classmethod
    def restore(cls, path: str):
        err = C.create_string_buffer(512)
        h = lib().orc_graph_restore(path.encode(), err, 512)
        if not h:
            raise RuntimeError(err.value.decode())
        return cls(h)

    def snapshot(self, path: str):
        err = C.create_string_buffer(512)
        if lib().orc_graph_snapshot(self.h, path.encode(), err, 512) != 0:
            raise RuntimeError(err.value.decode())

    @property
    def num_vertices(self):
        return lib().orc_graph_num_vertices(self.h)

    @property
    def top(self):
        return lib().orc_graph_top(self.h)

    @property
    def entry(self):
        return lib().orc_graph_entry(self.h)

    @property
    def ul_max(self):
        return lib().orc_graph_ul_max(self.h)

    @property
    def bl_max(self):
        return lib().orc_graph_bl_max(self.h)

    def highest(self, v):
        return lib().orc_graph_highest(self.h, v)

    def local(self, v):
        return lib().orc_graph_local(self.h, v)

    def level_nbrs(self, v, level) -> np.ndarray:
        cnt = u32()
        p = lib().orc_graph_level_nbrs(self.h, v, level, C.byref(cnt))
        return np.ctypeslib.as_array(p, shape=(cnt.value,)).copy() if cnt.value else np.zeros(0, np.uint32)

    def csr(self):
        """(highest[N], local[N], [group offsets arrays], nbrs) copies."""
        n = self.num_vertices
        hl = np.array([self.highest(v) for v in range(n)], np.uint32)
        lc = np.array([self.local(v) for v in range(n)], np.uint32)
        top = self.top if n else 0
        groups = []
        for h in range(top + 1):
            cnt = lib().orc_graph_group_count(self.h, h)
            p = lib().orc_graph_group_offsets(self.h, h)
            groups.append(np.ctypeslib.as_array(p, shape=(cnt * (h + 1) + 1,)).copy())
        nn = u64()
        p = lib().orc_graph_nbrs(self.h, C.byref(nn))
        nbrs = np.ctypeslib.as_array(p, shape=(nn.value,)).copy() if nn.value else np.zeros(0, np.uint32)
        return hl, lc, groups, nbrs


def batch_query(g: Graph, base, metric, queries, topk, queue, flags=0, thresholds=None, nthreads=0):
    base, queries = _f32(base), _f32(queries)
    nq = queries.shape[0]
    ov = np.empty((nq, topk), np.uint32)
    od = np.empty((nq, topk), np.float32)
    ndc = np.empty(nq, np.uint64)
    thr = _f32(thresholds if thresholds is not None else np.zeros(1, np.float32))
    err = C.create_string_buffer(512)
    rc = lib().orc_batch_query(g.h, _ptr(base), base.shape[1], metric, _ptr(queries), nq, topk, queue, flags,
                               _ptr(thr), _ptr(ov), _ptr(od), _ptr(ndc), nthreads, err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return ov, od, ndc


def query_trace(g: Graph, base, metric, query, topk, queue, flags=0, thresholds=None):
    base, query = _f32(base), _f32(query)
    ov = np.empty(topk, np.uint32)
    od = np.empty(topk, np.float32)
    cnt = u32()
    cur = np.full(max(g.top + 1, 1) if g.num_vertices else 1, 0xFFFFFFFF, np.uint32)
    ndc = u64()
    thr = _f32(thresholds if thresholds is not None else np.zeros(1, np.float32))
    rc = lib().orc_query_trace(g.h, _ptr(base), base.shape[1], metric, _ptr(query), topk, queue, flags, _ptr(thr),
                               _ptr(ov), _ptr(od), C.byref(cnt), _ptr(cur), C.byref(ndc))
    if rc != 0:
        raise RuntimeError("query_trace failed")
    return ov[:cnt.value], od[:cnt.value], cur, ndc.value


# --------------------------------------------------------------------- queue
class Queue:
    def __init__(self, capacity):
        self.h = lib().orc_queue_create(capacity)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_queue_free(self.h)
            self.h = None

    def try_push(self, vid, dist):
        return bool(lib().orc_queue_try_push(self.h, vid, dist))

    def pop_best(self):
        v, d = u32(), f32()
        if not lib().orc_queue_pop_best(self.h, C.byref(v), C.byref(d)):
            return None
        return v.value, d.value

    def should_terminate(self):
        return bool(lib().orc_queue_should_terminate(self.h))

    @property
    def lower_bound(self):
        return lib().orc_queue_lower_bound(self.h)

    @property
    def result_size(self):
        return lib().orc_queue_result_size(self.h)

    @property
    def unexplored_size(self):
        return lib().orc_queue_unexplored_size(self.h)

    def reset_exploration(self):
        lib().orc_queue_reset_exploration(self.h)

    def extract(self, k):
        v = np.empty(max(k, 1), np.uint32)
        d = np.empty(max(k, 1), np.float32)
        m = lib().orc_queue_extract(self.h, k, _ptr(v), _ptr(d))
        return v[:m], d[:m]


# -------------------------------------------------------------------- refine
def nbrs(pairs, new=False):
    """[(vid, dist), ...] -> structured array of Neighbor records."""
    a = np.zeros(len(pairs), NBR_DTYPE)
    for i, (v, d) in enumerate(pairs):
        a[i] = (v | (NBR_NEW if new else 0), d)
    return a


class Layer:
    def __init__(self, vecs, metric, cap, local_to_global=None):
        self.vecs = _f32(vecs)
        l2g = None if local_to_global is None else np.ascontiguousarray(local_to_global, np.uint32)
        self._l2g = l2g
        self.h = lib().orc_layer_create(_ptr(self.vecs), self.vecs.shape[0], self.vecs.shape[1], metric, cap,
                                        _ptr(l2g) if l2g is not None else None,
                                        0 if l2g is None else len(l2g))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_layer_free(self.h)
            self.h = None

    @property
    def num_rows(self):
        return lib().orc_layer_num_rows(self.h)

    def fill_row(self, local, entries):
        e = np.ascontiguousarray(entries, NBR_DTYPE)
        lib().orc_layer_fill_row(self.h, local, _ptr(e), len(e))

    def set_row_raw(self, local, entries):
        e = np.ascontiguousarray(entries, NBR_DTYPE)
        lib().orc_layer_set_row_raw(self.h, local, _ptr(e), len(e))

    def row(self, local):
        n = lib().orc_layer_row_size(self.h, local)
        out = np.zeros(max(n, 1), NBR_DTYPE)
        lib().orc_layer_get_row(self.h, local, _ptr(out), n)
        return out[:n]

    def export(self):
        n = self.num_rows
        counts = np.zeros(n, np.uint32)
        total = lib().orc_layer_export(self.h, _ptr(counts), None, 0)
        ent = np.zeros(max(total, 1), NBR_DTYPE)
        lib().orc_layer_export(self.h, _ptr(counts), _ptr(ent), total)
        return counts, ent[:total]

    def prefill(self, gen_size, start, end, row_begin=0, row_end=None, seed=0, stream=0, nthreads=0):
        lib().orc_next_prefill(self.h, gen_size, start, end, row_begin,
                               self.num_rows if row_end is None else row_end, seed, stream, nthreads)

    def triangle(self, scale, shift, nthreads=0):
        lib().orc_next_triangle(self.h, scale, shift, nthreads)

    def truncate(self, nthreads=0):
        lib().orc_next_truncate(self.h, nthreads)

    def reverse(self, nthreads=0):
        lib().orc_next_reverse(self.h, nthreads)

    def triangle_range(self, scale, shift, row_begin, row_end, nthreads=0):
        lib().orc_next_triangle_range(self.h, scale, shift, row_begin, row_end, nthreads)

    def reverse_range(self, row_begin, row_end, nthreads=0):
        lib().orc_next_reverse_range(self.h, row_begin, row_end, nthreads)

    def schedule(self, build_loops, triu_iters, scale, shift, nthreads=0):
        lib().orc_refine_schedule(self.h, build_loops, triu_iters, scale, shift, nthreads)

    def routing(self, topk, queue, nthreads=0):
        lib().orc_next_routing(self.h, topk, queue, nthreads)

    def routing_range(self, topk, queue, row_begin, row_end, nthreads=0):
        lib().orc_next_routing_range(self.h, topk, queue, row_begin, row_end, nthreads)

    def set_rows_raw(self, counts, entries):
        c = np.ascontiguousarray(counts, np.uint32)
        e = np.ascontiguousarray(entries, NBR_DTYPE)
        lib().orc_layer_set_rows_raw(self.h, _ptr(c), _ptr(e) if e.size else None)

    def pruning(self, scale, nthreads=0):
        lib().orc_next_pruning(self.h, scale, nthreads)

    def routing_loops(self, loops, topk, queue, scale, nthreads=0):
        lib().orc_routing_loops(self.h, loops, topk, queue, scale, nthreads)

    @property
    def last_logs(self):
        return lib().orc_layer_last_logs(self.h)

    @property
    def pair_distances(self):
        return lib().orc_layer_pair_distances(self.h)


def rand_draw(seed, stream, row, j, lo, hi):
    return lib().orc_rand_draw(seed, stream, row, j, lo, hi)


def bruteforce_knn(metric, base, queries, k, nthreads=0):
    base, queries = _f32(base), _f32(queries)
    nq = queries.shape[0]
    ov = np.empty((nq, k), np.uint32)
    od = np.empty((nq, k), np.float64)
    lib().orc_bruteforce_knn(metric, _ptr(base), base.shape[0], _ptr(queries), nq, base.shape[1], k, _ptr(ov),
                             _ptr(od), nthreads)
    return ov, od


def max_level(n):
    return lib().orc_max_level(n)


def build_full(vecs, metric=L2, ul_max=32, bl_max=64, max_level_=None, build_loops=15, triu_iters=4,
               prefill=0.4, scale=1.10, shift=1.5, l0_min_distance=1.0, seed=0, level_fraction=1.0, nthreads=0,
               routing_loops=0, routing_topk=64, routing_queue=96):
    """Hierarchy build: all vertices on every level (level_fraction 1) or the
    seeded geometric level assignment, then `routing_loops` routing loops per
    level; returns (Graph, seconds per level)."""
    v = _f32(vecs)
    n = v.shape[0]
    ml = max_level(n) if max_level_ is None else max_level_
    secs = np.zeros(ml + 1, np.float64)
    h = lib().orc_build_full_routing(_ptr(v), n, v.shape[1], metric, ul_max, bl_max, ml, build_loops, triu_iters,
                                     prefill, scale, shift, l0_min_distance, seed, level_fraction, routing_loops,
                                     routing_topk, routing_queue, nthreads, _ptr(secs))
    return Graph(h), secs


def build_level_of(seed: int, v: int, max_level_: int, level_fraction: float) -> int:
    return int(lib().orc_build_level_of(seed, v, max_level_, level_fraction))


def recall_matches(pred, gt, k):
    """Total RecallEstimator matches of rows pred[:, :k] against gt[:, :k]."""
    p = np.ascontiguousarray(pred, np.uint32)
    g = np.ascontiguousarray(gt, np.uint32)
    return int(lib().orc_recall_matches(_ptr(p), p.shape[1], _ptr(g), g.shape[1], p.shape[0], k))


def adr_at_1(metric, base, queries, pred, gt):
    """(ADR@1, queries counted) of the top-1 column of pred against gt's."""
    b, q = _f32(base), _f32(queries)
    p = np.ascontiguousarray(pred, np.uint32)
    g = np.ascontiguousarray(gt, np.uint32)
    c = C.c_uint64()
    a = lib().orc_adr_at_1(metric, _ptr(b), b.shape[0], _ptr(q), q.shape[0], b.shape[1], _ptr(p), p.shape[1], _ptr(g),
                           g.shape[1], C.byref(c))
    return float(a), c.value


def probe_knn(metric, base, samples, nthreads=0):
    b = _f32(base)
    s = np.ascontiguousarray(samples, np.uint32)
    out = np.zeros((len(s), 128), np.float32)
    lib().orc_probe_knn(metric, _ptr(b), b.shape[0], b.shape[1], _ptr(s), len(s), _ptr(out), nthreads)
    return out


def probe_farthest(metric, base, samples, nthreads=0):
    b = _f32(base)
    s = np.ascontiguousarray(samples, np.uint32)
    out = np.zeros(len(s), np.float32)
    lib().orc_probe_farthest(metric, _ptr(b), b.shape[0], b.shape[1], _ptr(s), len(s), _ptr(out), nthreads)
    return out


class RnetParams(C.Structure):
    """orc_rnet_params: RGraphConfig + the insertion path's RNG scale + waves."""
    _fields_ = [("beta", C.c_float), ("tau_k", C.c_float), ("l0_min_distance", C.c_float), ("scale", C.c_float),
                ("search_qs", C.c_uint32), ("ul_select_qs", C.c_uint32), ("bl_select_qs", C.c_uint32),
                ("ul_max", C.c_uint32), ("bl_max", C.c_uint32), ("insert_on_l0", C.c_uint32),
                ("wave_div", C.c_uint32), ("wave_max", C.c_uint32), ("seed", C.c_uint64)]


def rnet_params(beta=2.0, tau_k=0.0, l0_min_distance=1.0, scale=1.10, search_qs=64, ul_select_qs=100,
                bl_select_qs=100, ul_max=32, bl_max=64, insert_on_l0=False, wave_div=8, wave_max=32768, seed=0):
    return RnetParams(beta, tau_k, l0_min_distance, scale, search_qs, ul_select_qs, bl_select_qs, ul_max, bl_max,
                      int(insert_on_l0), wave_div, wave_max, seed)


class RNet:
    """The oracle's r-net insertion result: levels and rows per level."""

    def __init__(self, vecs, metric, params, nthreads=0):
        self.vecs = _f32(vecs)
        self.params = params
        self.h = lib().orc_rnet_build(_ptr(self.vecs), self.vecs.shape[0], self.vecs.shape[1], metric,
                                      C.byref(params), nthreads)
        self.n = self.vecs.shape[0]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_rnet_free(self.h)
            self.h = None

    @property
    def max_level(self):
        return lib().orc_rnet_max_level(self.h)

    @property
    def top(self):
        return lib().orc_rnet_top(self.h)

    @property
    def waves(self):
        return lib().orc_rnet_waves(self.h)

    def highest(self):
        p = lib().orc_rnet_highest(self.h)
        return np.ctypeslib.as_array(p, shape=(self.n,)).copy()

    def level_rows(self, level):
        counts = np.zeros(self.n, np.uint32)
        total = lib().orc_rnet_export_level(self.h, level, _ptr(counts), None, 0)
        ent = np.zeros(max(total, 1), NBR_DTYPE)
        lib().orc_rnet_export_level(self.h, level, _ptr(counts), _ptr(ent), total)
        return counts, ent[:total]


def rnet_radius(l0_min_distance, tau_k, beta, h):
    return float(lib().orc_rnet_radius(l0_min_distance, tau_k, beta, h))


def rnet_wave(inserted, wave_div, wave_max):
    return int(lib().orc_rnet_wave(inserted, wave_div, wave_max))


def build_full_rnet(vecs, params, metric=L2, build_loops=15, triu_iters=4, prefill=0.4, shift=1.5, routing_loops=0,
                    routing_topk=64, routing_queue=96, nthreads=0):
    """The artea_graph build: r-net insertion, refine_layer per level,
    compaction; returns (Graph, [r-net seconds, total seconds])."""
    v = _f32(vecs)
    secs = np.zeros(2, np.float64)
    h = lib().orc_build_full_rnet(_ptr(v), v.shape[0], v.shape[1], metric, C.byref(params), build_loops, triu_iters,
                                  prefill, shift, routing_loops, routing_topk, routing_queue, nthreads, _ptr(secs))
    return Graph(h), secs


def hier_prune(vecs, metric, entries, max_nbr, scale):
    v = _f32(vecs)
    e = np.ascontiguousarray(entries, NBR_DTYPE).copy()
    k = lib().orc_hier_prune(_ptr(v), v.shape[1], metric, _ptr(e), len(e), max_nbr, scale)
    return e[:k]

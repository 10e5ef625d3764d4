"""artea_amd -- MI355X (gfx950) implementation of ARTEA's search and refine path.

Python view of the C-ABI library libartea_hip.so.  The classes mirror the
reference's names (include/artea/cpu/framework/type_context/
default_context.hpp): ``DistanceMetrics``, ``DistFunc`` (dist_func_t),
``CompactHierarchicalGraph`` (compact::hierarchical_graph_t),
``HierarchicalGraphFileManager``, ``HierarchicalGraphRouter``
(hierarchical_graph_router_t) and ``PropagateEngine``.  Every compute call
runs on the GPU through the library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from ._lib import ArteaError, check, lib

__all__ = [
    "ArteaError", "DistanceMetrics", "DistFunc", "CompactHierarchicalGraph", "HierarchicalGraphFileManager",
    "DeviceIndex", "HierarchicalGraphRouter", "BatchedRouter", "LaunchGraph", "PropagateEngine", "generate", "pad_dim",
    "INVALID_VID", "EARLY_STOP", "L0_ONLY",
]

INVALID_VID = 0x7FFFFFFF
INVALID_LEVEL = 0xFFFFFFFF
EARLY_STOP, L0_ONLY = 1, 2


class DistanceMetrics:
    """DistanceMetricsT (framework/type_traits/computer_traits.hpp:38-43)."""
    EUCLIDEAN_SQR = 0
    DOT = 1
    COSINE = 2
    EUCLIDEAN = 3


def pad_dim(d: int) -> int:
    """VectorDataset::_pad_to_simd_alignment (containers/vector_dataset.hpp:242-269)."""
    return (d + 15) // 16 * 16


@dataclass(frozen=True)
class DistFunc:
    """dist_func_t<Metric, Dim>: stateless; Dim is the padded dimension."""
    metric: int
    dim: int


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dev_ptr(t) -> int:
    return int(t.data_ptr())


def generate(kind: int, seed: int, which: int, n: int, dim: int, dim_pad: Optional[int] = None, first: int = 0,
             nthreads: int = 0) -> np.ndarray:
    """Seeded synthetic workload rows (see artea_generate in include/artea_hip.h)."""
    dp = pad_dim(dim) if dim_pad is None else dim_pad
    out = np.empty((n, dp), np.float32)
    check(lib().artea_generate(kind, seed, which, first, n, dim, dp, _ptr(out), nthreads))
    return out


class CompactHierarchicalGraph:
    """Host copy of compact::HierarchicalGraph in its grouped CSR
    (index/compact_structure/hierarchical_graph.hpp:33-201)."""

    def __init__(self, num_vertices: int, top: int, ul_max: int, bl_max: int, entry: int,
                 vertex_info: np.ndarray, group_offsets: Sequence[np.ndarray], nbrs: np.ndarray):
        self.num_vertices = int(num_vertices)
        self.top = int(top)
        self.ul_max = int(ul_max)
        self.bl_max = int(bl_max)
        self.entry = int(entry)
        self.vertex_info = np.ascontiguousarray(vertex_info, np.uint32).reshape(-1, 2)
        self.group_offsets = [np.ascontiguousarray(g, np.uint64) for g in group_offsets]
        self.nbrs = np.ascontiguousarray(nbrs, np.uint32)

    # fetch_level_nbrs (:99-108)
    def fetch_level_nbrs(self, vid: int, level: int) -> np.ndarray:
        h, loc = self.vertex_info[vid]
        if h == INVALID_LEVEL or level > h:
            return np.zeros(0, np.uint32)
        r = int(loc) * (int(h) + 1) + (int(h) - level)
        o = self.group_offsets[h]
        return self.nbrs[o[r]:o[r + 1]]

    def top_occupied_level_id(self) -> int:
        return INVALID_LEVEL if self.num_vertices == 0 else self.top

    def entry_point_vid(self) -> int:
        return self.entry

    @classmethod
    def from_levels(cls, highest: Sequence[int], levels: List[Sequence[Sequence[int]]], ul_max: int, bl_max: int,
                    entry: int, top: Optional[int] = None):
        """Lay out per-(vid, level) rows in the grouped CSR, buckets in ascending vid."""
        n = len(highest)
        hl = np.asarray(highest, np.uint32)
        assigned = hl != INVALID_LEVEL
        t = int(hl[assigned].max()) if top is None and assigned.any() else (top or 0)
        vi = np.zeros((n, 2), np.uint32)
        vi[:, 0] = hl
        counts = [0] * (t + 1)
        for v in range(n):
            if hl[v] != INVALID_LEVEL:
                vi[v, 1] = counts[hl[v]]
                counts[hl[v]] += 1
            else:
                vi[v, 1] = INVALID_LEVEL
        gos, flat = [], []
        pos = 0
        members = [[] for _ in range(t + 1)]
        for v in range(n):
            if hl[v] != INVALID_LEVEL:
                members[hl[v]].append(v)
        for h in range(t + 1):
            o = np.zeros(counts[h] * (h + 1) + 1, np.uint64)
            for loc, v in enumerate(members[h]):
                for lvl in range(h, -1, -1):
                    row = list(levels[lvl][v]) if lvl < len(levels) and levels[lvl][v] is not None else []
                    o[loc * (h + 1) + (h - lvl)] = pos
                    flat.extend(row)
                    pos += len(row)
            o[-1] = pos
            gos.append(o)
        return cls(n, t, ul_max, bl_max, entry, vi, gos, np.array(flat, np.uint32))


class DeviceIndex:
    """Base vectors resident in HBM plus the device form of one attached graph."""

    def __init__(self, vecs, metric: int, device_ptr: Optional[int] = None, n: Optional[int] = None,
                 dim_pad: Optional[int] = None):
        h = C.c_void_p()
        if device_ptr is not None:
            check(lib().artea_index_create_device(C.c_void_p(device_ptr), n, dim_pad, metric, C.byref(h)))
            self.n, self.dim_pad = n, dim_pad
        else:
            v = np.ascontiguousarray(vecs, np.float32)
            check(lib().artea_index_create(_ptr(v), v.shape[0], v.shape[1], metric, C.byref(h)))
            self.n, self.dim_pad = v.shape
        self.h = h
        self.metric = metric
        self.graph = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_index_destroy(self.h)
            self.h = None

    def set_graph(self, g: CompactHierarchicalGraph):
        ptrs = (C.c_void_p * max(len(g.group_offsets), 1))(*[o.ctypes.data for o in g.group_offsets])
        nb = g.nbrs if g.nbrs.size else np.zeros(1, np.uint32)
        check(lib().artea_index_set_graph_csr(self.h, g.top, g.ul_max, g.bl_max, g.entry, _ptr(g.vertex_info), ptrs,
                                              _ptr(nb), g.nbrs.size))
        self.graph = g

    def load_hgra(self, path: str):
        check(lib().artea_index_set_graph_hgra(self.h, path.encode()))
        self.graph = path

    def snapshot_hgra(self, path: str):
        check(lib().artea_index_snapshot_hgra(self.h, path.encode()))

    def layout(self):
        """Device layout of the attached graph: ELL strides (level 0, upper
        levels) and the number of upper-level rows."""
        s0, su, ru = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().artea_index_layout(self.h, C.byref(s0), C.byref(su), C.byref(ru)))
        return {"stride0": s0.value, "stride_upper": su.value, "rows_upper": ru.value}

    def hgra_seconds(self):
        """Host-clock phases of the last device HGRA restore (file read and
        copy, segmented parse, checks and ELL fill) and snapshot (layout on
        the device, copy back and write, 0)."""
        load = (C.c_double * 3)()
        store = (C.c_double * 3)()
        check(lib().artea_index_hgra_seconds(self.h, load, store))
        return list(load), list(store)

    def compact_graph(self, highest, levels, ul_max: int, bl_max: int, order=None, metric: Optional[int] = None):
        """HierarchicalGraphCompactor::compact_graph on the device
        (compactor/hierarchical_graph_compactor.hpp:118-327): ``highest``
        per vid (INVALID_LEVEL = unassigned), ``levels[l][v]`` the rows of
        level l, ``order`` the order vertices were assigned to their buckets
        (None = ascending vid).  Attaches the compact graph to the index."""
        n = self.n
        hl = np.ascontiguousarray(highest, np.uint32)
        offs, ids = [], []
        for rows in levels:
            cnt = np.array([len(rows[v]) if v < len(rows) and rows[v] is not None else 0 for v in range(n)], np.uint64)
            off = np.zeros(n + 1, np.uint64)
            np.cumsum(cnt, out=off[1:])
            flat = np.zeros(max(int(off[-1]), 1), np.uint32)
            for v in range(n):
                if cnt[v]:
                    flat[off[v]:off[v + 1]] = rows[v]
            offs.append(off)
            ids.append(flat)
        op = (C.c_void_p * max(len(offs), 1))(*[o.ctypes.data for o in offs])
        ip = (C.c_void_p * max(len(ids), 1))(*[i.ctypes.data for i in ids])
        od = None if order is None else np.ascontiguousarray(order, np.uint32)
        check(lib().artea_index_compact_graph(self.h, len(levels) - 1, ul_max, bl_max, _ptr(hl), op, ip,
                                              _ptr(od) if od is not None else None,
                                              self.metric if metric is None else metric))

    def info(self):
        vals = [C.c_uint32() for _ in range(6)]
        check(lib().artea_index_info(self.h, *[C.byref(x) for x in vals]))
        return dict(zip(("n", "dim_pad", "top", "entry", "ul_max", "bl_max"), [x.value for x in vals]))

    def device_vectors(self) -> int:
        return lib().artea_index_device_vectors(self.h)

    @property
    def u8_rows(self) -> bool:
        """Whether the index keeps exact one-byte rows (every value a whole number in [0, 255])."""
        return bool(lib().artea_index_u8_rows(self.h))

    def set_u8_rows(self, enable: bool) -> bool:
        """Drop (False) or rebuild (True, when the values allow) the one-byte rows; returns whether they exist."""
        on = C.c_int()
        check(lib().artea_index_set_u8_rows(self.h, int(bool(enable)), C.byref(on)))
        return bool(on.value)

    def export_graph(self) -> CompactHierarchicalGraph:
        """Host copy of the attached graph in the grouped CSR."""
        g = C.c_void_p()
        check(lib().artea_index_get_graph(self.h, C.byref(g)))
        try:
            vals = [C.c_uint32() for _ in range(5)]
            nn = C.c_uint64()
            check(lib().artea_graph_info(g, *[C.byref(x) for x in vals], C.byref(nn)))
            n, top, ul, bl, entry = [x.value for x in vals]
            vi = np.ctypeslib.as_array(lib().artea_graph_vertex_info(g), (max(2 * n, 1),))[: 2 * n].copy()
            offs = []
            for h in range(top + 1 if n else 0):
                cnt = C.c_uint64()  # vertices in group h: count * (h + 1) + 1 offsets
                ptr = lib().artea_graph_group_offsets(g, h, C.byref(cnt))
                m = cnt.value * (h + 1) + 1
                offs.append(np.ctypeslib.as_array(ptr, (m,)).copy() if ptr else np.zeros(1, np.uint64))
            nb = (np.ctypeslib.as_array(lib().artea_graph_nbrs(g), (nn.value,)).copy() if nn.value
                  else np.zeros(0, np.uint32))
        finally:
            lib().artea_graph_destroy(g)
        return CompactHierarchicalGraph(n, top, ul, bl, entry, vi.reshape(-1, 2), offs, nb)


class HierarchicalGraphFileManager:
    """snapshot / restore of HGRA v1 through the device index
    (persistence/hierarchical_graph_file_manager.hpp:88-324)."""

    @staticmethod
    def snapshot(index: DeviceIndex, path: str):
        index.snapshot_hgra(path)

    @staticmethod
    def restore(index: DeviceIndex, path: str):
        index.load_hgra(path)


def _early_stop_thresholds(thresholds, top: int, metric: int):
    if thresholds is None:
        return None
    t = np.zeros(max(top + 1, 1), np.float32)
    t[: len(thresholds)] = np.asarray(thresholds, np.float32)[: len(t)]
    return t


class HierarchicalGraphRouter:
    """hierarchical_graph_router_t<Metric, Dim>
    (router/hierarchical_graph_router.hpp:93-430) on the GPU.

    ``index`` carries base_vecs (borrowed, as the reference borrows them) and
    the attached compact graph.  Results are (vids[nq, topk], dists[nq,
    topk]) ascending by (distance, vid), padded with (INVALID_VID, FLT_MAX).
    """

    def __init__(self, index: DeviceIndex, dist_func: Optional[DistFunc], topk: int, candidate_queue_size: int = 16):
        if dist_func is not None and (dist_func.metric != index.metric or dist_func.dim != index.dim_pad):
            raise ValueError("dist_func does not match the index metric / padded dimension")
        self.index = index
        self.topk = topk
        self.queue = candidate_queue_size
        h = C.c_void_p()
        check(lib().artea_router_create(index.h, topk, candidate_queue_size, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_router_destroy(self.h)
            self.h = None

    def initialize(self):
        """No per-thread pools on the device; kept for API parity."""

    def batch_query(self, queries, enable_early_stop: bool = False, thresholds=None, l0_only: bool = False):
        q = np.ascontiguousarray(queries, np.float32)
        nq = q.shape[0]
        ov = np.empty((nq, self.topk), np.uint32)
        od = np.empty((nq, self.topk), np.float32)
        flags = (EARLY_STOP if enable_early_stop else 0) | (L0_ONLY if l0_only else 0)
        top = self.index.info()["top"]
        thr = _early_stop_thresholds(thresholds, 0 if top == INVALID_LEVEL else top, self.index.metric)
        check(lib().artea_router_batch_query(self.h, _ptr(q), nq, flags, _ptr(thr) if thr is not None else None,
                                             _ptr(ov), _ptr(od)))
        return ov, od

    def query(self, q, **kw):
        v, d = self.batch_query(np.asarray(q, np.float32)[None, :], **kw)
        m = int(np.sum(v[0] != INVALID_VID))
        return v[0, :m], d[0, :m]

    def batch_query_l0_only(self, queries):
        return self.batch_query(queries, l0_only=True)

    def query_l0_only(self, q):
        return self.query(q, l0_only=True)

    def batch_query_device(self, d_queries: int, nq: int, d_out_vids: int, d_out_dists: int, stream: int = 0,
                           flags: int = 0, thresholds=None):
        thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.float32)
        check(lib().artea_router_batch_query_device(self.h, C.c_void_p(d_queries), nq, flags,
                                                    _ptr(thr) if thr is not None else None, C.c_void_p(d_out_vids),
                                                    C.c_void_p(d_out_dists), C.c_void_p(stream)))

    def last_stats(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().artea_router_last_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"ndc": a.value, "expansions": b.value, "tail_overflows": c.value}

    def last_overflows(self):
        """(count, batch indices) of the last launch's queries whose candidate
        list overflowed its tie tail (at most 255 indices listed)."""
        return _overflows(lib().artea_router_last_overflows, self.h)


def _overflows(fn, h):
    idx = np.zeros(255, np.uint32)
    cnt = C.c_uint32()
    check(fn(h, 255, C.c_void_p(idx.ctypes.data), C.byref(cnt)))
    return cnt.value, idx[:min(cnt.value, 255)].copy()


class BatchedRouter:
    """Persistent fused router with per-query topk / queue size."""

    def __init__(self, index: DeviceIndex, max_topk: int, max_queue: int):
        self.index = index
        self.max_topk = max_topk
        self.max_queue = max_queue
        h = C.c_void_p()
        check(lib().artea_batched_router_create(index.h, max_topk, max_queue, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_batched_router_destroy(self.h)
            self.h = None

    def query_device(self, d_queries: int, nq: int, d_topk: int, d_queue: int, d_out_vids: int, d_out_dists: int,
                     stream: int = 0, flags: int = 0, thresholds=None):
        thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.float32)
        check(lib().artea_batched_router_query_device(self.h, C.c_void_p(d_queries), nq, C.c_void_p(d_topk),
                                                      C.c_void_p(d_queue), flags,
                                                      _ptr(thr) if thr is not None else None,
                                                      C.c_void_p(d_out_vids), C.c_void_p(d_out_dists),
                                                      C.c_void_p(stream)))

    def last_stats(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().artea_batched_router_last_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"ndc": a.value, "expansions": b.value, "tail_overflows": c.value}

    def last_overflows(self):
        return _overflows(lib().artea_batched_router_last_overflows, self.h)

    # Kernel variant names by slot: FP32 rows (0..3), and one-byte rows
    # (4..7, while the index keeps them: every value a whole number in [0, 255]).
    VARIANTS = ("narrow", "wide", "wide-3waves", "wide-4waves")
    VARIANTS_U8 = ("u8-64rows", "u8-32rows", "u8-16rows", "u8-64rows-4waves")

    @property
    def variant_names(self):
        return self.VARIANTS + self.VARIANTS_U8 if self.index.u8_rows else self.VARIANTS

    def autotune(self, d_queries: int, nq: int, d_topk: int, d_queue: int, d_out_vids: int, d_out_dists: int,
                 flags: int = 0, thresholds=None):
        """Times every kernel variant on this batch and keeps the fastest;
        returns (chosen variant, per-variant milliseconds)."""
        thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.float32)
        ms = np.zeros(len(self.variant_names), np.float32)
        ch = C.c_int(-1)
        check(lib().artea_batched_router_autotune(self.h, C.c_void_p(d_queries), nq, C.c_void_p(d_topk),
                                                  C.c_void_p(d_queue), flags, _ptr(thr) if thr is not None else None,
                                                  C.c_void_p(d_out_vids), C.c_void_p(d_out_dists), _ptr(ms),
                                                  C.byref(ch)))
        return ch.value, ms

    def set_variant(self, variant: int):
        check(lib().artea_batched_router_set_variant(self.h, variant))

    def variant(self) -> int:
        v = C.c_int(-1)
        check(lib().artea_batched_router_get_variant(self.h, C.byref(v)))
        return v.value

    def capture(self, d_queries: int, nq: int, d_topk: int, d_queue: int, d_out_vids: int, d_out_dists: int,
                flags: int = 0, thresholds=None) -> "LaunchGraph":
        """Records one query over these device buffers as a HIP graph."""
        thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.float32)
        h = C.c_void_p()
        check(lib().artea_batched_router_capture(self.h, C.c_void_p(d_queries), nq, C.c_void_p(d_topk),
                                                 C.c_void_p(d_queue), flags, _ptr(thr) if thr is not None else None,
                                                 C.c_void_p(d_out_vids), C.c_void_p(d_out_dists), C.byref(h)))
        return LaunchGraph(h, self)


class LaunchGraph:
    """An instantiated HIP graph of one batched-router query."""

    def __init__(self, h, owner):
        self.h = h
        self.owner = owner  # keeps the router (and its buffers) alive

    def launch(self, stream: int = 0):
        check(lib().artea_launch_graph_launch(self.h, C.c_void_p(stream)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_launch_graph_destroy(self.h)
            self.h = None


from .refine import PropagateEngine, build_full  # noqa: E402

"""PropagateEngine on the GPU (refiner/propagate_engine.hpp) and the
full-hierarchy build, through libartea_hip.so."""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np

from ._lib import check, lib

NBR_NEW = 0x80000000
JOIN_FP32, JOIN_BF16 = 0, 1


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class PropagateEngine:
    """One single-layer graph plus the propagate engine over it.

    ``rows`` are Neighbor records as a (count, 2) uint32 view: vid:31|new:1
    and the FP32 distance bits.  Every ``next_*`` call is propagate followed
    by merge_logs (propagate_engine.hpp:200-225)."""

    def __init__(self, index, cap: int, metric: Optional[int] = None, local_to_global: Optional[Sequence[int]] = None,
                 seed: int = 0, join: int = JOIN_FP32):
        m = index.metric if metric is None else metric
        h = C.c_void_p()
        l2g = None if local_to_global is None else np.ascontiguousarray(local_to_global, np.uint32)
        self._l2g = l2g
        self.n_local = index.n if l2g is None else len(l2g)
        self.cap = cap
        check(lib().artea_refine_create(C.c_void_p(index.device_vectors()), index.n, index.dim_pad, m, cap,
                                        _ptr(l2g) if l2g is not None else None, self.n_local, seed, C.byref(h)))
        self.h = h
        self.index = index
        self.set_join(join)

    def set_join(self, precision: int):
        """JOIN_FP32: every needed pair by the FP32 chain; JOIN_BF16: MFMA
        Gram tiles with FP32 rechecks (identical rows and logs)."""
        check(lib().artea_refine_set_join(self.h, precision))

    def set_recheck_queue(self, entries: int):
        """Per-row FP32 recheck queue of the BF16 join (0..1024)."""
        check(lib().artea_refine_set_recheck_queue(self.h, entries))

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_refine_destroy(self.h)
            self.h = None

    def upload(self, counts, entries):
        c = np.ascontiguousarray(counts, np.uint32)
        e = np.ascontiguousarray(entries).view(np.uint32).reshape(-1)
        if e.size == 0:
            e = np.zeros(2, np.uint32)
        check(lib().artea_refine_upload_rows(self.h, _ptr(c), _ptr(e)))

    def download(self):
        c = np.zeros(self.n_local, np.uint32)
        e = np.zeros(max(self.n_local * self.cap, 1) * 2, np.uint32)
        check(lib().artea_refine_download_rows(self.h, _ptr(c), _ptr(e)))
        total = int(c.sum())
        return c, e[: 2 * total].reshape(-1, 2)

    def next_prefill(self, gen_size, start, end, row_begin=0, row_end=None, stream=0):
        check(lib().artea_refine_prefill(self.h, gen_size, start, end, row_begin,
                                         self.n_local if row_end is None else row_end, stream))

    def next_triangle(self, scale_coeffs, shifted_coeffs):
        check(lib().artea_refine_triangle(self.h, scale_coeffs, shifted_coeffs))

    def next_triangle_range(self, scale_coeffs, shifted_coeffs, row_begin, row_end):
        check(lib().artea_refine_triangle_range(self.h, scale_coeffs, shifted_coeffs, row_begin, row_end))

    def next_reverse_range(self, row_begin, row_end):
        check(lib().artea_refine_reverse_range(self.h, row_begin, row_end))

    def next_truncate(self):
        check(lib().artea_refine_truncate(self.h))

    def next_reverse(self):
        check(lib().artea_refine_reverse(self.h))

    def schedule(self, build_loops, triu_iters, scale_coeffs, effective_shift):
        check(lib().artea_refine_schedule(self.h, build_loops, triu_iters, scale_coeffs, effective_shift))

    def routing_loops(self, loops, topk, queue, scale_coeffs):
        """refine_layer's routing loops: routing + pruning, reverse + arc
        between loops (index/artea_graph/index_factory.hpp:400-404)."""
        check(lib().artea_refine_routing_loops(self.h, loops, topk, queue, scale_coeffs))

    def next_pruning(self, scale_coeffs):
        """next(PruningUpdater): rows pruned in place, no logs."""
        check(lib().artea_refine_prune(self.h, scale_coeffs))

    def routing_stats(self):
        a, b, d = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().artea_refine_routing_stats(self.h, C.byref(a), C.byref(b), C.byref(d)))
        return {"distances": a.value, "expansions": b.value, "unresolved": d.value}

    def last_stats(self):
        a, b, d = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().artea_refine_last_stats(self.h, C.byref(a), C.byref(b), C.byref(d)))
        return {"logs": a.value, "pair_distances": b.value, "approx_pair_distances": d.value}


def max_level_for(n: int) -> int:
    """compute_max_allowed_level_id (index/stacked_rgraph/configs.hpp:241-249)."""
    if n == 0:
        return 1
    ratio = n / 1000.0
    if ratio <= 1.0:
        return 1
    return max(int(math.ceil(math.log(ratio))), 1)


def build_full(index, ul_max=32, bl_max=64, max_level=None, build_loops=15, triu_iters=4, prefill=0.4, scale=1.10,
               shift=1.5, l0_min_distance=1.0, seed=0, join=JOIN_FP32, level_fraction=1.0, routing_loops=0,
               routing_topk=64, routing_queue=96):
    """Hierarchy build on the GPU; attaches the graph to ``index``.
    ``join`` picks the triangle join's pair-distance path (the graph is the
    same either way); ``level_fraction`` < 1 draws each vertex's level
    geometrically (sparse upper layers) instead of putting every vertex on
    every level; ``routing_loops`` routing loops (routing topk and queue)
    follow each level's build loops.  Returns seconds per level."""
    ml = max_level_for(index.n) if max_level is None else max_level
    secs = np.zeros(ml + 1, np.float64)
    check(lib().artea_build_full_routing(index.h, ul_max, bl_max, ml, build_loops, triu_iters, prefill, scale, shift,
                                         l0_min_distance, seed, join, level_fraction, routing_loops, routing_topk,
                                         routing_queue, _ptr(secs)))
    return secs


class RnetParams(C.Structure):
    """artea_rnet_params: RGraphConfig (index/stacked_rgraph/configs.hpp
    :65-138), the insertion path's RNG scale and the wave schedule."""
    _fields_ = [("rnet_beta", C.c_float), ("tau_k", C.c_float), ("l0_min_distance", C.c_float),
                ("scale_coeffs", C.c_float), ("search_nn_qs", C.c_uint32), ("ul_select_nbrs_qs", C.c_uint32),
                ("bl_select_nbrs_qs", C.c_uint32), ("ul_max_nbr_size", C.c_uint32), ("bl_max_nbr_size", C.c_uint32),
                ("insert_on_l0", C.c_uint32), ("wave_div", C.c_uint32), ("wave_max", C.c_uint32),
                ("seed", C.c_uint64)]


def rnet_params(rnet_beta=2.0, tau_k=0.0, l0_min_distance=1.0, scale_coeffs=1.10, search_nn_qs=64,
                ul_select_nbrs_qs=100, bl_select_nbrs_qs=100, ul_max_nbr_size=32, bl_max_nbr_size=64,
                insert_on_l0=False, wave_div=8, wave_max=32768, seed=0) -> RnetParams:
    return RnetParams(rnet_beta, tau_k, l0_min_distance, scale_coeffs, search_nn_qs, ul_select_nbrs_qs,
                      bl_select_nbrs_qs, ul_max_nbr_size, bl_max_nbr_size, int(insert_on_l0), wave_div, wave_max,
                      seed)


class RNet:
    """R-net insertion on the GPU (stacked_rgraph::IndexFactory): levels and
    rows per level of the dynamic hierarchy, before refinement."""

    def __init__(self, index, params: RnetParams, metric: Optional[int] = None):
        m = (3 if index.metric == 0 else index.metric) if metric is None else metric
        h = C.c_void_p()
        check(lib().artea_rnet_build(index.h, m, C.byref(params), C.byref(h)))
        self.h = h
        self.n = index.n
        self.params = params

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_rnet_destroy(self.h)
            self.h = None

    def info(self):
        ml, top, waves = C.c_uint32(), C.c_uint32(), C.c_uint32()
        dist, unres = C.c_uint64(), C.c_uint64()
        check(lib().artea_rnet_info(self.h, C.byref(ml), C.byref(top), C.byref(waves), C.byref(dist), C.byref(unres)))
        return {"max_level": ml.value, "top": top.value, "waves": waves.value, "distances": dist.value,
                "unresolved": unres.value}

    def highest(self):
        out = np.zeros(max(self.n, 1), np.uint32)
        check(lib().artea_rnet_levels(self.h, _ptr(out)))
        return out[: self.n]

    def level_rows(self, level):
        counts = np.zeros(max(self.n, 1), np.uint32)
        total = C.c_uint64()
        check(lib().artea_rnet_level_rows(self.h, level, _ptr(counts), None, 0, C.byref(total)))
        ent = np.zeros(max(total.value, 1) * 2, np.uint32)
        check(lib().artea_rnet_level_rows(self.h, level, _ptr(counts), _ptr(ent), total.value, C.byref(total)))
        return counts[: self.n], ent[: 2 * total.value].reshape(-1, 2)


def build_full_rnet(index, params: RnetParams, build_loops=15, triu_iters=4, prefill=0.4, shift=1.5, join=JOIN_FP32,
                    routing_loops=0, routing_topk=64, routing_queue=96):
    """The artea_graph build on the GPU (index/artea_graph/index_factory.hpp
    :135-410): r-net insertion, refine_layer per level, compaction; attaches
    the graph to ``index``.  Returns (r-net seconds, total seconds)."""
    secs = np.zeros(2, np.float64)
    check(lib().artea_build_full_rnet(index.h, C.byref(params), build_loops, triu_iters, prefill, shift, join,
                                      routing_loops, routing_topk, routing_queue, _ptr(secs)))
    return float(secs[0]), float(secs[1])

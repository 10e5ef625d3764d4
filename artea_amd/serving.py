"""Continuous-batching request server over a resident index.

`Server` wraps the library's native server (artea_server_* in
include/artea_hip.h): requests (a query row, topk, queue size) are submitted
as they arrive and completed ones are collected by polling.  The native side
keeps one queue per queue-size class and runs `lanes` streams, each holding
at most `budget` requests in flight; an idle lane takes the class of the
oldest waiting request and launches up to `budget` of that class's requests
with its batched router, so short searches never wait behind long ones in
the same launch.

Per-request results do not depend on what they are batched with: every query
is routed by its own wavefront with no shared state besides the work counter.
`Server.run` replays an arrival schedule against the server and measures each
request's latency from its arrival to the poll that returned it.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._lib import check, lib

# Queue-size classes of the router's register layouts (candidate list of K
# blocks of 64 entries, 16 entries of tie headroom).
QUEUE_CLASSES = (48, 112, 240, 496, 816, 1008)


def queue_class(q: int) -> int:
    for c in QUEUE_CLASSES:
        if q <= c:
            return c
    raise ValueError("queue size %d exceeds the device maximum" % q)


_CLASSES = np.asarray(QUEUE_CLASSES, np.int64)


def queue_classes(q: np.ndarray) -> np.ndarray:
    """queue_class over an array of queue sizes."""
    q = np.asarray(q, np.int64)
    if q.size and int(q.max()) > QUEUE_CLASSES[-1]:
        raise ValueError("queue size %d exceeds the device maximum" % int(q.max()))
    return _CLASSES[np.searchsorted(_CLASSES, q, side="left")]


class Server:
    """Open-loop request server.  ``budget`` bounds the requests one lane
    holds in flight; ``lanes`` bounds how many launches overlap.  With
    ``hbm_bytes`` the admission comes from memory instead
    (artea_server_create_budget): the lanes' device buffers share
    ``hbm_bytes``, and a launch of a queue-size class takes at most
    ``lds_rounds`` times the wavefronts that class's kernel keeps resident."""

    def __init__(self, index, max_topk: int = 100, max_queue: int = 800, lanes: int = 4, budget: int = 2048,
                 hbm_bytes: int = 0, lds_rounds: int = 4):
        h = C.c_void_p()
        if hbm_bytes:
            check(lib().artea_server_create_budget(index.h, max_topk, max_queue, lanes, hbm_bytes, lds_rounds,
                                                   C.byref(h)))
        else:
            check(lib().artea_server_create(index.h, max_topk, max_queue, lanes, budget, C.byref(h)))
        self.h = h
        self.index = index  # the server reads the index's device rows and graph
        self.dim_pad = index.dim_pad
        self.max_topk = max_topk
        self.lanes = lanes
        if hbm_bytes:
            budget = self.admission()["lane_requests"]
        self.budget = budget
        self._cap = max(lanes * budget, 1)
        self._t = np.empty(self._cap, np.uint64)
        self._v = np.empty((self._cap, max_topk), np.uint32)
        self._d = np.empty((self._cap, max_topk), np.float32)

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_server_destroy(self.h)
            self.h = None

    def submit(self, queries: np.ndarray, topk: np.ndarray, queue: np.ndarray) -> int:
        """Queues requests (rows of dim_pad floats); returns the first ticket
        (the others follow consecutively)."""
        q = np.ascontiguousarray(queries, np.float32)
        k = np.ascontiguousarray(topk, np.uint32)
        L = np.ascontiguousarray(queue, np.uint32)
        n = q.shape[0]
        if q.ndim != 2 or q.shape[1] != self.dim_pad or k.shape != (n,) or L.shape != (n,):
            raise ValueError("submit: queries must be n x %d with n topk and n queue values" % self.dim_pad)
        first = C.c_uint64()
        check(lib().artea_server_submit(self.h, C.c_void_p(q.ctypes.data), n, C.c_void_p(k.ctypes.data),
                                        C.c_void_p(L.ctypes.data), C.byref(first)))
        return first.value

    def poll(self):
        """Completed requests since the last poll: (tickets, vids, dists) views
        valid until the next poll."""
        got = C.c_uint32()
        check(lib().artea_server_poll(self.h, self._cap, C.c_void_p(self._t.ctypes.data),
                                      C.c_void_p(self._v.ctypes.data), C.c_void_p(self._d.ctypes.data),
                                      C.byref(got)))
        m = got.value
        return self._t[:m], self._v[:m], self._d[:m]

    def _poll_into(self, base: int, table_v: np.ndarray, table_d: np.ndarray) -> int:
        """Completed requests written to rows (ticket - base) of the tables;
        their tickets land in self._t[:count]."""
        got = C.c_uint32()
        check(lib().artea_server_poll_into(self.h, base, table_v.shape[0], C.c_void_p(table_v.ctypes.data),
                                           C.c_void_p(table_d.ctypes.data), self._cap,
                                           C.c_void_p(self._t.ctypes.data), C.byref(got)))
        return got.value

    def admission(self):
        """Requests per launch and resident wavefronts of each queue-size
        class, the requests one lane holds, and the lanes' device bytes."""
        admit = np.zeros(6, np.uint32)
        resident = np.zeros(6, np.uint32)
        lane, hbm = C.c_uint32(), C.c_uint64()
        check(lib().artea_server_admission(self.h, C.c_void_p(admit.ctypes.data), C.c_void_p(resident.ctypes.data),
                                           C.byref(lane), C.byref(hbm)))
        return {"classes": _CLASSES.tolist(), "admit": admit.tolist(), "resident_waves": resident.tolist(),
                "lane_requests": lane.value, "hbm_bytes": hbm.value}

    def status(self):
        """(queued, in flight, completed but not yet polled) request counts."""
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().artea_server_status(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def overflows(self):
        """(re-run, unresolved) counts of requests whose candidate list
        overflowed its tie tail (re-run on a larger list before returning)."""
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().artea_server_overflows(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def run(self, queries: np.ndarray, topk: np.ndarray, queue: np.ndarray, arrivals: np.ndarray):
        """Serve requests arriving at `arrivals` (seconds from now, sorted).
        Arrived requests are submitted, at most one lane budget per call,
        while fewer than two launches' worth per lane are inside the server
        (queued or in flight); the rest wait on the host, and their wait
        counts in their latency.
        Returns (vids, dists, latency_seconds, wall_seconds)."""
        n = len(arrivals)
        queries = np.ascontiguousarray(queries, np.float32)
        topk = np.ascontiguousarray(topk, np.uint32)
        queue = np.ascontiguousarray(queue, np.uint32)
        out_v = np.full((n, self.max_topk), 0x7FFFFFFF, np.uint32)
        out_d = np.full((n, self.max_topk), np.finfo(np.float32).max, np.float32)
        lat = np.zeros(n, np.float64)
        window = 2 * self.lanes * self.budget
        base = 0
        nxt = 0
        done = 0
        t0 = time.perf_counter()
        while done < n:
            now = time.perf_counter() - t0
            if nxt < n and arrivals[nxt] <= now and nxt - done < window:
                # at most one lane budget per submit, so a backlog reaches
                # the lanes in launch-sized pieces
                end = min(int(np.searchsorted(arrivals, now, side="right")), done + window, nxt + self.budget)
                first = self.submit(queries[nxt:end], topk[nxt:end], queue[nxt:end])
                if nxt == 0:
                    base = first  # request i has ticket base + i
                nxt = end
            m = self._poll_into(base, out_v, out_d) if nxt else 0
            if m:
                t = time.perf_counter() - t0
                ids = (self._t[:m] - np.uint64(base)).astype(np.int64)
                lat[ids] = t - arrivals[ids]
                done += m
            elif nxt < n and done == nxt:
                # nothing in flight: sleep towards the next arrival
                wait = arrivals[nxt] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
        return out_v, out_d, lat, time.perf_counter() - t0


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def percentile_nearest_rank(x: np.ndarray, p: float) -> float:
    """Nearest-rank percentile (router/profiler/profile_1nn_result.hpp:62-69)."""
    s = np.sort(x)
    if len(s) == 0:
        return float("nan")
    r = int(np.ceil(p / 100.0 * len(s)))
    return float(s[max(r, 1) - 1])

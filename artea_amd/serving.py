"""Continuous-batching request loop over a resident index.

Requests (a query row, topk, queue size) arrive over time.  The server keeps
the index and every request's device state resident, and runs `lanes`
independent streams; whenever a lane is free it admits the requests that have
arrived since, up to the lane's budget, and launches them.  Work of different
lanes overlaps on the device, so requests that arrive while others are on the
GPU start as soon as a lane frees up.  Each admitted batch is split by queue
size class so each class runs the register layout sized for it.

Per-request results do not depend on what they are batched with: every query
is routed by its own wavefront with no shared state besides the work counter.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import BatchedRouter
from ._lib import check, lib
from .device import DeviceBuffer, Event, Stream

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


class PinnedBuffer:
    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape)
        n = int(np.prod(shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().artea_host_alloc(max(n, 1), C.byref(p)))
        self.ptr = p.value
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(n, 1)).from_address(self.ptr))[:n].view(self.dtype)
        self.array = self.array.reshape(self.shape)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().artea_host_free(C.c_void_p(self.ptr))
            self.ptr = None


@dataclass
class _Launch:
    ids: object       # request ids of one queue-size class segment (index array, or a slice when consecutive)
    count: int        # requests in the segment
    rows: int         # first row of the segment in the lane's buffers
    event: Event      # recorded after the segment's result copies


class _Lane:
    def __init__(self, index, max_topk: int, budget: int, classes):
        self.stream = Stream()
        self.routers = {c: BatchedRouter(index, max_topk, c) for c in classes}
        d = index.dim_pad
        self.budget = budget
        self.hq = PinnedBuffer((budget, d), np.float32)
        self.hk = PinnedBuffer((budget,), np.uint32)
        self.hl = PinnedBuffer((budget,), np.uint32)
        self.hv = PinnedBuffer((budget, max_topk), np.uint32)
        self.hd = PinnedBuffer((budget, max_topk), np.float32)
        self.dq = DeviceBuffer.empty((budget, d), np.float32)
        self.dk = DeviceBuffer.empty((budget,), np.uint32)
        self.dl = DeviceBuffer.empty((budget,), np.uint32)
        self.dv = DeviceBuffer.empty((budget, max_topk), np.uint32)
        self.dd = DeviceBuffer.empty((budget, max_topk), np.float32)
        self.max_topk = max_topk
        self.inflight: List[_Launch] = []

    def classes(self, q: np.ndarray) -> np.ndarray:
        # smallest router class that holds the queue and the row stride
        return queue_classes(np.maximum(q, self.max_topk))

    def launch(self, lo: int, hi: int, queries, topk, queue):
        """Launches requests lo..hi-1, grouped by queue-size class."""
        m = hi - lo
        cl = self.classes(queue[lo:hi])
        if m < 2 or bool(np.all(cl[1:] >= cl[:-1])):
            # already grouped: every segment is a run of consecutive requests
            ids, classes = None, cl
            self.hq.array[:m] = queries[lo:hi]
            self.hk.array[:m] = topk[lo:hi]
            self.hl.array[:m] = queue[lo:hi]
        else:
            order = np.argsort(cl, kind="stable")
            ids = lo + order
            classes = cl[order]
            self.hq.array[:m] = queries[ids]
            self.hk.array[:m] = topk[ids]
            self.hl.array[:m] = queue[ids]
        s = self.stream.h
        d = self.hq.array.shape[1]
        check(lib().artea_memcpy_htod(C.c_void_p(self.dq.ptr), C.c_void_p(self.hq.ptr), m * d * 4, C.c_void_p(s)))
        check(lib().artea_memcpy_htod(C.c_void_p(self.dk.ptr), C.c_void_p(self.hk.ptr), m * 4, C.c_void_p(s)))
        check(lib().artea_memcpy_htod(C.c_void_p(self.dl.ptr), C.c_void_p(self.hl.ptr), m * 4, C.c_void_p(s)))
        # Smallest classes first, each with its own result copies and event,
        # so a request completes when its class does, not when the lane's
        # slowest class does.
        start = 0
        while start < m:
            c = classes[start]
            end = start + int(np.searchsorted(classes[start:], c, side="right"))
            self.routers[c].query_device(self.dq.ptr + start * d * 4, end - start, self.dk.ptr + start * 4,
                                         self.dl.ptr + start * 4, self.dv.ptr + start * self.max_topk * 4,
                                         self.dd.ptr + start * self.max_topk * 4, stream=s)
            off, nb = start * self.max_topk * 4, (end - start) * self.max_topk * 4
            check(lib().artea_memcpy_dtoh(C.c_void_p(self.hv.ptr + off), C.c_void_p(self.dv.ptr + off), nb,
                                          C.c_void_p(s)))
            check(lib().artea_memcpy_dtoh(C.c_void_p(self.hd.ptr + off), C.c_void_p(self.dd.ptr + off), nb,
                                          C.c_void_p(s)))
            ev = Event()
            ev.record(self.stream)
            seg = slice(lo + start, lo + end) if ids is None else ids[start:end]
            self.inflight.append(_Launch(seg, end - start, start, ev))
            start = end

    def poll(self) -> List[_Launch]:
        """Segments completed since the last poll, in launch order."""
        out = []
        while self.inflight:
            done = C.c_int()
            check(lib().artea_event_query(C.c_void_p(self.inflight[0].event.h), C.byref(done)))
            if not done.value:
                break
            out.append(self.inflight.pop(0))
        return out


class Server:
    """Open-loop request server.  ``budget`` bounds the queries one lane
    holds in flight (its device buffers are sized for it); ``lanes`` bounds
    how many launches overlap."""

    def __init__(self, index, max_topk: int = 100, max_queue: int = 800, lanes: int = 4, budget: int = 2048):
        top = queue_class(max(max_queue, max_topk))
        classes = [c for c in QUEUE_CLASSES if queue_class(max_topk) <= c <= top]
        self.lanes = [_Lane(index, max_topk, budget, classes) for _ in range(lanes)]
        self.max_topk = max_topk
        self.budget = budget

    def run(self, queries: np.ndarray, topk: np.ndarray, queue: np.ndarray, arrivals: np.ndarray):
        """Serve requests arriving at `arrivals` (seconds from now, sorted).
        Returns (vids, dists, latency_seconds, wall_seconds)."""
        n = len(arrivals)
        queries = np.ascontiguousarray(queries, np.float32)
        topk = np.asarray(topk, np.uint32)
        queue = np.asarray(queue, np.uint32)
        out_v = np.full((n, self.max_topk), 0x7FFFFFFF, np.uint32)
        out_d = np.full((n, self.max_topk), np.finfo(np.float32).max, np.float32)
        lat = np.zeros(n, np.float64)
        nxt = 0
        done = 0
        t0 = time.perf_counter()
        while done < n:
            now = time.perf_counter() - t0
            progressed = False
            for ln in self.lanes:
                for seg in ln.poll():
                    t = time.perf_counter() - t0
                    ids, r0, m = seg.ids, seg.rows, seg.count
                    out_v[ids] = ln.hv.array[r0:r0 + m]
                    out_d[ids] = ln.hd.array[r0:r0 + m]
                    lat[ids] = t - arrivals[ids]
                    done += m
                    progressed = True
                if not ln.inflight and nxt < n and arrivals[nxt] <= now:
                    end = min(n, int(np.searchsorted(arrivals, now, side="right")), nxt + ln.budget)
                    ln.launch(nxt, end, queries, topk, queue)
                    nxt = end
                    progressed = True
            if not progressed:
                if nxt < n and all(not ln.inflight for ln in self.lanes):
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

"""artea_batched_query_noalloc (the allocation-free form of
artea_batched_router_query_device) returns that entry point's results:
mixed topk / queue, early stop, two calls in a row with the second on a
smaller batch, and its argument errors."""
import ctypes as C

import numpy as np
import pytest

import artea_amd as aa
from artea_amd._lib import lib

pytestmark = pytest.mark.gpu


def noalloc(idx, max_topk, max_queue, tq, nq, tk, tl, ov, od, stream, flags=0, thresholds=None):
    thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.float32)
    return lib().artea_batched_query_noalloc(idx.h, max_topk, max_queue, C.c_void_p(tq.ptr), nq, C.c_void_p(tk.ptr),
                                             C.c_void_p(tl.ptr), flags,
                                             thr.ctypes.data_as(C.c_void_p) if thr is not None else None,
                                             C.c_void_p(ov.ptr), C.c_void_p(od.ptr), C.c_void_p(stream))


def batch(kind, seed, nq, dim, max_topk, queues):
    qs = aa.generate(kind, seed, 1, nq, dim)
    rng = np.random.default_rng(seed)
    topk = rng.choice(np.array([1, 10, max_topk], np.uint32), nq)
    queue = np.maximum(rng.choice(np.array(queues, np.uint32), nq), topk).astype(np.uint32)
    return qs, topk, queue


@pytest.mark.parametrize("kind,dim,metric", [(0, 128, 0), (2, 96, 3)])
def test_noalloc_equals_batched_router_twice(kind, dim, metric):
    from artea_amd.device import DeviceBuffer, Stream

    vecs = aa.generate(kind, 51, 0, 4000, dim)
    idx = aa.DeviceIndex(vecs, metric)
    aa.build_full(idx, max_level=2, build_loops=3, triu_iters=2, seed=5, level_fraction=0.25)
    s = Stream()
    br = aa.BatchedRouter(idx, 100, 400)
    # first call on 500 queries, the second on 137 other ones
    for seed, nq in ((61, 500), (62, 137)):
        qs, topk, queue = batch(kind, seed, nq, dim, 100, [100, 200, 400])
        tq, tk, tl = DeviceBuffer.from_numpy(qs), DeviceBuffer.from_numpy(topk), DeviceBuffer.from_numpy(queue)
        want_v, want_d = DeviceBuffer.empty((nq, 100), np.uint32), DeviceBuffer.empty((nq, 100), np.float32)
        got_v, got_d = DeviceBuffer.empty((nq, 100), np.uint32), DeviceBuffer.empty((nq, 100), np.float32)
        br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, want_v.ptr, want_d.ptr, stream=s.h)
        s.synchronize()
        got_v.zero()
        got_d.zero()
        assert noalloc(idx, 100, 400, tq, nq, tk, tl, got_v, got_d, s.h) == 0, lib().artea_last_error()
        s.synchronize()
        mask = np.arange(100)[None, :] < topk[:, None]
        assert np.array_equal(np.where(mask, got_v.numpy(), 0), np.where(mask, want_v.numpy(), 0)), nq
        assert np.array_equal(np.where(mask, got_d.numpy(), 0).view(np.uint32),
                              np.where(mask, want_d.numpy(), 0).view(np.uint32)), nq
        # early stop: thresholds written into the scratch array by a kernel
        top = idx.info()["top"]
        thr = np.array([0.0] + [1e-3 * (h + 1) if metric == 3 else 1e-6 * (h + 1) for h in range(top)], np.float32)
        if kind == 0:
            thr = np.array([0.0] + [4000.0 * (h + 1) for h in range(top)], np.float32)
        br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, want_v.ptr, want_d.ptr, stream=s.h, flags=aa.EARLY_STOP,
                        thresholds=thr)
        s.synchronize()
        assert noalloc(idx, 100, 400, tq, nq, tk, tl, got_v, got_d, s.h, flags=aa.EARLY_STOP, thresholds=thr) == 0
        s.synchronize()
        assert np.array_equal(np.where(mask, got_v.numpy(), 0), np.where(mask, want_v.numpy(), 0)), nq
    # argument errors, before anything is launched
    nq = 4
    qs, topk, queue = batch(kind, 63, nq, dim, 10, [10])
    tq, tk, tl = DeviceBuffer.from_numpy(qs), DeviceBuffer.from_numpy(topk), DeviceBuffer.from_numpy(queue)
    ov, od = DeviceBuffer.empty((nq, 10), np.uint32), DeviceBuffer.empty((nq, 10), np.float32)
    assert noalloc(idx, 10, 2000, tq, nq, tk, tl, ov, od, s.h) == 1
    assert b"exceeds the device maximum" in lib().artea_last_error()
    assert noalloc(idx, 20, 10, tq, nq, tk, tl, ov, od, s.h) == 1
    assert b"must be >= topk" in lib().artea_last_error()

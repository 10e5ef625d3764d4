"""Continuous-batching server: each request's result is the same alone and
in the mix (bit-exact), and equals the plain router's for its (topk, queue)."""
import numpy as np
import pytest

import artea_amd as aa

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix(orc):
    from artea_amd.serving import Server, poisson_arrivals

    vecs = aa.generate(0, 41, 0, 5000, 128)
    idx = aa.DeviceIndex(vecs, 0)
    aa.build_full(idx, max_level=2, build_loops=4, triu_iters=2, seed=3)
    n = 600
    qs = aa.generate(0, 41, 1, n, 128)
    rng = np.random.default_rng(9)
    topk = rng.choice([10, 50, 100], n).astype(np.uint32)
    queue = np.maximum(topk, rng.choice([64, 128, 256, 512, 800], n)).astype(np.uint32)
    srv = Server(idx, max_topk=100, max_queue=800, lanes=3, budget=128)
    v, d, lat, wall = srv.run(qs, topk, queue, poisson_arrivals(n, 20000.0, seed=1))
    assert np.all(lat > 0)
    for i in range(0, n, 12):
        va, da, _, _ = srv.run(qs[i:i + 1], topk[i:i + 1], queue[i:i + 1], np.zeros(1))
        k = int(topk[i])
        assert np.array_equal(va[0, :k], v[i, :k]) and np.array_equal(da[0, :k], d[i, :k]), i
        pv, pd = aa.HierarchicalGraphRouter(idx, None, k, int(queue[i])).batch_query(qs[i:i + 1])
        assert np.array_equal(pv[0], v[i, :k]) and np.array_equal(pd[0], d[i, :k]), i


def test_grouped_batches_take_consecutive_segments():
    """Requests whose queue classes already arrive grouped (ascending here)
    are launched from consecutive rows; results equal the plain router's."""
    from artea_amd.serving import Server, poisson_arrivals

    vecs = aa.generate(0, 43, 0, 5000, 128)
    idx = aa.DeviceIndex(vecs, 0)
    aa.build_full(idx, max_level=2, build_loops=4, triu_iters=2, seed=4)
    n = 400
    qs = aa.generate(0, 43, 1, n, 128)
    topk = np.full(n, 100, np.uint32)
    topk[::3] = 10
    queue = np.sort(np.random.default_rng(2).choice([100, 200, 300, 600], n)).astype(np.uint32)
    srv = Server(idx, max_topk=100, max_queue=800, lanes=2, budget=96)
    v, d, lat, _ = srv.run(qs, topk, queue, poisson_arrivals(n, 50000.0, seed=2))
    assert np.all(lat > 0)
    for k in (10, 100):
        for L in np.unique(queue):
            sel = np.where((topk == k) & (queue == L))[0]
            pv, pd = aa.HierarchicalGraphRouter(idx, None, k, int(L)).batch_query(qs[sel])
            assert np.array_equal(pv, v[sel, :k]) and np.array_equal(pd, d[sel, :k]), (k, L)


def test_native_server_submit_poll_through_the_c_abi():
    """artea_server_*: consecutive tickets, each result equal to the plain
    router's for its (topk, queue) with the slots beyond topk invalid,
    partial polls, status counts, and argument errors."""
    import ctypes as C

    from artea_amd._lib import lib
    from artea_amd.serving import Server

    vecs = aa.generate(0, 47, 0, 4000, 128)
    idx = aa.DeviceIndex(vecs, 0)
    aa.build_full(idx, max_level=2, build_loops=4, triu_iters=2, seed=6)
    n = 300
    qs = aa.generate(0, 47, 1, n, 128)
    rng = np.random.default_rng(5)
    topk = rng.choice([1, 10, 64], n).astype(np.uint32)
    queue = np.maximum(topk, rng.choice([40, 100, 230, 500], n)).astype(np.uint32)
    srv = Server(idx, max_topk=64, max_queue=500, lanes=2, budget=50)
    t0 = srv.submit(qs[:200], topk[:200], queue[:200])
    t1 = srv.submit(qs[200:], topk[200:], queue[200:])
    assert t1 == t0 + 200
    queued, inflight, ready = srv.status()
    assert queued + inflight + ready == n and inflight <= 2 * 50
    tickets = np.empty(7, np.uint64)
    vids = np.empty((7, 64), np.uint32)
    dists = np.empty((7, 64), np.float32)
    got = C.c_uint32()
    seen_v = np.zeros((n, 64), np.uint32)
    seen_d = np.zeros((n, 64), np.float32)
    seen = np.zeros(n, bool)
    import time

    deadline = time.time() + 60
    while time.time() < deadline:
        assert lib().artea_server_poll(srv.h, 7, C.c_void_p(tickets.ctypes.data), C.c_void_p(vids.ctypes.data),
                                       C.c_void_p(dists.ctypes.data), C.byref(got)) == 0
        m = got.value
        assert m <= 7
        ids = (tickets[:m] - np.uint64(t0)).astype(np.int64)
        assert not seen[ids].any()
        seen[ids] = True
        seen_v[ids], seen_d[ids] = vids[:m], dists[:m]
        if seen.all():
            break
    assert seen.all()
    assert srv.status() == (0, 0, 0)
    for k in np.unique(topk):
        for L in np.unique(queue[topk == k]):
            sel = np.where((topk == k) & (queue == L))[0]
            pv, pd = aa.HierarchicalGraphRouter(idx, None, int(k), int(L)).batch_query(qs[sel])
            assert np.array_equal(seen_v[sel, :k], pv) and np.array_equal(seen_d[sel, :k], pd), (k, L)
            assert np.all(seen_v[sel, k:] == 0x7FFFFFFF) and np.all(seen_d[sel, k:] == np.finfo(np.float32).max)
    # poll_into: the same requests again, each result written to its row of
    # a per-request table; equal to the compact poll's rows.
    t2 = srv.submit(qs, topk, queue)
    tv = np.zeros((n, 64), np.uint32)
    td = np.zeros((n, 64), np.float32)
    tix = np.empty(n, np.uint64)
    left = n
    deadline = time.time() + 60
    while left and time.time() < deadline:
        assert lib().artea_server_poll_into(srv.h, t2, n, C.c_void_p(tv.ctypes.data), C.c_void_p(td.ctypes.data),
                                            n, C.c_void_p(tix.ctypes.data), C.byref(got)) == 0
        left -= got.value
    assert left == 0 and np.array_equal(tv, seen_v) and np.array_equal(td.view(np.uint32), seen_d.view(np.uint32))
    # a ticket outside the table is refused and stays ready for a later poll
    t3 = srv.submit(qs[:1], topk[:1], queue[:1])
    while srv.status()[2] == 0 and time.time() < deadline:
        # max_results 0: only advances the server (harvest, launch)
        assert lib().artea_server_poll(srv.h, 0, None, None, None, C.byref(got)) == 0
    assert srv.status()[2] == 1
    assert lib().artea_server_poll_into(srv.h, t3 + 1, 4, C.c_void_p(tv.ctypes.data), C.c_void_p(td.ctypes.data), 4,
                                        C.c_void_p(tix.ctypes.data), C.byref(got)) != 0
    assert srv.status()[2] == 1
    assert lib().artea_server_poll_into(srv.h, t3, 1, C.c_void_p(tv.ctypes.data), C.c_void_p(td.ctypes.data), 4,
                                        C.c_void_p(tix.ctypes.data), C.byref(got)) == 0
    assert got.value == 1 and int(tix[0]) == t3 and np.array_equal(tv[0], seen_v[0])
    assert srv.status() == (0, 0, 0)
    one = qs[:1]
    for k, L in [(0, 10), (65, 100), (10, 501), (20, 10)]:
        with pytest.raises(aa.ArteaError):
            srv.submit(one, np.array([k], np.uint32), np.array([L], np.uint32))
    assert srv.status() == (0, 0, 0)
    with pytest.raises(aa.ArteaError):
        Server(idx, max_topk=64, max_queue=2000)
    with pytest.raises(aa.ArteaError):
        Server(idx, max_topk=64, max_queue=500, lanes=0)


def test_admission_by_device_memory_budget():
    """artea_server_create_budget: a lane holds hbm_bytes / (lanes x request
    bytes) requests, a launch of a queue-size class at most lds_rounds times
    the wavefronts its kernel keeps resident; results equal the plain
    router's and are the same alone and in the mix; a budget below one
    request per lane is an argument error."""
    from artea_amd._lib import ArteaError
    from artea_amd.serving import QUEUE_CLASSES, Server, poisson_arrivals, queue_class

    vecs = aa.generate(0, 53, 0, 4000, 128)
    idx = aa.DeviceIndex(vecs, 0)
    aa.build_full(idx, max_level=2, build_loops=4, triu_iters=2, seed=8)
    per_request = 128 * 4 + 100 * 8 + 8
    lanes, rounds = 3, 1
    hbm = lanes * per_request * 700 + 17
    srv = Server(idx, max_topk=100, max_queue=800, lanes=lanes, hbm_bytes=hbm, lds_rounds=rounds)
    a = srv.admission()
    assert a["lane_requests"] == 700 and a["hbm_bytes"] == lanes * 700 * per_request <= hbm
    lo, hi = QUEUE_CLASSES.index(queue_class(100)), QUEUE_CLASSES.index(queue_class(800))
    for c in range(lo, hi + 1):
        assert a["resident_waves"][c] > 0
        assert a["admit"][c] == min(700, rounds * a["resident_waves"][c]), a
    n = 900
    qs = aa.generate(0, 53, 1, n, 128)
    rng = np.random.default_rng(11)
    topk = rng.choice([10, 50, 100], n).astype(np.uint32)
    queue = np.maximum(topk, rng.choice([64, 128, 256, 512, 800], n)).astype(np.uint32)
    v, d, lat, _ = srv.run(qs, topk, queue, poisson_arrivals(n, 40000.0, seed=3))
    assert np.all(lat > 0)
    for k in (10, 50, 100):
        for L in np.unique(queue):
            sel = np.where((topk == k) & (queue == L))[0]
            if len(sel) == 0:
                continue
            pv, pd = aa.HierarchicalGraphRouter(idx, None, k, int(L)).batch_query(qs[sel])
            assert np.array_equal(pv, v[sel, :k]) and np.array_equal(pd, d[sel, :k]), (k, L)
    for i in range(0, n, 97):
        va, da, _, _ = srv.run(qs[i:i + 1], topk[i:i + 1], queue[i:i + 1], np.zeros(1))
        assert np.array_equal(va[0], v[i]) and np.array_equal(da[0], d[i]), i
    with pytest.raises(ArteaError, match="less than one request per lane"):
        Server(idx, max_topk=100, max_queue=800, lanes=2, hbm_bytes=2 * per_request - 1)

"""Launch ordering, argument checks and the tie-tail overflow path.

* One router used from two streams: its launches share a work counter, so
  the library orders them; both results equal sequential launches.
* artea_refine_upload_rows rejects ids outside the vector array and null
  entries on the host, before any kernel runs.
* A query whose candidate list overflows its tie tail is listed by
  last_overflows, and the server re-runs it on a larger list: its result
  equals the oracle's (the reference's heaps are unbounded,
  router/data_structures/std_candidate_queue.hpp:358-388)."""
import ctypes as C

import numpy as np
import pytest

import artea_amd as aa

pytestmark = pytest.mark.gpu


def csr_of(g):
    hl, lc, groups, nbrs = g.csr()
    return aa.CompactHierarchicalGraph(g.num_vertices, g.top if g.num_vertices else 0, g.ul_max, g.bl_max, g.entry,
                                       np.stack([hl, lc], 1), groups, nbrs)


@pytest.fixture(scope="module")
def small(orc):
    vecs = aa.generate(0, 13, 0, 4000, 128)
    g, _ = orc.build_full(vecs, metric=orc.L2, seed=6, nthreads=8)
    idx = aa.DeviceIndex(vecs, 0)
    idx.set_graph(csr_of(g))
    return vecs, g, idx


def test_one_router_two_streams_equals_sequential(small):
    from artea_amd.device import DeviceBuffer, Stream

    vecs, g, idx = small
    nq = 600
    qa = aa.generate(0, 13, 1, nq, 128)
    qb = aa.generate(0, 13, 2, nq, 128)
    rng = np.random.default_rng(3)
    topk = rng.choice(np.array([10, 100], np.uint32), nq)
    queue = np.maximum(rng.choice(np.array([100, 200, 400], np.uint32), nq), topk).astype(np.uint32)
    tk, tl = DeviceBuffer.from_numpy(topk), DeviceBuffer.from_numpy(queue)
    ta, tb = DeviceBuffer.from_numpy(qa), DeviceBuffer.from_numpy(qb)
    outs = [(DeviceBuffer.empty((nq, 100), np.uint32), DeviceBuffer.empty((nq, 100), np.float32)) for _ in range(4)]
    br = aa.BatchedRouter(idx, 100, 400)
    s1, s2 = Stream(), Stream()
    # sequential references
    br.query_device(ta.ptr, nq, tk.ptr, tl.ptr, outs[0][0].ptr, outs[0][1].ptr, stream=s1.h)
    s1.synchronize()
    br.query_device(tb.ptr, nq, tk.ptr, tl.ptr, outs[1][0].ptr, outs[1][1].ptr, stream=s1.h)
    s1.synchronize()
    mask = np.arange(100)[None, :] < topk[:, None]
    for rep in range(3):
        for o in outs[2:]:
            o[0].zero()
            o[1].zero()
        # enqueued back to back on two streams, no host synchronisation between
        br.query_device(ta.ptr, nq, tk.ptr, tl.ptr, outs[2][0].ptr, outs[2][1].ptr, stream=s1.h)
        br.query_device(tb.ptr, nq, tk.ptr, tl.ptr, outs[3][0].ptr, outs[3][1].ptr, stream=s2.h)
        s1.synchronize()
        s2.synchronize()
        for got, want in ((outs[2], outs[0]), (outs[3], outs[1])):
            assert np.array_equal(np.where(mask, got[0].numpy(), 0), np.where(mask, want[0].numpy(), 0)), rep
            assert np.array_equal(np.where(mask, got[1].numpy(), 0), np.where(mask, want[1].numpy(), 0)), rep
    # a captured launch sequence replayed on s2 while eager launches use s1
    graph = br.capture(ta.ptr, nq, tk.ptr, tl.ptr, outs[2][0].ptr, outs[2][1].ptr)
    outs[2][0].zero()
    br.query_device(tb.ptr, nq, tk.ptr, tl.ptr, outs[3][0].ptr, outs[3][1].ptr, stream=s1.h)
    graph.launch(s2.h)
    s1.synchronize()
    s2.synchronize()
    assert np.array_equal(np.where(mask, outs[2][0].numpy(), 0), np.where(mask, outs[0][0].numpy(), 0))
    assert np.array_equal(np.where(mask, outs[3][0].numpy(), 0), np.where(mask, outs[1][0].numpy(), 0))
    del graph


def test_refine_upload_rejects_bad_rows(small):
    from artea_amd._lib import lib
    from artea_amd.refine import PropagateEngine

    vecs, g, idx = small
    eng = PropagateEngine(idx, cap=8)
    counts = np.zeros(idx.n, np.uint32)
    counts[5] = 2
    ok = np.array([[7, np.float32(1.0).view(np.uint32)], [9, np.float32(2.0).view(np.uint32)]], np.uint32)
    eng.upload(counts, ok)
    c, e = eng.download()
    assert c[5] == 2 and np.array_equal(e, ok)
    bad = ok.copy()
    bad[1, 0] = idx.n  # one past the last vector
    with pytest.raises(aa.ArteaError, match="not below the vector count"):
        eng.upload(counts, bad)
    bad[1, 0] = 0x80000000 | (idx.n + 5)  # the new flag does not hide the id
    with pytest.raises(aa.ArteaError, match="not below the vector count"):
        eng.upload(counts, bad)
    rc = lib().artea_refine_upload_rows(eng.h, counts.ctypes.data_as(C.c_void_p), None)
    assert rc == 1 and "null entries" in lib().artea_last_error().decode()
    counts[5] = 9  # longer than cap
    with pytest.raises(aa.ArteaError, match="longer than cap"):
        eng.upload(counts, np.zeros((9, 2), np.uint32))
    # the rows are unchanged by the rejected uploads
    c, e = eng.download()
    assert c[5] == 2 and np.array_equal(e, ok)


@pytest.fixture(scope="module")
def tied(orc):
    """100 copies of one point (vids 0..99, x = 100) and 20 distinct points
    (vids 100..119, x = 1..20) on a fully connected layer 0 listed in vid
    order.  A query at x <= 0 with queue 48 fills its top-48 set with copies
    from the first chunk; the 20 distinct points then evict 20 copies tied
    with the bound, four more than the 16 entries of tie tail a 64-entry
    list leaves."""
    n = 120
    vecs = np.zeros((n, 16), np.float32)
    vecs[:100, 0] = 100.0
    vecs[100:, 0] = np.arange(1, 21)
    rows = [[u for u in range(n) if u != v] for v in range(n)]
    g = orc.Graph.compact([0] * n, [rows], vecs, orc.L2SQR, n, n)
    idx = aa.DeviceIndex(vecs, 0)
    idx.set_graph(csr_of(g))
    qs = np.zeros((8, 16), np.float32)
    qs[:, 0] = -np.arange(8, dtype=np.float32) * 0.5
    return vecs, g, idx, qs


def test_tail_overflow_is_listed_per_query(orc, tied):
    vecs, g, idx, qs = tied
    r = aa.HierarchicalGraphRouter(idx, None, 10, 48)
    gv, gd = r.batch_query(qs)
    cnt, which = r.last_overflows()
    assert cnt == 8 and list(which) == list(range(8))
    assert r.last_stats()["tail_overflows"] >= 8
    ov, od, _ = orc.batch_query(g, vecs, orc.L2SQR, qs, 10, 48)
    # here the dropped copies could not change the answer (every vertex is
    # found by the first expansion)
    assert np.array_equal(gv, ov) and np.array_equal(gd, od)
    # a router that has not launched yet lists nothing
    r2 = aa.HierarchicalGraphRouter(idx, None, 10, 48)
    assert r2.last_overflows()[0] == 0 and r2.last_stats()["tail_overflows"] == 0


def test_server_reruns_overflowed_requests(orc, tied):
    from artea_amd.serving import Server

    vecs, g, idx, qs = tied
    n = qs.shape[0]
    topk = np.full(n, 10, np.uint32)
    queue = np.full(n, 48, np.uint32)
    srv = Server(idx, max_topk=10, max_queue=48, lanes=1, budget=16)
    v, d, lat, _ = srv.run(qs, topk, queue, np.zeros(n))
    rerun, unresolved = srv.overflows()
    assert rerun == n and unresolved == 0
    ov, od, _ = orc.batch_query(g, vecs, orc.L2SQR, qs, 10, 48)
    assert np.array_equal(v, ov) and np.array_equal(d, od)

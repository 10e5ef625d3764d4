"""GPU router parity: libartea_hip.so's HIP routers against the CPU oracle.

Bar: bit-exact ids and FP32 distances (the device distance kernel follows
the reference's lane order and the queue emulation its try_push rule), on
the reference's KATs, on oracle-built random indexes for every metric, and
for the batched persistent router against the plain one.
"""
import os

import numpy as np
import pytest

import artea_amd as aa

pytestmark = pytest.mark.gpu


def gmm(n, dim=128, seed=1, which=0, kind=0):
    return aa.generate(kind, seed, which, n, dim)


def oracle_graph_to_csr(orc, g):
    hl, lc, groups, nbrs = g.csr()
    vi = np.stack([hl, lc], 1)
    top = g.top if g.num_vertices else 0
    return aa.CompactHierarchicalGraph(g.num_vertices, top, g.ul_max, g.bl_max, g.entry, vi, groups, nbrs)


def device_index(orc, g, vecs, metric):
    idx = aa.DeviceIndex(vecs, metric)
    idx.set_graph(oracle_graph_to_csr(orc, g))
    return idx


def assert_same(gv, gd, ov, od, what=""):
    assert gv.shape == ov.shape
    bad = np.argwhere((gv != ov) | (gd != od))
    assert bad.size == 0, "%s: %d slot mismatches, first %s: gpu (%s, %s) oracle (%s, %s)" % (
        what, len(bad), bad[0], gv[tuple(bad[0])], gd[tuple(bad[0])], ov[tuple(bad[0])], od[tuple(bad[0])])


def test_line_graph_matches_exact_sort_and_restore(orc, tmp_path):
    # test_compact_csr_query.cpp:20-87 on the GPU.
    n = 80
    vecs = np.zeros((n, 128), np.float32)
    vecs[:, 0] = np.arange(n)
    highest = [1 if v < 40 else 0 for v in range(n)]
    g = orc.Graph.compact(highest, [[list(range(n))] * n, [list(range(40)) if v < 40 else [] for v in range(n)]],
                          vecs, orc.L2SQR, n, n)
    idx = device_index(orc, g, vecs, aa.DistanceMetrics.EUCLIDEAN_SQR)
    router = aa.HierarchicalGraphRouter(idx, aa.DistFunc(0, 128), 10, n)
    qs = np.zeros((5, 128), np.float32)
    qs[:, 0] = [-2.0, 12.25, 39.5, 60.75, 90.0]
    v, d = router.batch_query(qs)
    for i in range(5):
        dist = orc.distance_matrix(orc.L2SQR, qs[i:i + 1], vecs)[0]
        order = sorted(range(n), key=lambda u: (dist[u], u))[:10]
        assert list(v[i]) == order and np.array_equal(d[i], dist[order])
    # restored from the oracle's HGRA file into the device index
    path = str(tmp_path / "line.graph")
    g.snapshot(path)
    idx2 = aa.DeviceIndex(vecs, 0)
    idx2.load_hgra(path)
    v2, d2 = aa.HierarchicalGraphRouter(idx2, None, 10, n).batch_query(qs)
    assert_same(v2, d2, v, d, "restored")
    # device snapshot is byte-identical to the oracle's
    p2 = str(tmp_path / "line2.graph")
    idx2.snapshot_hgra(p2)
    assert open(p2, "rb").read() == open(path, "rb").read()


def early_stop_index(orc, metric):
    from test_oracle_kats import early_stop_graph

    g, vecs = early_stop_graph(orc)
    return g, vecs, device_index(orc, g, vecs, metric)


@pytest.mark.parametrize("metric,thr", [(3, [0.0, 8.0, 16.0]), (0, [0.0, 64.0, 256.0])])
def test_early_stop_distance_counts(orc, metric, thr):
    # test_router_early_stop.cpp:122-137: 7 calls plain, 5 with early stop.
    g, vecs, idx = early_stop_index(orc, metric)
    r = aa.HierarchicalGraphRouter(idx, None, 1, 4)
    v, d = r.batch_query(vecs[3:4])
    assert v[0, 0] == 3 and r.last_stats()["ndc"] == 7
    v, d = r.batch_query(vecs[3:4], enable_early_stop=True, thresholds=thr)
    assert v[0, 0] == 3 and d[0, 0] == 0 and r.last_stats()["ndc"] == 5
    v, _ = r.batch_query(vecs)
    assert list(v[:, 0]) == [0, 1, 2, 3]


def test_threshold_overflow_and_queue_errors(orc):
    g, vecs, idx = early_stop_index(orc, 0)
    r = aa.HierarchicalGraphRouter(idx, None, 1, 4)
    with pytest.raises(OverflowError):
        r.batch_query(vecs[3:4], enable_early_stop=True, thresholds=[0, 0.0, 16.0])
    with pytest.raises(aa.ArteaError, match="must be >= topk"):
        aa.HierarchicalGraphRouter(idx, None, 5, 4)


def test_empty_graph_returns_invalid(orc):
    idx = aa.DeviceIndex(np.zeros((0, 16), np.float32), 0)
    idx.set_graph(aa.CompactHierarchicalGraph(0, 0, 4, 4, 0xFFFFFFFF, np.zeros((0, 2), np.uint32), [np.zeros(1, np.uint64)],
                                              np.zeros(0, np.uint32)))
    r = aa.HierarchicalGraphRouter(idx, None, 2, 4)
    v, d = r.batch_query(np.zeros((3, 16), np.float32))
    assert np.all(v == aa.INVALID_VID) and np.all(d == np.finfo(np.float32).max)


@pytest.fixture(scope="module")
def small_l2(orc):
    vecs = gmm(4000, seed=3)
    g, _ = orc.build_full(vecs, metric=orc.L2, seed=5, nthreads=8)
    qs = gmm(300, seed=3, which=1)
    return g, vecs, qs


@pytest.mark.parametrize("rows", ["u8", "f32"])
@pytest.mark.parametrize("topk,queue", [(1, 1), (10, 10), (10, 64), (100, 100), (100, 200), (50, 400), (100, 800),
                                        (100, 1008)])
def test_random_index_parity_l2sqr(orc, small_l2, topk, queue, rows):
    # The data are whole numbers in [0, 255]: the index keeps one-byte rows
    # unless they are dropped.  queue 1008 leaves 16 entries of tail room in
    # the 16-block list, so merges that run past it are frequent.
    g, vecs, qs = small_l2
    idx = device_index(orc, g, vecs, 0)
    assert idx.set_u8_rows(rows == "u8") == (rows == "u8")
    r = aa.HierarchicalGraphRouter(idx, None, topk, queue)
    gv, gd = r.batch_query(qs)
    ov, od, ndc = orc.batch_query(g, vecs, orc.L2SQR, qs, topk, queue)
    assert_same(gv, gd, ov, od, "topk=%d queue=%d" % (topk, queue))
    st = r.last_stats()
    assert st["tail_overflows"] == 0
    assert st["ndc"] >= ndc.sum()  # the lossy visited cache can only add work


def test_random_index_parity_l0_only_and_early_stop(orc, small_l2):
    g, vecs, qs = small_l2
    idx = device_index(orc, g, vecs, 0)
    r = aa.HierarchicalGraphRouter(idx, None, 20, 40)
    gv, gd = r.batch_query_l0_only(qs)
    ov, od, _ = orc.batch_query(g, vecs, orc.L2SQR, qs, 20, 40, flags=orc.L0_ONLY)
    assert_same(gv, gd, ov, od, "l0_only")
    thr = np.array([0] + [40000.0 * (h + 1) for h in range(g.top)], np.float32)
    gv, gd = r.batch_query(qs, enable_early_stop=True, thresholds=thr)
    ov, od, _ = orc.batch_query(g, vecs, orc.L2SQR, qs, 20, 40, flags=orc.EARLY_STOP, thresholds=thr)
    assert_same(gv, gd, ov, od, "early_stop")


@pytest.mark.parametrize("metric,kind,dim", [(2, 1, 100), (1, 1, 100), (2, 2, 96)])
def test_random_index_parity_cosine_dot(orc, metric, kind, dim):
    vecs = gmm(3000, dim=dim, seed=11, kind=kind)
    g, _ = orc.build_full(vecs, metric=metric, seed=2, nthreads=8, shift=0.0)
    qs = gmm(200, dim=dim, seed=11, which=1, kind=kind)
    idx = device_index(orc, g, vecs, metric)
    for topk, queue in ((10, 32), (100, 160)):
        r = aa.HierarchicalGraphRouter(idx, None, topk, queue)
        gv, gd = r.batch_query(qs)
        ov, od, _ = orc.batch_query(g, vecs, metric, qs, topk, queue)
        assert_same(gv, gd, ov, od, "metric=%d topk=%d" % (metric, topk))


def test_distance_matrix_bit_exact(orc):
    from artea_amd.device import DeviceBuffer, Stream

    s = Stream()
    for metric, kind, dim in ((0, 0, 128), (3, 0, 128), (1, 1, 100), (2, 1, 100), (2, 2, 96)):
        a = gmm(37, dim=dim, seed=21, kind=kind)
        b = gmm(53, dim=dim, seed=21, which=1, kind=kind)
        da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
        out = DeviceBuffer.empty((37, 53), np.float32)
        aa._lib.check(aa._lib.lib().artea_distance_matrix_device(metric, da.ptr, 37, db.ptr, 53, a.shape[1], out.ptr,
                                                                 s.h))
        s.synchronize()
        ref = orc.distance_matrix(metric, a, b)
        assert np.array_equal(out.numpy(), ref), "metric %d" % metric


@pytest.mark.parametrize("metric,kind,dim", [(0, 0, 128), (2, 1, 100), (1, 2, 96)])
def test_bruteforce_knn_matches_oracle(orc, metric, kind, dim):
    from artea_amd.device import DeviceBuffer, bruteforce_knn

    base = gmm(2000, dim=dim, seed=31, kind=kind)
    qs = gmm(50, dim=dim, seed=31, which=1, kind=kind)
    ids, d = bruteforce_knn(metric, DeviceBuffer.from_numpy(base), 2000, DeviceBuffer.from_numpy(qs), 50,
                            base.shape[1], 100)
    oi, od = orc.bruteforce_knn(metric, base, qs, 100)
    if metric == 0:  # integer data: FP64 distances are exact, ties broken by vid
        assert np.array_equal(ids, oi) and np.array_equal(d, od)
    else:
        assert np.mean(ids == oi) > 0.999 and np.allclose(d, od, rtol=1e-12, atol=1e-12)


def test_batched_router_mixed_work_matches_plain(orc, small_l2):
    from artea_amd.device import DeviceBuffer

    g, vecs, qs = small_l2
    idx = device_index(orc, g, vecs, 0)
    nq = qs.shape[0]
    rng = np.random.default_rng(4)
    topk = rng.choice([10, 100], nq).astype(np.uint32)
    queue = np.maximum(topk, rng.choice([100, 200, 400], nq)).astype(np.uint32)
    br = aa.BatchedRouter(idx, 100, 400)
    tq, tk, tl = DeviceBuffer.from_numpy(qs), DeviceBuffer.from_numpy(topk), DeviceBuffer.from_numpy(queue)
    ov = DeviceBuffer.empty((nq, 100), np.uint32)
    od = DeviceBuffer.empty((nq, 100), np.float32)
    from artea_amd.device import Stream

    s = Stream()
    br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr, stream=s.h)
    s.synchronize()
    bv = ov.numpy()
    bd = od.numpy()
    for k in (10, 100):
        for L in (100, 200, 400):
            sel = np.where((topk == k) & (queue == L))[0]
            if sel.size == 0:
                continue
            pv, pd = aa.HierarchicalGraphRouter(idx, None, k, L).batch_query(qs[sel])
            assert_same(bv[sel, :k], bd[sel, :k], pv, pd, "k=%d L=%d" % (k, L))
    assert br.last_stats()["tail_overflows"] == 0


def test_device_entry_point_matches_host_buffers(orc, small_l2):
    from artea_amd.device import DeviceBuffer, Stream

    g, vecs, qs = small_l2
    idx = device_index(orc, g, vecs, 0)
    r = aa.HierarchicalGraphRouter(idx, None, 10, 64)
    hv, hd = r.batch_query(qs)
    s = Stream()
    tq = DeviceBuffer.from_numpy(qs)
    ov = DeviceBuffer.empty((qs.shape[0], 10), np.uint32)
    od = DeviceBuffer.empty((qs.shape[0], 10), np.float32)
    r.batch_query_device(tq.ptr, qs.shape[0], ov.ptr, od.ptr, stream=s.h)
    s.synchronize()
    assert_same(ov.numpy(), od.numpy(), hv, hd, "device api")


def test_batched_router_graph_replays_match_eager(orc, small_l2):
    """Captured launch sequence: every replay, with eager launches and host
    synchronisations in between, returns the eager result; changing the
    query buffer's contents between replays changes the result accordingly."""
    from artea_amd.device import DeviceBuffer, Stream

    g, vecs, qs = small_l2
    idx = device_index(orc, g, vecs, 0)
    nq = qs.shape[0]
    rng = np.random.default_rng(11)
    topk = rng.choice([10, 100], nq).astype(np.uint32)
    queue = np.maximum(topk, rng.choice([100, 200], nq)).astype(np.uint32)
    br = aa.BatchedRouter(idx, 100, 200)
    tq, tk, tl = DeviceBuffer.from_numpy(qs), DeviceBuffer.from_numpy(topk), DeviceBuffer.from_numpy(queue)
    ov = DeviceBuffer.empty((nq, 100), np.uint32)
    od = DeviceBuffer.empty((nq, 100), np.float32)
    ev = DeviceBuffer.empty((nq, 100), np.uint32)
    ed = DeviceBuffer.empty((nq, 100), np.float32)
    s = Stream()
    br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, ev.ptr, ed.ptr, stream=s.h)
    s.synchronize()
    eager_v, eager_d = ev.numpy(), ed.numpy()
    eager_stats = br.last_stats()
    graph = br.capture(tq.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr)
    for rep in range(4):
        ov.zero()
        graph.launch(s.h)
        s.synchronize()
        mask = np.arange(100)[None, :] < topk[:, None]
        assert np.array_equal(np.where(mask, ov.numpy(), 0), np.where(mask, eager_v, 0)), "replay %d" % rep
        assert np.array_equal(np.where(mask, od.numpy(), 0), np.where(mask, eager_d, 0)), "replay %d" % rep
        assert br.last_stats() == eager_stats
        # eager work between replays
        br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, ev.ptr, ed.ptr, stream=s.h)
        s.synchronize()
    # new query contents, same buffers: the replay follows them
    qs2 = qs[::-1].copy()
    tq.upload(qs2)
    graph.launch(s.h)
    s.synchronize()
    for k in (10, 100):
        sel = np.where(topk == k)[0]
        for L in (100, 200):
            sub = sel[queue[sel] == L]
            if sub.size == 0:
                continue
            pv, pd = aa.HierarchicalGraphRouter(idx, None, k, L).batch_query(qs2[sub])
            assert_same(ov.numpy()[sub, :k], od.numpy()[sub, :k], pv, pd, "replay after upload k=%d L=%d" % (k, L))


@pytest.mark.parametrize("metric,kind,dim", [(0, 0, 128), (2, 1, 100), (1, 2, 96)])
def test_batched_router_variants_and_autotune_agree(orc, metric, kind, dim):
    """Every kernel variant (gather width, register budget) and the
    autotuner's pick return the plain router's results, for each metric
    family and for list sizes of 2, 4 and 8 register blocks."""
    from artea_amd.device import DeviceBuffer, Stream

    vecs = aa.generate(kind, 5, 0, 3000, dim)
    qs = aa.generate(kind, 5, 3, 48, dim)
    idx = aa.DeviceIndex(vecs, metric)
    aa.build_full(idx, max_level=2, build_loops=3, triu_iters=2, seed=7)
    nq = qs.shape[0]
    s = Stream()
    tq = DeviceBuffer.from_numpy(qs)
    for L in (100, 200, 400):
        topk = np.full(nq, 10, np.uint32)
        queue = np.full(nq, L, np.uint32)
        tk, tl = DeviceBuffer.from_numpy(topk), DeviceBuffer.from_numpy(queue)
        ov = DeviceBuffer.empty((nq, 10), np.uint32)
        od = DeviceBuffer.empty((nq, 10), np.float32)
        pv, pd = aa.HierarchicalGraphRouter(idx, None, 10, L).batch_query(qs)
        br = aa.BatchedRouter(idx, 10, L)
        for v in range(len(br.variant_names)):
            br.set_variant(v)
            assert br.variant() == v
            br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr, stream=s.h)
            s.synchronize()
            assert_same(ov.numpy(), od.numpy(), pv, pd, "metric %d L %d variant %d" % (metric, L, v))
        chosen, ms = br.autotune(tq.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr)
        assert 0 <= chosen < len(ms) and np.all(ms > 0) and br.variant() == chosen
        assert_same(ov.numpy(), od.numpy(), pv, pd, "metric %d L %d autotuned" % (metric, L))
        br.set_variant(-1)
        with pytest.raises(aa.ArteaError):
            br.set_variant(len(ms))


@pytest.mark.parametrize("metric,kind,dim,k", [(0, 0, 128, 100), (0, 0, 128, 1), (3, 2, 96, 10), (1, 1, 100, 50),
                                               (2, 1, 100, 100), (0, 2, 96, 256)])
def test_bruteforce_mfma_equals_fp64(metric, kind, dim, k):
    """The BF16 MFMA filter with FP64 rechecks returns the FP64 kernel's ids
    and distances exactly (every metric, row widths 96 / 112 / 128)."""
    from artea_amd.device import DeviceBuffer, Stream, bruteforce_knn

    n, nq = 40000, 300
    base = aa.generate(kind, 21, 0, n, dim)
    qs = aa.generate(kind, 21, 1, nq, dim)
    D = base.shape[1]
    s = Stream()
    tb, tq = DeviceBuffer.from_numpy(base), DeviceBuffer.from_numpy(qs)
    fi, fd = bruteforce_knn(metric, tb, n, tq, nq, D, k, s)
    st = {}
    mi, md = bruteforce_knn(metric, tb, n, tq, nq, D, k, s, method="mfma", stats=st)
    assert st["sample_rows"] > 0, "the MFMA path was not taken"
    assert st["kept_pairs"] >= nq * k
    assert np.array_equal(mi, fi) and np.array_equal(md, fd)


def test_bruteforce_mfma_overflow_falls_back_to_fp64():
    """Every pair tied: each query keeps all 12000 rows, more than the 4096
    candidate slots, so every query is recomputed by the FP64 kernel."""
    from artea_amd.device import DeviceBuffer, Stream, bruteforce_knn

    row = aa.generate(0, 3, 0, 1, 128)
    base = np.repeat(row, 12000, axis=0)
    qs = aa.generate(0, 3, 1, 40, 128)
    s = Stream()
    tb, tq = DeviceBuffer.from_numpy(base), DeviceBuffer.from_numpy(qs)
    st = {}
    mi, md = bruteforce_knn(0, tb, 12000, tq, 40, 128, 10, s, method="mfma", stats=st)
    fi, fd = bruteforce_knn(0, tb, 12000, tq, 40, 128, 10, s)
    assert st["fp64_fallback_queries"] == 40
    assert np.array_equal(mi, fi) and np.array_equal(md, fd)
    assert np.array_equal(mi, np.tile(np.arange(10, dtype=np.uint32), (40, 1)))

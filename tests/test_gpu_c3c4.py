"""Parity of the unit-norm 96-wide L2 path the 10M x 96 build and serving
workloads run on: the 1024-component unit-norm mixture (generator kind 2),
row width 96 (six 16-float chunks), L2SQR and L2 search, with the build's
settings (geometric levels f = 1/32, 15 x 4 schedule, prefill 0.4, scale
1.10, shift 1.5, caps 32 / 64).  These values are not whole numbers, so the
FP32 chains' order of operations is what makes the distances equal: every
comparison here is bit-exact against the oracle.

Reference call sites: router/hierarchical_graph_router.hpp:293-349
(batch_query), refiner/updaters/triangle_updater.hpp:114-237 (the join
whose BF16 form must give the FP32 join's edges)."""
import hashlib

import numpy as np
import pytest

import artea_amd as aa
from artea_amd.refine import JOIN_BF16, JOIN_FP32, build_full

pytestmark = pytest.mark.gpu

SEED = 20261203 + 3
BUILD = dict(ul_max=32, bl_max=64, build_loops=15, triu_iters=4, prefill=0.4, scale=1.10, shift=1.5,
             l0_min_distance=1.0, level_fraction=1.0 / 32)
QUEUES = (64, 128, 256, 512, 800)
TOPKS = (10, 50, 100)


def assert_same(gv, gd, ov, od, what=""):
    assert gv.shape == ov.shape, what
    bad = np.argwhere((gv != ov) | (gd.view(np.uint32) != od.view(np.uint32)))
    assert bad.size == 0, "%s: %d slot mismatches, first %s: gpu (%s, %r) oracle (%s, %r)" % (
        what, len(bad), bad[0], gv[tuple(bad[0])], gd[tuple(bad[0])], ov[tuple(bad[0])], od[tuple(bad[0])])


@pytest.fixture(scope="module")
def unit96(orc, tmp_path_factory):
    """A 6,000-vector unit-norm index built by the oracle with the C3
    settings (L2 build), restored from its HGRA file into a device index
    per search metric."""
    n, nq = 6000, 240
    vecs = aa.generate(2, SEED, 0, n, 96)
    qs = aa.generate(2, SEED, 1, nq, 96)
    g, _ = orc.build_full(vecs, metric=orc.L2, seed=SEED, nthreads=8, **BUILD)
    path = str(tmp_path_factory.mktemp("c3") / "unit96.graph")
    g.snapshot(path)
    return vecs, qs, g, path


def device_index(vecs, path, metric):
    idx = aa.DeviceIndex(vecs, metric)
    assert not idx.u8_rows  # real-valued data: the FP32 rows only
    idx.load_hgra(path)
    return idx


def test_gpu_build_equals_oracle_build(orc, unit96, tmp_path):
    """The GPU build with either join writes the oracle build's HGRA file
    byte for byte at these settings."""
    vecs, _, g, path = unit96
    ref = open(path, "rb").read()
    for join in (JOIN_FP32, JOIN_BF16):
        idx = aa.DeviceIndex(vecs, aa.DistanceMetrics.EUCLIDEAN_SQR)
        build_full(idx, seed=SEED, join=join, **BUILD)
        p = str(tmp_path / ("gpu%d.graph" % join))
        idx.snapshot_hgra(p)
        assert open(p, "rb").read() == ref, "join %d" % join


@pytest.mark.parametrize("metric", [0, 3])
def test_plain_router_equals_oracle(orc, unit96, metric):
    vecs, qs, g, path = unit96
    idx = device_index(vecs, path, metric)
    for k in TOPKS:
        for L in QUEUES:
            if L < k:
                continue
            r = aa.HierarchicalGraphRouter(idx, None, k, L)
            gv, gd = r.batch_query(qs)
            ov, od, ndc = orc.batch_query(g, vecs, metric, qs, k, L)
            assert_same(gv, gd, ov, od, "metric %d topk %d queue %d" % (metric, k, L))
            assert r.last_stats()["tail_overflows"] == 0
            assert r.last_stats()["ndc"] >= ndc.sum()


@pytest.mark.parametrize("metric", [0, 3])
def test_batched_router_every_slot_and_autotune_equal_oracle(orc, unit96, metric):
    """Mixed topk {10, 50, 100} and queue {64 .. 800} per query in one
    launch: every FP32 kernel slot and the autotuned pick return the
    oracle's lists for every query."""
    from artea_amd.device import DeviceBuffer, Stream

    vecs, qs, g, path = unit96
    idx = device_index(vecs, path, metric)
    nq = qs.shape[0]
    rng = np.random.default_rng(17)
    topk = rng.choice(np.array(TOPKS, np.uint32), nq)
    queue = np.maximum(rng.choice(np.array(QUEUES, np.uint32), nq), topk).astype(np.uint32)
    want_v = np.zeros((nq, 100), np.uint32)
    want_d = np.zeros((nq, 100), np.float32)
    for k in TOPKS:
        for L in np.unique(queue):
            sel = np.where((topk == k) & (queue == L))[0]
            if sel.size:
                ov, od, _ = orc.batch_query(g, vecs, metric, qs[sel], int(k), int(L))
                want_v[sel, :k], want_d[sel, :k] = ov, od
    mask = np.arange(100)[None, :] < topk[:, None]
    s = Stream()
    tq, tk, tl = DeviceBuffer.from_numpy(qs), DeviceBuffer.from_numpy(topk), DeviceBuffer.from_numpy(queue)
    ov = DeviceBuffer.empty((nq, 100), np.uint32)
    od = DeviceBuffer.empty((nq, 100), np.float32)
    br = aa.BatchedRouter(idx, 100, 800)
    assert len(br.variant_names) == 4
    for v in range(4):
        br.set_variant(v)
        ov.zero()
        br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr, stream=s.h)
        s.synchronize()
        assert_same(np.where(mask, ov.numpy(), 0), np.where(mask, od.numpy(), 0), np.where(mask, want_v, 0),
                    np.where(mask, want_d, 0), "metric %d slot %s" % (metric, br.variant_names[v]))
        assert br.last_stats()["tail_overflows"] == 0 and br.last_overflows()[0] == 0
    chosen, ms = br.autotune(tq.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr)
    assert 0 <= chosen < 4 and br.variant() == chosen
    assert_same(np.where(mask, ov.numpy(), 0), np.where(mask, od.numpy(), 0), np.where(mask, want_v, 0),
                np.where(mask, want_d, 0), "metric %d autotuned" % metric)


def test_serving_mix_equals_oracle(orc, unit96):
    """The serving mix (topk {10, 50, 100}, queue {64 .. 800}, Poisson
    arrivals) through the native server: every request's list equals the
    oracle's for its (topk, queue), alone and in the mix."""
    from artea_amd.serving import Server, poisson_arrivals

    vecs, qs, g, path = unit96
    idx = device_index(vecs, path, 0)
    n = 480
    rng = np.random.default_rng(23)
    pick = rng.integers(0, qs.shape[0], n)
    topk = rng.choice(np.array(TOPKS, np.uint32), n)
    queue = np.maximum(rng.choice(np.array(QUEUES, np.uint32), n), topk).astype(np.uint32)
    srv = Server(idx, max_topk=100, max_queue=800, lanes=4, budget=64)
    v, d, lat, _ = srv.run(qs[pick], topk, queue, poisson_arrivals(n, 20000.0, seed=5))
    assert np.all(lat > 0)
    for k in TOPKS:
        for L in np.unique(queue):
            sel = np.where((topk == k) & (queue == L))[0]
            if sel.size == 0:
                continue
            ov, od, _ = orc.batch_query(g, vecs, orc.L2SQR, qs[pick[sel]], int(k), int(L))
            assert_same(v[sel, :k], d[sel, :k], ov, od, "serving topk %d queue %d" % (k, L))
            assert np.all(v[sel, k:] == aa.INVALID_VID)
    for i in range(0, n, 40):
        va, da, _, _ = srv.run(qs[pick[i:i + 1]], topk[i:i + 1], queue[i:i + 1], np.zeros(1))
        assert np.array_equal(va[0], v[i]) and np.array_equal(da[0].view(np.uint32), d[i].view(np.uint32)), i
    assert srv.overflows() == (0, 0)


def test_million_scale_build_fp32_join_equals_bf16_join():
    """1,000,000 x 96 unit-norm vectors, the 10M build's settings: the BF16
    MFMA join and the FP32 join give the same graph (edge digest and every
    CSR array), and the index routes (recall@100 at queue 400 against the
    exact FP64 ground truth: 0.947 at this size when measured)."""
    from artea_amd.device import DeviceBuffer, Stream, bruteforce_knn

    n, nq = 1_000_000, 400
    base = aa.generate(2, SEED, 0, n, 96)
    qs = aa.generate(2, SEED, 1, nq, 96)
    idx = aa.DeviceIndex(base, aa.DistanceMetrics.EUCLIDEAN_SQR)
    digests, graphs = {}, {}
    for join in (JOIN_FP32, JOIN_BF16):
        build_full(idx, seed=SEED, join=join, **BUILD)
        gr = idx.export_graph()
        digests[join] = hashlib.sha256(gr.nbrs.tobytes() + b"".join(o.tobytes() for o in gr.group_offsets) +
                                       gr.vertex_info.tobytes()).hexdigest()
        graphs[join] = gr
    a, b = graphs[JOIN_FP32], graphs[JOIN_BF16]
    assert digests[JOIN_FP32] == digests[JOIN_BF16]
    assert a.entry == b.entry and a.top == b.top and np.array_equal(a.nbrs, b.nbrs)
    assert a.nbrs.size > 50 * n  # layer 0 rows near their 64 cap
    s = Stream()
    gi, _ = bruteforce_knn(0, DeviceBuffer.borrow(idx.device_vectors(), base.shape, np.float32), n,
                           DeviceBuffer.from_numpy(qs), nq, 96, 100, s, method="mfma")
    v, _ = aa.HierarchicalGraphRouter(idx, None, 100, 400).batch_query(qs)
    rec = np.mean([len(set(v[i]) & set(gi[i])) / 100.0 for i in range(nq)])
    assert rec >= 0.93, rec


@pytest.mark.parametrize("metric,kind,dim", [(0, 1, 100), (3, 1, 100)])
def test_l2_on_power_law_rows_equals_oracle(orc, metric, kind, dim):
    """L2SQR / L2 on the 112-wide power-law rows (real values)."""
    vecs = aa.generate(kind, 29, 0, 3000, dim)
    qs = aa.generate(kind, 29, 1, 160, dim)
    g, _ = orc.build_full(vecs, metric=orc.L2, seed=4, nthreads=8, shift=0.0)
    idx = aa.DeviceIndex(vecs, metric)
    hl, lc, groups, nbrs = g.csr()
    idx.set_graph(aa.CompactHierarchicalGraph(g.num_vertices, g.top, g.ul_max, g.bl_max, g.entry,
                                              np.stack([hl, lc], 1), groups, nbrs))
    for k, L in ((10, 64), (100, 200), (50, 800)):
        gv, gd = aa.HierarchicalGraphRouter(idx, None, k, L).batch_query(qs)
        ov, od, _ = orc.batch_query(g, vecs, metric, qs, k, L)
        assert_same(gv, gd, ov, od, "metric %d topk %d queue %d" % (metric, k, L))

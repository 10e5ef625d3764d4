"""Pins the CPU oracle against the reference's own known-answer tests.

Each test restates one reference test case (cited) on the oracle, or checks
an invariant that the mathematics fixes (closed forms, brute force on tiny
inputs).  The oracle is the parity checker for the HIP path, so a mistake
here would hide a mistake there.
"""
import math
import os
import struct

import numpy as np
import pytest


def vec16(*xs, dim=16):
    v = np.zeros(dim, np.float32)
    v[: len(xs)] = xs
    return v


# ------------------------------------------------------------------ distance
def test_euclidean_345(orc):
    # unit_tests/test_euclidean_distance.cpp:27-41: (0,0) vs (3,4) -> 5 / 25.
    for dim in (16, 128, 960):
        a, b = vec16(0, 0, dim=dim), vec16(3, 4, dim=dim)
        assert orc.distance(orc.L2, a, b) == np.float32(5.0)
        assert orc.distance(orc.L2SQR, a, b) == np.float32(25.0)


def test_euclidean_matches_fp64(orc):
    # unit_tests/test_euclidean_distance.cpp:43-65: FP64 reference, rel 2e-6.
    rng = np.random.default_rng(7)
    for dim in (16, 32, 96, 112, 128, 304, 384, 960):
        for _ in range(20):
            a = rng.standard_normal(dim).astype(np.float32)
            b = rng.standard_normal(dim).astype(np.float32)
            gt = float(np.sum((a.astype(np.float64) - b) ** 2))
            got = float(orc.distance(orc.L2SQR, a, b))
            assert abs(got - gt) <= 2e-6 * max(1.0, gt)
            assert abs(float(orc.distance(orc.L2, a, b)) - math.sqrt(gt)) <= 2e-6 * max(1.0, math.sqrt(gt))


def test_cosine_edge_cases(orc):
    # unit_tests/test_cosine_distance.cpp:90-114.
    rng = np.random.default_rng(3)
    a = rng.standard_normal(112).astype(np.float32)
    assert abs(orc.distance(orc.COSINE, a, a)) < 1e-5
    assert abs(orc.distance(orc.COSINE, a, a * np.float32(3.5))) < 1e-5
    assert abs(orc.distance(orc.COSINE, a, -a) - 2.0) < 1e-5
    z = np.zeros(112, np.float32)
    assert orc.distance(orc.COSINE, z, a) == np.float32(1.0)
    assert orc.distance(orc.COSINE, z, z) == np.float32(1.0)


def test_inner_product_is_negated_dot(orc):
    # unit_tests/test_cosine_distance.cpp:116-135 (tolerance 1e-3 * max(1,|dot|)).
    rng = np.random.default_rng(5)
    for _ in range(16):
        a = rng.standard_normal(208).astype(np.float32)
        b = rng.standard_normal(208).astype(np.float32)
        dot = float(np.dot(a.astype(np.float64), b))
        assert abs(float(orc.distance(orc.DOT, a, b)) + dot) <= 1e-3 * max(1.0, abs(dot))


def test_integer_data_l2sqr_is_exact(orc):
    # Integer-valued rows in [0,255]: every partial sum is an integer < 2^24,
    # so the FP32 chain equals the exact FP64 value (configs 0 and 1).
    rng = np.random.default_rng(11)
    a = rng.integers(0, 256, (50, 128)).astype(np.float32)
    b = rng.integers(0, 256, (40, 128)).astype(np.float32)
    m = orc.distance_matrix(orc.L2SQR, a, b)
    gt = ((a[:, None, :].astype(np.float64) - b[None]) ** 2).sum(-1)
    assert np.array_equal(m.astype(np.float64), gt)


def _numpy_lane_tree(a, b, kind):
    """Independent numpy restatement: 16 lane accumulators in FP32 with
    FMA emulated in FP64 (exact product, one rounding), GCC reduce tree."""
    D = a.shape[0]
    acc = np.zeros(16, np.float32)
    for c in range(D // 16):
        x = a[c * 16:(c + 1) * 16].astype(np.float64)
        y = b[c * 16:(c + 1) * 16].astype(np.float64)
        if kind == "l2":
            d = (a[c * 16:(c + 1) * 16] - b[c * 16:(c + 1) * 16]).astype(np.float64)
            acc = (d * d + acc.astype(np.float64)).astype(np.float32)
        else:
            acc = (x * y + acc.astype(np.float64)).astype(np.float32)
    t8 = (acc[8:] + acc[:8]).astype(np.float32)
    t4 = (t8[4:] + t8[:4]).astype(np.float32)
    return np.float32(np.float32(t4[0] + t4[2]) + np.float32(t4[1] + t4[3]))


def test_lane_order_matches_independent_restatement(orc):
    # Bit order of the 16-lane FMA chain + reduce tree.  FP64 fma emulation is
    # exact for a product of two FP32 plus an FP32 addend except in double
    # rounding corner cases, which these random inputs do not hit in practice;
    # any reordering of chunks or lanes changes many of these results.
    rng = np.random.default_rng(17)
    mismatches = 0
    for _ in range(200):
        a = rng.standard_normal(128).astype(np.float32)
        b = rng.standard_normal(128).astype(np.float32)
        mismatches += orc.distance(orc.L2SQR, a, b) != _numpy_lane_tree(a, b, "l2")
        mismatches += orc.distance(orc.DOT, a, b) != -_numpy_lane_tree(a, b, "dot")
    assert mismatches <= 2


# --------------------------------------------------------------------- queue
def test_queue_eviction_and_threshold(orc):
    # StdCandidateQueue semantics (std_candidate_queue.hpp:358-388, 503-513):
    # distance-only rejection at the lower bound, stale evictees stay in U.
    q = orc.Queue(3)
    assert q.try_push(10, 5.0)
    assert q.try_push(11, 3.0)
    assert q.try_push(12, 4.0)
    assert q.lower_bound == 5.0
    assert not q.try_push(13, 5.0)  # equal to lb: rejected (distance only)
    assert not q.try_push(2, 5.0)  # smaller vid, same distance: still rejected
    assert q.try_push(14, 1.0)  # evicts (5.0, 10)
    assert q.result_size == 3
    assert q.unexplored_size == 4  # evictee stays unexplored
    assert q.lower_bound == 4.0
    v, d = q.extract(3)
    assert list(v) == [14, 11, 12] and list(d) == [1.0, 3.0, 4.0]


def test_queue_termination_and_pop_order(orc):
    q = orc.Queue(2)
    q.try_push(5, 2.0)
    q.try_push(3, 2.0)  # tie in distance, order by vid
    assert q.pop_best() == (3, 2.0)
    assert not q.should_terminate()  # U top (5,2.0) == lb, not >
    q.try_push(7, 1.0)  # evicts (2.0, 5)
    assert q.pop_best() == (7, 1.0)
    assert q.pop_best() == (5, 2.0)  # stale evictee still pops
    assert q.should_terminate()  # U empty


def test_queue_extract_fewer(orc):
    q = orc.Queue(8)
    for vid, d in [(4, 0.5), (1, 0.5), (9, 0.25)]:
        q.try_push(vid, d)
    v, d = q.extract(2)
    assert list(v) == [9, 1] and list(d) == [0.25, 0.5]


# -------------------------------------------------------------------- router
def line_graph(orc, tmp_path=None):
    """unit_tests/test_compact_csr_query.cpp:20-39: 80 points on a line,
    vids < 40 at level 1; L0 rows complete (80), L1 rows the first 40."""
    n = 80
    vecs = np.zeros((n, 128), np.float32)
    vecs[:, 0] = np.arange(n)
    highest = [1 if v < 40 else 0 for v in range(n)]
    l0 = [list(range(n)) for _ in range(n)]
    l1 = [list(range(40)) if v < 40 else [] for v in range(n)]
    g = orc.Graph.compact(highest, [l0, l1], vecs, orc.L2SQR, ul_max=n, bl_max=n)
    return g, vecs


def test_compact_csr_query_matches_exact_sort(orc, tmp_path):
    g, vecs = line_graph(orc)
    assert g.top == 1
    path = str(tmp_path / "greedy.graph")
    g.snapshot(path)
    r = orc.Graph.restore(path)
    qs = np.zeros((5, 128), np.float32)
    qs[:, 0] = [-2.0, 12.25, 39.5, 60.75, 90.0]
    for graph in (g, r):
        v, d, _ = orc.batch_query(graph, vecs, orc.L2SQR, qs, topk=10, queue=80)
        for i in range(5):
            dist = orc.distance_matrix(orc.L2SQR, qs[i:i + 1], vecs)[0]
            order = sorted(range(80), key=lambda u: (dist[u], u))[:10]
            assert list(v[i]) == order
            assert np.array_equal(d[i], dist[order])


def early_stop_graph(orc, top=2):
    """unit_tests/test_router_early_stop.cpp:64-94: L2 singleton entry 0;
    L1 chain 0->1->2->3; L0 complete."""
    counts = [0] * (top + 1)
    counts[top] = 1
    counts[1 if top else 0] += 3
    vi = []
    locals_ = [0] * (top + 1)
    rows = {}
    for vid in range(4):
        h = top if vid == 0 else min(top, 1)
        vi += [h, locals_[h]]
        locals_[h] += 1
        for lvl in range(h + 1):
            if lvl == 0:
                rows[(vid, 0)] = [o for o in range(4) if o != vid]
            elif lvl == 1 and vid < 3:
                rows[(vid, 1)] = [vid + 1]
            else:
                rows[(vid, lvl)] = []
    # lay out CSR: groups in order h = 0..top, rows local*(h+1)+(h-l)
    group_offsets, nbrs = [], []
    vi_arr = np.array(vi, np.uint32).reshape(4, 2)
    total = 0
    for h in range(top + 1):
        members = sorted([v for v in range(4) if vi_arr[v, 0] == h], key=lambda v: vi_arr[v, 1])
        offs = [0] * (len(members) * (h + 1) + 1)
        for v in members:
            loc = vi_arr[v, 1]
            for lvl in range(h, -1, -1):
                ri = loc * (h + 1) + (h - lvl)
                offs[ri] = total
                nbrs += rows[(v, lvl)]
                total += len(rows[(v, lvl)])
        offs[-1] = total
        group_offsets.append(offs)
    g = orc.Graph.from_csr(4, top, 4, 4, 0, vi_arr, group_offsets, nbrs)
    vecs = np.zeros((4, 16), np.float32)
    vecs[:, 0] = [12, 8, 4, 0]
    return g, vecs


@pytest.mark.parametrize("metric,thr", [(3, [0.0, 8.0, 16.0]), (0, [0.0, 64.0, 256.0])])
def test_router_early_stop_distance_calls(orc, metric, thr):
    # test_router_early_stop.cpp:122-137: 7 distance calls without early stop,
    # 5 with T1=8, T2=16 (squared for L2SQR); the result is vid 3 either way.
    g, vecs = early_stop_graph(orc)
    v, d, cur, ndc = orc.query_trace(g, vecs, metric, vecs[3], topk=1, queue=4)
    assert list(v) == [3] and ndc == 7
    assert cur[1] == 3
    v, d, cur, ndc = orc.query_trace(g, vecs, metric, vecs[3], topk=1, queue=4, flags=orc.EARLY_STOP,
                                     thresholds=np.array(thr, np.float32))
    assert list(v) == [3] and d[0] == 0.0 and ndc == 5
    assert cur[2] == 0 and cur[1] == 1


def test_router_early_stop_inclusive_threshold(orc):
    # test_router_early_stop.cpp:99-120 via the hierarchy: a level-1 threshold
    # just below 8 walks one more hop (cursor 2) than threshold 8 (cursor 1).
    g, vecs = early_stop_graph(orc)
    _, _, cur, _ = orc.query_trace(g, vecs, orc.L2, vecs[3], 1, 4, orc.EARLY_STOP,
                                   np.array([0, 8.0, 1e30], np.float32))
    assert cur[1] == 1
    _, _, cur, _ = orc.query_trace(g, vecs, orc.L2, vecs[3], 1, 4, orc.EARLY_STOP,
                                   np.array([0, np.nextafter(np.float32(8), np.float32(0)), 1e30], np.float32))
    assert cur[1] == 2


def test_router_batch_and_bottom_only(orc):
    g, vecs = early_stop_graph(orc)
    v, _, _ = orc.batch_query(g, vecs, orc.L2, vecs, topk=1, queue=4)
    assert list(v[:, 0]) == [0, 1, 2, 3]
    b, vecs = early_stop_graph(orc, top=0)
    v, _, _ = orc.batch_query(b, vecs, orc.L2, vecs[3:4], topk=1, queue=4)
    assert v[0, 0] == 3


def test_router_empty_graph_pads_invalid(orc):
    g = orc.Graph.from_csr(0, 0, 4, 4, 0xFFFFFFFF, np.zeros(0, np.uint32), [[0]], [])
    vecs = np.zeros((4, 16), np.float32)
    v, d, ndc = orc.batch_query(g, vecs, orc.L2, vecs, topk=2, queue=4)
    assert np.all(v == orc.INVALID_VID) and np.all(d == np.finfo(np.float32).max) and np.all(ndc == 0)


def test_router_rejects_queue_below_topk(orc):
    g, vecs = early_stop_graph(orc)
    with pytest.raises(RuntimeError, match="must be >= topk"):
        orc.batch_query(g, vecs, orc.L2, vecs, topk=5, queue=4)


def test_router_complete_graph_equals_brute_force(orc):
    # With a complete L0 and queue >= N, the beam visits every vertex, so the
    # result equals brute force by (dist, vid) including ties.
    rng = np.random.default_rng(1)
    n = 60
    vecs = rng.integers(0, 6, (n, 16)).astype(np.float32)  # many ties
    highest = [0] * n
    l0 = [[u for u in range(n) if u != v] for v in range(n)]
    g = orc.Graph.compact(highest, [l0], vecs, orc.L2SQR, 64, 64)
    qs = rng.integers(0, 6, (20, 16)).astype(np.float32)
    v, d, _ = orc.batch_query(g, vecs, orc.L2SQR, qs, topk=10, queue=n)
    bv, bd = orc.bruteforce_knn(orc.L2SQR, vecs, qs, 10)
    assert np.array_equal(v, bv) and np.array_equal(d.astype(np.float64), bd)


# -------------------------------------------------------------- HGRA format
def test_hgra_wire_format_and_round_trip(orc, tmp_path):
    g, vecs = early_stop_graph(orc)
    path = str(tmp_path / "g.graph")
    g.snapshot(path)
    raw = open(path, "rb").read()
    # header 7 x u32 (hierarchical_graph_file_manager.hpp:26-33)
    hdr = struct.unpack("<7I", raw[:28])
    assert hdr == (0x48475241, 1, 4, 2, 4, 4, 0)
    # vid 0: highest 2; levels 2,1,0 with counts 0, 1 [1], 3 [1,2,3]
    words = struct.unpack("<%dI" % ((len(raw) - 28) // 4), raw[28:])
    assert list(words[:9]) == [2, 0, 1, 1, 3, 1, 2, 3, 1]
    r = orc.Graph.restore(path)
    path2 = str(tmp_path / "g2.graph")
    r.snapshot(path2)
    assert open(path2, "rb").read() == raw


def test_hgra_rejects_corrupt_files(orc, tmp_path):
    g, vecs = early_stop_graph(orc)
    path = str(tmp_path / "g.graph")
    g.snapshot(path)
    raw = bytearray(open(path, "rb").read())
    cases = {
        "magic": (lambda b: b.__setitem__(slice(0, 4), b"\0\0\0\0"), "bad magic"),
        "version": (lambda b: b.__setitem__(slice(4, 8), struct.pack("<I", 2)), "unsupported version"),
        "trunc": (lambda b: b.__delitem__(slice(len(b) - 6, len(b))), "truncated"),
        "entry": (lambda b: b.__setitem__(slice(24, 28), struct.pack("<I", 9)), "invalid header"),
        "count": (lambda b: b.__setitem__(slice(28 + 8, 28 + 12), struct.pack("<I", 99)), "capacity"),
        "nbr": (lambda b: b.__setitem__(slice(28 + 12, 28 + 16), struct.pack("<I", 77)), "invalid neighbor"),
        "level": (lambda b: b.__setitem__(slice(28, 32), struct.pack("<I", 5)), "top_level_id"),
    }
    for name, (mut, msg) in cases.items():
        b = bytearray(raw)
        mut(b)
        p = str(tmp_path / (name + ".graph"))
        open(p, "wb").write(bytes(b))
        with pytest.raises(RuntimeError, match=msg):
            orc.Graph.restore(p)


def test_compactor_trims_small_top_buckets(orc):
    # hierarchical_graph_compactor.hpp:146-179: buckets above the first one
    # with >= 32 vertices are demoted into it, appended in source order.
    n = 40
    vecs = np.zeros((n, 16), np.float32)
    vecs[:, 0] = np.arange(n)
    highest = [2 if v in (3, 7) else 1 for v in range(n)]
    lv = [[[u for u in range(n) if u != v][:4] for v in range(n)] for _ in range(3)]
    lv[2] = [[7] if v == 3 else ([3] if v == 7 else []) for v in range(n)]
    g = orc.Graph.compact(highest, lv, vecs, orc.L2, 8, 8)
    assert g.top == 1
    assert g.highest(3) == 1 and g.highest(7) == 1
    assert g.local(3) == 38 and g.local(7) == 39  # appended after the level-1 bucket
    # entry = vertex closest to the FP64 centroid (x = 19.5): vids 19 and 20
    # tie at 0.5; the earlier bucket position wins.
    assert g.entry == 19


# -------------------------------------------------------------------- refine
GRID = [(c, r) for r in range(3) for c in range(4)]  # test_propagate_engine.cpp:47-58


def grid_vecs():
    v = np.zeros((12, 16), np.float32)
    for i, (x, y) in enumerate(GRID):
        v[i, 0], v[i, 1] = x, y
    return v


def test_triangle_prunes_farthest_retained_first(orc):
    # test_propagate_engine.cpp:235-259.
    v = grid_vecs()
    L = orc.Layer(v, orc.L2, cap=8)
    L.set_row_raw(0, orc.nbrs([(u, orc.distance(orc.L2, v[0], v[u])) for u in (1, 8, 10)], new=True))
    L.triangle(1.0, 0.0)
    row = L.row(0)
    assert [e["vid_status"] for e in row] == [1, 8]  # both old
    r8 = L.row(8)
    assert len(r8) == 1 and r8[0]["vid_status"] == (10 | orc.NBR_NEW) and r8[0]["dist"] == 2.0
    assert len(L.row(1)) == 0


def test_triangle_selects_closest_recommendation(orc):
    # test_propagate_engine.cpp:261-284.
    v = grid_vecs()
    L = orc.Layer(v, orc.L2, cap=8)
    L.set_row_raw(0, orc.nbrs([(u, orc.distance(orc.L2, v[0], v[u])) for u in (1, 8, 11)], new=True))
    L.triangle(2.0, 0.0)
    assert [e["vid_status"] for e in L.row(0)] == [1, 8, 11]
    r1 = L.row(1)
    assert len(r1) == 1 and r1[0]["vid_status"] & 0x7FFFFFFF == 11
    assert r1[0]["dist"] == np.float32(math.sqrt(8.0))
    assert len(L.row(8)) == 0


@pytest.mark.parametrize("alpha,shift,retain", [(2.0, 0.0, False), (2.0, 0.2, False), (2.0, 0.4, True),
                                                (2.0, 1.0 / 3.0, True), (1.0, 1.25, True), (1.0, 1.5, True),
                                                (1.0, 2.0, True)])
def test_triangle_alpha_dependent_shift(orc, alpha, shift, retain):
    # test_propagate_engine.cpp:286-318.
    v = grid_vecs()
    L = orc.Layer(v, orc.L2, cap=8)
    L.set_row_raw(0, orc.nbrs([(u, orc.distance(orc.L2, v[0], v[u])) for u in (2, 3)], new=True))
    L.triangle(alpha, shift)
    vids = [e["vid_status"] & 0x7FFFFFFF for e in L.row(0)]
    assert 2 in vids and (3 in vids) == retain
    r2 = L.row(2)
    assert len(r2) == 1 and r2[0]["vid_status"] & 0x7FFFFFFF == 3 and r2[0]["dist"] == 1.0


def rng_property_ok(v, rows, metric, scale=1.0, shift=0.0):
    n = v.shape[0]
    D = np.array([[float(np.float32(np.sqrt(np.sum((v[a] - v[b]).astype(np.float64) ** 2)))) for b in range(n)]
                  for a in range(n)])
    for u in range(n):
        for e in rows[u]:
            w_ = e["vid_status"] & 0x7FFFFFFF
            thr = (e["dist"] - (scale + 1) * shift) / scale
            for w in range(n):
                if w in (u, w_):
                    continue
                if D[u, w] < thr and D[w_, w] < thr:
                    return False
    return True


def test_triangle_rng_property_after_five_passes(orc):
    # test_propagate_engine.cpp:146-233: complete graph, 5 triangle passes,
    # RNG property holds, edges drop, rows stay sorted.
    v = grid_vecs()
    L = orc.Layer(v, orc.L2, cap=6)
    for u in range(12):
        ents = sorted([(w, float(orc.distance(orc.L2, v[u], v[w]))) for w in range(12) if w != u],
                      key=lambda t: (t[1], t[0]))
        L.set_row_raw(u, orc.nbrs(ents, new=True))
    for _ in range(5):
        L.triangle(1.0, 0.0)
    rows = [L.row(u) for u in range(12)]
    assert rng_property_ok(v, rows, orc.L2)
    assert sum(len(r) for r in rows) < 12 * 11
    for r in rows:
        assert len(r) <= 6
        keys = [(e["dist"], e["vid_status"] & 0x7FFFFFFF) for e in r]
        assert keys == sorted(keys)


@pytest.mark.parametrize("loops,expect", [(0, 0), (1, 0), (2, 1)])
def test_reverse_runs_between_build_rounds(orc, loops, expect):
    # test_graph_build_regressions.cpp:97-121: only edge 0->1, triu_iters=0,
    # prefill 0: the reverse edge 1->0 appears only with build_loops=2.
    n = 32
    v = np.zeros((n, 16), np.float32)
    v[1, 0] = 1.0
    L = orc.Layer(v, orc.L2, cap=12)
    L.fill_row(0, orc.nbrs([(1, 1.0)]))
    L.schedule(loops, 0, 1.0, 0.0)
    assert L.row(0)[0]["vid_status"] & 0x7FFFFFFF == 1
    r1 = L.row(1)
    assert len(r1) == expect
    if expect:
        assert r1[0]["vid_status"] & 0x7FFFFFFF == 0


def test_random_then_reverse_is_bidirectional(orc):
    # test_propagate_engine.cpp:320-415.
    v = grid_vecs()
    L = orc.Layer(v, orc.L2, cap=64)
    L.prefill(5, 0, 12, seed=1, stream=0)
    rows = [L.row(u) for u in range(12)]
    missing = sum(1 for u in range(12) for e in rows[u]
                  if not any((f["vid_status"] & 0x7FFFFFFF) == u for f in rows[e["vid_status"] & 0x7FFFFFFF]))
    assert missing > 0
    before = sum(len(r) for r in rows)
    L.reverse()
    rows = [L.row(u) for u in range(12)]
    assert sum(len(r) for r in rows) > before
    for u in range(12):
        for e in rows[u]:
            w = e["vid_status"] & 0x7FFFFFFF
            back = [f for f in rows[w] if (f["vid_status"] & 0x7FFFFFFF) == u]
            assert back and back[0]["dist"] == e["dist"]


def test_merge_is_sorted_unique_and_row_wins_ties(orc):
    # nbr_log_table.hpp:127-182: stable merge keeps the row's entry (its flag)
    # when a log carries the same (dist, vid); adjacent unique by vid.
    v = grid_vecs()
    L = orc.Layer(v, orc.L2, cap=64)
    L.set_row_raw(0, orc.nbrs([(1, 1.0), (4, 1.0)], new=False))
    L.set_row_raw(1, orc.nbrs([(0, 1.0)], new=False))
    L.set_row_raw(4, orc.nbrs([(0, 1.0), (5, 1.0)], new=False))
    L.reverse()
    r0 = L.row(0)
    assert [(e["vid_status"], e["dist"]) for e in r0] == [(1, 1.0), (4, 1.0)]  # old flags kept
    r5 = L.row(5)
    assert [(e["vid_status"], e["dist"]) for e in r5] == [(4 | orc.NBR_NEW, 1.0)]


def test_prefill_draws_are_uniform_and_in_range(orc):
    # Property of the counter-based draws (the reference's MKL stream is not
    # reproducible, random_seq.hpp:46-50): range, self-skip, rough uniformity.
    draws = np.array([orc.rand_draw(9, 0, r, j, 0, 100) for r in range(200) for j in range(50)])
    assert draws.min() >= 0 and draws.max() < 100
    counts = np.bincount(draws, minlength=100)
    assert counts.min() > 50 and counts.max() < 150
    v = np.zeros((100, 16), np.float32)
    v[:, 0] = np.arange(100)
    L = orc.Layer(v, orc.L2, cap=96)
    L.prefill(38, 0, 100, seed=9)
    for u in range(100):
        r = L.row(u)
        vids = [e["vid_status"] & 0x7FFFFFFF for e in r]
        assert u not in vids and len(r) <= 38 and len(set(vids)) == len(vids)
        keys = [(e["dist"], e["vid_status"] & 0x7FFFFFFF) for e in r]
        assert keys == sorted(keys)


def orc_rows(n, dim=32, seed=3):
    """Small integer-valued clustered rows (8 centres), float32."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 128, (8, dim))
    x = centres[rng.integers(0, 8, n)] + rng.normal(0, 12, (n, dim))
    return np.clip(np.rint(x), 0, 255).astype(np.float32)


def test_geometric_levels_and_compaction_trim(orc):
    """Seeded geometric level draw: fraction f^h of vertices at level >= h;
    the compactor keeps the first bucket from the top with >= 32 vertices
    as the top (index_traits.hpp:79), demotes the rest to it, and picks the
    entry from the top bucket."""
    n, f, ml = 20000, 0.25, 6
    lv = np.array([orc.build_level_of(5, v, ml, f) for v in range(n)])
    for h in range(1, 4):
        frac = np.mean(lv >= h)
        assert abs(frac - f ** h) < 4 * np.sqrt(f ** h / n) + 1e-3, (h, frac)
    assert all(orc.build_level_of(5, v, ml, 1.0) == ml for v in range(100))
    vecs = orc_rows(3000)
    g, _ = orc.build_full(vecs, metric=orc.L2, max_level_=5, build_loops=2, triu_iters=1, seed=9,
                          level_fraction=0.25, nthreads=4)
    src = np.array([orc.build_level_of(9, v, 5, 0.25) for v in range(3000)])
    counts = np.bincount(src, minlength=6)
    top = max(h for h in range(6) if counts[h] >= 32)
    assert g.top == top
    hl = np.array([g.highest(v) for v in range(3000)])
    assert np.array_equal(hl, np.minimum(src, top))
    assert hl[g.entry] == top
    # rows exist exactly on levels 0..highest, every id in range and no self loops
    for v in range(0, 3000, 37):
        for lvl in range(top + 1):
            ids = g.level_nbrs(v, lvl)
            if lvl > hl[v]:
                assert ids.size == 0
            else:
                assert ids.size > 0 and np.all(ids < 3000) and v not in ids
                assert np.all(hl[ids] >= lvl)  # neighbours live on the same level


def test_oracle_row_window_passes(orc):
    """next_range (refiner/propagate_engine.hpp:217-225) in the oracle: the
    full window equals next(); a window that leaves out the only populated
    row changes nothing; a window holding it equals next()."""
    import artea_amd as aa

    vecs = aa.generate(0, 3, 0, 400, 128)

    def layer_after(fn):
        L = orc.Layer(vecs, orc.L2, 16)
        L.prefill(6, 0, 400, seed=1, stream=0, nthreads=2)
        fn(L)
        return [L.row(i).tolist() for i in range(400)]

    full = layer_after(lambda L: L.triangle(1.1, 0.5, nthreads=2))
    assert layer_after(lambda L: L.triangle_range(1.1, 0.5, 0, 400, nthreads=2)) == full
    assert layer_after(lambda L: L.triangle_range(1.1, 0.5, 0, 10 ** 6, nthreads=2)) == full
    base = layer_after(lambda L: None)
    assert layer_after(lambda L: L.triangle_range(1.1, 0.5, 7, 7, nthreads=2)) == base
    assert layer_after(lambda L: L.reverse_range(0, 0, nthreads=2)) == base
    # reverse of rows [0, 200) only: every row >= 200 keeps its prefill
    # entries, and an edge u -> v with u >= 200 is reversed only if some row
    # below 200 also lists it
    rev = layer_after(lambda L: L.reverse_range(0, 200, nthreads=2))
    allrev = layer_after(lambda L: L.reverse(nthreads=2))
    assert rev != allrev and rev != base
    for u in range(200):
        for e in base[u]:
            v = e[0] & 0x7FFFFFFF
            assert any((x[0] & 0x7FFFFFFF) == u for x in rev[v]) or len(rev[v]) == 16

    # one populated row: a window without it leaves the layer as it was
    L1 = orc.Layer(vecs, orc.L2, 16)
    L1.set_row_raw(5, orc.nbrs([(u, orc.distance(orc.L2, vecs[5], vecs[u])) for u in (9, 40, 77, 123)], new=True))
    before = [L1.row(i).tolist() for i in range(400)]
    L1.triangle_range(1.0, 0.0, 6, 400, nthreads=2)
    assert [L1.row(i).tolist() for i in range(400)] == before

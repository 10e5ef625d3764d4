"""Routing loops of refine_layer (index/artea_graph/index_factory.hpp:378-403):
the warm-start RoutingUpdater (refiner/updaters/routing_updater.hpp:84-114)
and the PruningUpdater (refiner/updaters/pruning_updater.hpp:63-177).

CPU: the oracle against the reference's PruningUpdater known answers
(unit_tests/test_exact_artea_arc_candidates.cpp:382-399), against the exact
k nearest of every vertex when the queue holds the whole layer (the beam
then visits every reachable vertex), against a plain merge of the row with
those results, and the loop schedule's row bounds.  GPU: the product's fused
routing pass (artea_refine_routing_loops) and its pruning step
(artea_refine_prune) against the oracle, row for row."""
import numpy as np
import pytest

NEW = 0x80000000


def line(values, dim=16):
    v = np.zeros((len(values), dim), np.float32)
    v[:, 0] = values
    return v


def entries(pairs):
    import pyoracle

    return np.array([(vid, d) for vid, d in pairs], dtype=pyoracle.NBR_DTYPE)


def row_ids(L, r):
    return [int(x) & 0x7FFFFFFF for x in L.row(r)["vid_status"]]


# ----------------------------------------------------------------- oracle


def test_oracle_pruning_reference_kats(orc):
    # legacy(0, old) keeps both old entries (old pairs are skipped); with new
    # entries vertex 2 is pruned by vertex 1 (|1.25 - 1| < 1.25 / 2); the
    # shift is not used by the refiner's scaled rule; an empty row stays
    # empty; survivors are marked old.
    for metric in (orc.L2, orc.L2SQR):
        L = orc.Layer(line([0, 1, 1.25]), metric, 4)
        L.set_row_raw(0, entries([(1, 1.0), (2, 1.25)]))
        L.set_row_raw(1, entries([(1 | 0, 0.0)])[:0])
        L.pruning(2.0)
        assert row_ids(L, 0) == [1, 2] and L.row(1).size == 0
        L.set_row_raw(0, entries([(1 | NEW, 1.0), (2 | NEW, 1.25)]))
        L.pruning(2.0)
        r = L.row(0)
        assert row_ids(L, 0) == [1] and int(r["vid_status"][0]) & NEW == 0
        # one new and one old: the pair is checked
        L.set_row_raw(0, entries([(1, 1.0), (2 | NEW, 1.25)]))
        L.pruning(2.0)
        assert row_ids(L, 0) == [1]
        # strict boundary: dist == threshold keeps (1.25 / 1.25 = 1 scale)
        L2 = orc.Layer(line([0, 1, 2]), orc.L2, 4)
        L2.set_row_raw(0, entries([(1 | NEW, 1.0), (2 | NEW, 2.0)]))
        L2.pruning(2.0)  # |2 - 1| = 1 is not < 2 / 2
        assert row_ids(L2, 0) == [1, 2]


def random_layer(orc, n, D, cap, seed, metric=None):
    rng = np.random.default_rng(seed)
    vecs = rng.standard_normal((n, D)).astype(np.float32)
    metric = orc.L2 if metric is None else metric
    L = orc.Layer(vecs, metric, cap)
    L.prefill(int(cap * 0.6), 0, n, seed=seed)
    return vecs, L


def test_oracle_routing_full_queue_equals_exact_knn(orc):
    """queue >= n: the beam visits every reachable vertex, so the results are
    the exact topk nearest (by (distance, vid)) of each pivot; the logs are
    those minus the pivot and the row-full cut, merged into the row."""
    n, D, cap, topk = 120, 16, 12, 10
    vecs, L = random_layer(orc, n, D, cap, 3)
    before = [L.row(r).copy() for r in range(n)]
    dm = orc.distance_matrix(orc.L2, vecs, vecs)
    L.routing(topk, n)
    for p in range(n):
        order = sorted(range(n), key=lambda u: (dm[p, u], u))[:topk]
        row = before[p]
        full_thr = row["dist"][cap - 1] if len(row) >= cap else np.inf
        logs = [(u, dm[p, u]) for u in order if u != p and not (len(row) >= cap and full_thr <= dm[p, u])]
        merged = [(int(e["vid_status"]), float(e["dist"])) for e in row]
        merged += [(u | NEW, float(d)) for u, d in logs]
        merged.sort(key=lambda e: (e[1], e[0] & 0x7FFFFFFF, (e[0] & NEW) != 0))
        want, seen = [], set()
        for vs, d in merged:
            if want and (want[-1][0] & 0x7FFFFFFF) == (vs & 0x7FFFFFFF):
                continue
            want.append((vs, d))
        got = [(int(e["vid_status"]), float(e["dist"])) for e in L.row(p)]
        assert got == want, p


def grid_vecs():
    v = np.zeros((12, 16), np.float32)
    for i in range(12):
        v[i, 0], v[i, 1] = i % 4, i // 4
    return v


def grid_chain_rows(orc, v):
    """Rows of the chain 0-1-...-11 over the 3 x 4 grid (old entries)."""
    d = lambda a, b: float(np.sqrt(((v[a] - v[b]) ** 2).sum()))
    rows = []
    for u in range(12):
        nb = sorted([x for x in (u - 1, u + 1) if 0 <= x < 12], key=lambda x: (d(u, x), x))
        rows.append(np.array([(x, d(u, x)) for x in nb], dtype=orc.NBR_DTYPE))
    return rows


def grid_neighbours(u):
    r, c = u // 4, u % 4
    return sorted(rr * 4 + cc for rr, cc in ((r - 1, c), (r + 1, c), (r, c - 1), (r, c + 1))
                  if 0 <= rr < 3 and 0 <= cc < 4)


def test_oracle_routing_loop_grid_closed_form(orc):
    """A routing loop with topk = queue = n over a connected layer merges
    every other vertex into each row; the scale-1 pruning then leaves exactly
    the 3 x 4 grid's axis neighbours (a diagonal or farther vertex always has
    a retained axis neighbour strictly closer to it than the pivot is)."""
    v = grid_vecs()
    L = orc.Layer(v, orc.L2, 16)
    for u, row in enumerate(grid_chain_rows(orc, v)):
        L.set_row_raw(u, row)
    L.routing_loops(1, 12, 12, 1.0)
    for u in range(12):
        assert sorted(row_ids(L, u)) == grid_neighbours(u), u


def test_oracle_routing_loops_bounds_and_flags(orc):
    n, D, cap = 300, 16, 24
    _, L = random_layer(orc, n, D, cap, 5)
    L.schedule(2, 2, 1.1, 0.0)
    L.routing_loops(2, 16, 24, 1.1)
    lens = [len(L.row(r)) for r in range(n)]
    assert max(lens) <= cap + 16 and min(lens) >= 1
    for r in range(n):
        row = L.row(r)
        assert all(int(x) & NEW == 0 for x in row["vid_status"])  # pruning marks survivors old
        ids = row_ids(L, r)
        assert r not in ids and len(set(ids)) == len(ids)
        d = row["dist"]
        assert all(d[i] <= d[i + 1] for i in range(len(d) - 1))


# -------------------------------------------------------------------- GPU


def orc_rows(L, cap=None):
    c, e = L.export()
    out, pos = [], 0
    for n in c:
        m = n if cap is None else min(int(n), cap)
        out.append([(int(e[pos + k]["vid_status"]), int(np.float32(e[pos + k]["dist"]).view(np.uint32)))
                    for k in range(m)])
        pos += n
    return out


def dev_rows(eng):
    c, e = eng.download()
    out, pos = [], 0
    for n in c:
        out.append([(int(e[pos + k, 0]), int(e[pos + k, 1])) for k in range(n)])
        pos += n
    return out


def upload_from(eng, L):
    c, e = L.export()
    eng.upload(c, e.view(np.uint32).reshape(-1, 2) if e.size else np.zeros((0, 2), np.uint32))


def first_difference(a, b):
    for r, (x, y) in enumerate(zip(a, b)):
        if x != y:
            return r, x, y
    return None


@pytest.mark.gpu
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_gpu_pruning_matches_oracle(orc, metric):
    """next(PruningUpdater) on rows of new and old entries (a build loop,
    then reverse edges), every metric."""
    import artea_amd as aa
    from artea_amd.refine import PropagateEngine

    n, D, cap = 900, 32, 48
    vecs = aa.generate(0, 17 + metric, 0, n, D)
    idx = aa.DeviceIndex(vecs, metric)
    L = orc.Layer(vecs, metric, cap)
    L.prefill(int(cap * 0.5), 0, n, seed=3)
    L.schedule(1, 1, 1.1, 0.0)
    L.reverse()
    L.truncate()
    eng = PropagateEngine(idx, cap=cap, metric=metric)
    upload_from(eng, L)
    assert dev_rows(eng) == orc_rows(L)
    L.pruning(1.1)
    eng.next_pruning(1.1)
    got, want = dev_rows(eng), orc_rows(L)
    assert got == want, first_difference(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("metric,sparse,topk,queue,cap", [(3, False, 32, 48, 48), (0, False, 32, 48, 48),
                                                          (1, False, 24, 40, 48), (2, False, 32, 48, 48),
                                                          (3, True, 32, 48, 48), (3, False, 64, 128, 48),
                                                          (3, False, 32, 48, 12), (0, True, 16, 16, 8)])
def test_gpu_routing_loops_match_oracle(orc, metric, sparse, topk, queue, cap):
    """One and two routing loops after two build loops: identical rows (ids,
    distances, flags; the oracle's rows cut to cap as the writeback does).
    Small caps fill the rows (the row-full cut and the truncation)."""
    import artea_amd as aa
    from artea_amd.refine import PropagateEngine

    rng = np.random.default_rng(40 + metric)
    n, D = 3000, 48
    vecs = aa.generate(0, 29 + metric, 0, n, D)
    idx = aa.DeviceIndex(vecs, metric)
    l2g = np.sort(rng.choice(n, 1200, replace=False)).astype(np.uint32) if sparse else None
    L = orc.Layer(vecs, metric, cap, local_to_global=l2g)
    nl = L.num_rows
    L.prefill(int(cap * 0.4), 0, nl, seed=7)
    L.schedule(2, 2, 1.1, 0.3)
    base_c, base_e = L.export()
    for loops in (1, 2):
        L2 = orc.Layer(vecs, metric, cap, local_to_global=l2g)
        off = np.concatenate([[0], np.cumsum(base_c)]).astype(np.int64)
        for r in range(nl):
            L2.set_row_raw(r, base_e[off[r]:off[r + 1]])
        eng = PropagateEngine(idx, cap=cap, metric=metric, local_to_global=l2g)
        upload_from(eng, L2)
        L2.routing_loops(loops, topk, queue, 1.1)
        eng.routing_loops(loops, topk, queue, 1.1)
        st = eng.routing_stats()
        assert st["unresolved"] == 0 and st["expansions"] > 0
        got, want = dev_rows(eng), orc_rows(L2, cap)
        assert got == want, (loops, first_difference(got, want))


@pytest.mark.gpu
def test_gpu_routing_loop_grid_closed_form(orc):
    import artea_amd as aa
    from artea_amd.refine import PropagateEngine

    v = grid_vecs()
    eng = PropagateEngine(aa.DeviceIndex(v, 3), cap=16, metric=3)
    rows = grid_chain_rows(orc, v)
    counts = np.array([len(r) for r in rows], np.uint32)
    eng.upload(counts, np.concatenate(rows).view(np.uint32).reshape(-1, 2))
    eng.routing_loops(1, 12, 12, 1.0)
    got = dev_rows(eng)
    for u in range(12):
        assert sorted(e[0] & 0x7FFFFFFF for e in got[u]) == grid_neighbours(u), u
        assert all(e[0] & NEW == 0 for e in got[u])


@pytest.mark.gpu
def test_gpu_routing_argument_errors():
    import artea_amd as aa
    from artea_amd.refine import PropagateEngine

    vecs = aa.generate(0, 5, 0, 200, 16)
    eng = PropagateEngine(aa.DeviceIndex(vecs, 3), cap=16)
    with pytest.raises(aa.ArteaError, match=r"candidate_queue_size \(8\) must be >= topk \(16\)"):
        eng.routing_loops(1, 16, 8, 1.1)
    with pytest.raises(aa.ArteaError, match="topk must be in"):
        eng.routing_loops(1, 0, 8, 1.1)
    with pytest.raises(aa.ArteaError, match="scale must be positive"):
        eng.next_pruning(0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("fraction,loops", [(1.0, 1), (0.25, 2)])
def test_gpu_full_build_with_routing_loops(orc, fraction, loops):
    """build_full with routing loops after every level's build loops: the
    same compacted graph and HGRA bytes as the oracle's build."""
    import os
    import tempfile

    import artea_amd as aa
    from artea_amd.refine import build_full

    n = 3000
    vecs = aa.generate(0, 13, 0, n, 128)
    idx = aa.DeviceIndex(vecs, 0)
    build_full(idx, max_level=3, build_loops=3, triu_iters=2, seed=17, level_fraction=fraction, routing_loops=loops,
               routing_topk=32, routing_queue=48)
    g, _ = orc.build_full(vecs, metric=orc.L2, max_level_=3, build_loops=3, triu_iters=2, seed=17,
                          level_fraction=fraction, nthreads=8, routing_loops=loops, routing_topk=32, routing_queue=48)
    plain, _ = orc.build_full(vecs, metric=orc.L2, max_level_=3, build_loops=3, triu_iters=2, seed=17,
                              level_fraction=fraction, nthreads=8)
    with tempfile.TemporaryDirectory() as d:
        a, b, c = (os.path.join(d, x) for x in ("gpu.graph", "cpu.graph", "plain.graph"))
        idx.snapshot_hgra(a)
        g.snapshot(b)
        plain.snapshot(c)
        assert open(a, "rb").read() == open(b, "rb").read()
        assert open(b, "rb").read() != open(c, "rb").read()  # the loops changed the graph

"""R-net insertion (index/stacked_rgraph/index_factory.hpp:192-583) and the
artea_graph build on top of it (index/artea_graph/index_factory.hpp:135-410).

The product inserts in waves: the vertices of a wave search the graph as it
stood when the wave began and are applied in insertion order; a wave ends
after the first vertex that grows the hierarchy.  One-vertex waves are the
reference's serial insertion.

CPU: the oracle's HierarchicalPruningUpdater against closed forms on a line
(hierarchical_pruning_updater.hpp:96-138); with queues that hold the whole
data set, serial insertion's levels against the greedy hierarchical r-net
computed by brute force (vertex q stops below the first level h whose
nearest earlier member lies within radius_at(h)), independence from the apex
draws, the last vertex's rows against the pruned exact neighbour lists, and
row invariants.  GPU: levels, rows and the built graph identical to the
oracle's for the same waves."""
import numpy as np
import pytest

NEW = 0x80000000
MASK = 0x7FFFFFFF


def line(values, dim=16):
    v = np.zeros((len(values), dim), np.float32)
    v[:, 0] = values
    return v


def plane(n, seed, span=30.0, dim=16):
    rng = np.random.default_rng(seed)
    v = np.zeros((n, dim), np.float32)
    v[:, :2] = rng.uniform(0, span, (n, 2))
    return v


def test_oracle_hier_prune_closed_forms(orc):
    v = line([0, 1, 2, 3])
    cand = np.array([(1 | NEW, 1.0), (2 | NEW, 2.0), (3 | NEW, 3.0)], dtype=orc.NBR_DTYPE)
    ids = lambda e: [int(x) & MASK for x in e["vid_status"]]
    # scale 1: |2-1| = 1 < 2 and |3-1| = 2 < 3 prune both
    assert ids(orc.hier_prune(v, orc.L2, cand, 8, 1.0)) == [1]
    # scale 2: thresholds 1 and 1.5: 2 stays (|2-1| = 1 is not < 1), 3 goes (|3-2| = 1 < 1.5)
    out = orc.hier_prune(v, orc.L2, cand, 8, 2.0)
    assert ids(out) == [1, 2] and all(int(x) & NEW == 0 for x in out["vid_status"])
    # the degree cap stops the scan; the first entry is always kept
    assert ids(orc.hier_prune(v, orc.L2, cand, 1, 100.0)) == [1]
    # two old entries are never compared
    old = np.array([(1, 1.0), (2, 2.0), (3, 3.0)], dtype=orc.NBR_DTYPE)
    assert ids(orc.hier_prune(v, orc.L2, old, 8, 1.0)) == [1, 2, 3]
    assert orc.hier_prune(v, orc.L2, old[:0], 8, 1.0).size == 0


def greedy_levels(orc, v, max_level, params):
    """Serial r-net levels with exhaustive searches, by brute force."""
    n = len(v)
    dm = orc.distance_matrix(orc.L2, v, v)
    radii = [orc.rnet_radius(params.l0_min_distance, params.tau_k, params.beta, h) for h in range(max_level + 1)]
    lv = np.full(n, -1)
    top = None
    for q in range(n):
        if top is None:
            h, top = 1, 1
        else:
            h = min(top + 1, max_level)
            for k in range(1, top + 1):
                members = np.nonzero(lv[:q] >= k)[0]
                if len(members) and dm[q, members].min() <= radii[k]:
                    h = k - 1
                    break
            top = max(top, h)
        lv[q] = h
    return lv


@pytest.mark.parametrize("n,seed", [(900, 1), (3000, 2)])
def test_oracle_rnet_serial_levels_are_the_greedy_net(orc, n, seed):
    v = plane(n, seed)
    p = orc.rnet_params(search_qs=n, ul_select_qs=n, bl_select_qs=n, wave_max=1, seed=5)
    R = orc.RNet(v, orc.L2, p, nthreads=8)
    want = greedy_levels(orc, v, R.max_level, p)
    assert np.array_equal(R.highest(), want)
    assert R.waves == n
    # the apex draw does not matter when every search is exhaustive
    p2 = orc.rnet_params(search_qs=n, ul_select_qs=n, bl_select_qs=n, wave_max=1, seed=77)
    R2 = orc.RNet(v, orc.L2, p2, nthreads=8)
    assert np.array_equal(R2.highest(), want)
    for lvl in range(1, R.max_level + 1):
        c1, e1 = R.level_rows(lvl)
        c2, e2 = R2.level_rows(lvl)
        assert np.array_equal(c1, c2) and np.array_equal(e1, e2)


def test_oracle_rnet_last_vertex_rows_are_pruned_exact_lists(orc):
    """With exhaustive searches the last vertex's select phase sees every
    member of each of its levels; its rows are those lists, sorted by
    (distance, vid), pruned (nothing is inserted after it)."""
    n = 1200
    v = plane(n, 4)
    p = orc.rnet_params(search_qs=n, ul_select_qs=n, bl_select_qs=n, wave_max=1, insert_on_l0=True)
    R = orc.RNet(v, orc.L2, p, nthreads=8)
    lv = R.highest()
    dm = orc.distance_matrix(orc.L2, v, v)
    q = n - 1
    for lvl in range(0, lv[q] + 1):
        members = [u for u in range(q) if lv[u] >= lvl]
        members.sort(key=lambda u: (dm[q, u], u))
        cand = np.array([(u | NEW, dm[q, u]) for u in members], dtype=orc.NBR_DTYPE)
        want = orc.hier_prune(v, orc.L2, cand, p.bl_max if lvl == 0 else p.ul_max, p.scale)
        c, e = R.level_rows(lvl)
        off = int(c[:q].sum())
        got = e[off:off + c[q]]
        assert np.array_equal(got, want), lvl


@pytest.mark.parametrize("insert_on_l0", [False, True])
def test_oracle_rnet_waves_row_invariants(orc, insert_on_l0):
    n = 5000
    v = plane(n, 9, span=60.0)
    p = orc.rnet_params(insert_on_l0=insert_on_l0)
    R = orc.RNet(v, orc.L2, p, nthreads=8)
    lv = R.highest()
    assert lv.max() == R.top and R.waves < n // 4
    for lvl in range(0, R.max_level + 1):
        c, e = R.level_rows(lvl)
        cap = p.bl_max if lvl == 0 else p.ul_max
        if lvl == 0 and not insert_on_l0:
            assert c.sum() == 0
            continue
        assert c.max() <= cap
        off = np.concatenate([[0], np.cumsum(c)]).astype(np.int64)
        for u in range(0, n, 13):
            ids = [int(x) & MASK for x in e["vid_status"][off[u]:off[u + 1]]]
            assert u not in ids and len(set(ids)) == len(ids)
            assert all(lv[x] >= lvl for x in ids)
            if c[u]:
                assert lv[u] >= lvl


def test_oracle_wave_schedule():
    import pyoracle as orc

    assert [orc.rnet_wave(i, 8, 100) for i in (0, 9, 10, 15, 16, 80, 799, 800, 10 ** 6)] == \
        [1, 1, 1, 1, 2, 10, 99, 100, 100]


# -------------------------------------------------------------------- GPU


def oracle_params(orc, dev):
    """The same parameters for the oracle."""
    return orc.rnet_params(beta=dev.rnet_beta, tau_k=dev.tau_k, l0_min_distance=dev.l0_min_distance,
                           scale=dev.scale_coeffs, search_qs=dev.search_nn_qs, ul_select_qs=dev.ul_select_nbrs_qs,
                           bl_select_qs=dev.bl_select_nbrs_qs, ul_max=dev.ul_max_nbr_size, bl_max=dev.bl_max_nbr_size,
                           insert_on_l0=bool(dev.insert_on_l0), wave_div=dev.wave_div, wave_max=dev.wave_max,
                           seed=dev.seed)


RNET_CASES = [
    # (data, n, dim, metric, insert_on_l0, wave_max)
    ("plane", 600, 16, 3, True, 1),
    ("plane", 5000, 16, 3, False, 32768),
    ("plane", 5000, 16, 0, True, 64),
    ("kind0", 3000, 128, 0, False, 32768),
    ("kind2", 20000, 96, 0, False, 32768),
    ("kind2", 8000, 96, 2, True, 256),
]


def case_data(kind, n, dim):
    import artea_amd as aa

    if kind == "plane":
        return plane(n, 21, span=60.0, dim=dim)
    return aa.generate(0 if kind == "kind0" else 2, 13, 0, n, dim)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,dim,metric,on_l0,wave_max", RNET_CASES)
def test_gpu_rnet_matches_oracle(orc, kind, n, dim, metric, on_l0, wave_max):
    """Levels and every level's rows (ids, distances, flags, stored order)
    identical to the oracle's insertion with the same waves."""
    import artea_amd as aa
    from artea_amd.refine import RNet, rnet_params

    v = case_data(kind, n, dim)
    idx = aa.DeviceIndex(v, metric)
    bmetric = 3 if metric == 0 else metric
    p = rnet_params(insert_on_l0=on_l0, wave_max=wave_max, seed=3)
    g = RNet(idx, p)
    R = orc.RNet(v, bmetric, oracle_params(orc, p), nthreads=8)
    info = g.info()
    assert info["max_level"] == R.max_level and info["top"] == R.top and info["waves"] == R.waves
    assert info["unresolved"] == 0
    assert np.array_equal(g.highest(), R.highest())
    for lvl in range(R.max_level + 1):
        gc, ge = g.level_rows(lvl)
        oc, oe = R.level_rows(lvl)
        assert np.array_equal(gc, oc), lvl
        assert np.array_equal(ge, oe.view(np.uint32).reshape(-1, 2)), lvl


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,dim,metric", [("kind0", 3000, 128, 0), ("kind2", 20000, 96, 0)])
def test_gpu_full_build_rnet_matches_oracle(orc, tmp_path, kind, n, dim, metric):
    """r-net + refine_layer per level + compaction: the same HGRA bytes."""
    import artea_amd as aa
    from artea_amd.refine import build_full_rnet, rnet_params

    v = case_data(kind, n, dim)
    idx = aa.DeviceIndex(v, metric)
    p = rnet_params(seed=5)
    build_full_rnet(idx, p, build_loops=3, triu_iters=2)
    g, _ = orc.build_full_rnet(v, oracle_params(orc, p), metric=3 if metric == 0 else metric, build_loops=3,
                               triu_iters=2, nthreads=8)
    a, b = str(tmp_path / "gpu.graph"), str(tmp_path / "cpu.graph")
    idx.snapshot_hgra(a)
    g.snapshot(b)
    assert open(a, "rb").read() == open(b, "rb").read()

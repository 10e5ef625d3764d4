"""The reference's compaction fixture, restated
(unit_tests/compaction_baseline_fixture.hpp:45-163 and the checks of
unit_tests/test_hierarchical_graph_compactor.cpp:17-127, 133-143, 208-260):
seven topologies of apex counts per level, vertices assigned to their
buckets in descending vid order, one-coordinate vectors whose final top
bucket sums to zero with a unique vertex at the origin, and rows that mix
empty, partial and full lists at both degree limits.  The oracle's compactor
(CPU) and the product's (artea_index_compact_graph, GPU centroid) must give
the expected top, buckets, entry point, rows, CSR layout and HGRA bytes."""
import os

import numpy as np
import pytest

FLOOR = 32  # min_layer_cap (framework/type_traits/index_traits.hpp:79)
UPPER, BOTTOM = 3, 5
INVALID = 0xFFFFFFFF

CASES = [
    ("single_l0", [1], 0),
    ("l0_only", [5], 0),
    ("retained_layers", [5, FLOOR + 1, FLOOR + 1], 2),
    ("threshold_top", [5, FLOOR + 1, FLOOR], 2),
    ("multiple_trim", [5, FLOOR + 1, 2, 2], 1),
    ("trim_to_l0", [5, 3, 2, 2], 0),
    ("empty_middle_group", [5, 0, FLOOR + 1], 2),
]


def fixture(apex, expected_top):
    """Inputs and expectations of compaction_baseline::Fixture."""
    n = sum(apex)
    highest = np.zeros(n, np.uint32)
    order, source_buckets = [], []
    begin = 0
    for h, cnt in enumerate(apex):
        end = begin + cnt
        bucket = list(range(end - 1, begin - 1, -1))  # descending assignment
        order += bucket
        source_buckets.append(bucket)
        highest[begin:end] = h
        begin = end
    expected_buckets = [[] for _ in range(expected_top + 1)]
    for h, b in enumerate(source_buckets):
        expected_buckets[min(h, expected_top)] += b
    top = expected_buckets[-1]
    entry = top[0]
    x = np.arange(1, n + 1, dtype=np.float32)
    x[entry] = 0.0
    i = 1
    if len(top) % 2 == 0:
        x[top[1]], x[top[2]], x[top[3]] = 1.0, 2.0, -3.0
        i = 4
    coord = 4
    while i < len(top):
        x[top[i]], x[top[i + 1]] = coord, -coord
        i += 2
        coord += 1
    vecs = np.zeros((n, 16), np.float32)
    vecs[:, 0] = x
    levels = [[[] for _ in range(n)] for _ in range(len(apex))]
    for v in range(n):
        for h in range(int(highest[v]) + 1):
            cand = [o for o in range(n - 1, -1, -1) if o != v and highest[o] >= h]
            cap = BOTTOM if h == 0 else UPPER
            cnt = min((v + h) % (cap + 1), len(cand))
            levels[h][v] = [cand[(v + h + j) % len(cand)] for j in range(cnt)]
    return dict(n=n, highest=highest, order=order, vecs=vecs, levels=levels, buckets=expected_buckets, entry=entry,
                top=expected_top)


def check_topology(f, top, ul, bl, entry, level_of, local_of, level_nbrs, buckets_of):
    assert top == f["top"] and (ul, bl) == (UPPER, BOTTOM) and entry == f["entry"]
    for h, want in enumerate(f["buckets"]):
        assert buckets_of(h) == want, h
    for v in range(f["n"]):
        hv = min(int(f["highest"][v]), f["top"])
        assert level_of(v) == hv
        assert f["buckets"][hv][local_of(v)] == v
        for h in range(hv + 1):
            assert list(level_nbrs(v, h)) == f["levels"][h][v], (v, h)


def csr_layout(n, top, group_offsets, nbrs, vertex_info):
    """csr_layout (test_hierarchical_graph_compactor.cpp:70-121)."""
    seen = np.zeros(n, bool)
    g_off = 0
    for h in range(top + 1):
        off = group_offsets[h]
        cnt = (len(off) - 1) // (h + 1)
        assert len(off) == cnt * (h + 1) + 1 and off[0] == g_off, h
        members = [v for v in range(n) if vertex_info[v][0] == h]
        assert len(members) == cnt
        for v in sorted(members, key=lambda u: vertex_info[u][1]):
            assert not seen[v]
            seen[v] = True
            loc = vertex_info[v][1]
            for lvl in range(h, -1, -1):
                r = loc * (h + 1) + h - lvl
                assert off[r] == g_off and off[r + 1] >= off[r] and off[r + 1] <= len(nbrs)
                assert all(x < n for x in nbrs[off[r]:off[r + 1]])
                g_off = int(off[r + 1])
    assert seen.sum() == sum(1 for v in range(n) if vertex_info[v][0] != INVALID) and g_off == len(nbrs)


@pytest.mark.parametrize("name,apex,top", CASES)
def test_oracle_compaction_topology(orc, tmp_path, name, apex, top):
    f = fixture(apex, top)
    g = orc.Graph.compact(f["highest"], f["levels"], f["vecs"], orc.L2SQR, UPPER, BOTTOM, order=f["order"])
    hl, lc, groups, nbrs = g.csr()
    buckets = lambda h: [v for v in sorted(range(f["n"]), key=lambda u: lc[u]) if hl[v] == h]
    check_topology(f, g.top, g.ul_max, g.bl_max, g.entry, g.highest, g.local, g.level_nbrs, buckets)
    csr_layout(f["n"], g.top, groups, nbrs, np.stack([hl, lc], 1))
    # version-1 round trip preserves rows and bytes (:208-227)
    a, b = str(tmp_path / "a.graph"), str(tmp_path / "b.graph")
    g.snapshot(a)
    r = orc.Graph.restore(a)
    assert r.entry == g.entry
    for v in range(f["n"]):
        assert r.highest(v) == g.highest(v)
        for h in range(g.highest(v) + 1):
            assert list(r.level_nbrs(v, h)) == list(g.level_nbrs(v, h))
    r.snapshot(b)
    assert open(a, "rb").read() == open(b, "rb").read()


def test_oracle_compaction_empty_and_edgeless(orc):
    # :133-143: no vertices -> no entry and no occupied layer
    g = orc.Graph.compact([], [[]], np.zeros((0, 16), np.float32), orc.L2SQR, UPPER, BOTTOM)
    assert g.num_vertices == 0 and g.top == INVALID and g.entry == INVALID
    # :241-260: 65 vertices, the first 32 on level 1, no edges
    n = 65
    vecs = np.zeros((n, 16), np.float32)
    vecs[:, 0] = np.arange(n)
    highest = [1 if v < FLOOR else 0 for v in range(n)]
    g = orc.Graph.compact(highest, [[[]] * n, [[]] * n], vecs, orc.L2SQR, UPPER, BOTTOM)
    assert g.top == 1 and g.entry < FLOOR
    hl, lc, groups, nbrs = g.csr()
    assert len(nbrs) == 0
    csr_layout(n, 1, groups, nbrs, np.stack([hl, lc], 1))


@pytest.mark.gpu
@pytest.mark.parametrize("name,apex,top", CASES)
def test_device_compaction_topology(orc, tmp_path, name, apex, top):
    """The product's compactor (centroid and argmin on the GPU) on the same
    inputs: the expectations above, and the oracle's HGRA bytes."""
    import artea_amd as aa

    f = fixture(apex, top)
    idx = aa.DeviceIndex(f["vecs"], aa.DistanceMetrics.EUCLIDEAN_SQR)
    idx.compact_graph(f["highest"], f["levels"], UPPER, BOTTOM, order=f["order"])
    info = idx.info()
    ex = idx.export_graph()
    hl, lc = ex.vertex_info[:, 0], ex.vertex_info[:, 1]
    buckets = lambda h: [v for v in sorted(range(f["n"]), key=lambda u: lc[u]) if hl[v] == h]
    check_topology(f, info["top"], info["ul_max"], info["bl_max"], info["entry"], lambda v: int(hl[v]),
                   lambda v: int(lc[v]), ex.fetch_level_nbrs, buckets)
    csr_layout(f["n"], info["top"], ex.group_offsets, ex.nbrs, ex.vertex_info)
    g = orc.Graph.compact(f["highest"], f["levels"], f["vecs"], orc.L2SQR, UPPER, BOTTOM, order=f["order"])
    a, b = str(tmp_path / "gpu.graph"), str(tmp_path / "cpu.graph")
    idx.snapshot_hgra(a)
    g.snapshot(b)
    assert open(a, "rb").read() == open(b, "rb").read()


@pytest.mark.gpu
def test_device_compaction_edgeless_and_errors():
    import artea_amd as aa

    n = 65
    vecs = np.zeros((n, 16), np.float32)
    vecs[:, 0] = np.arange(n)
    idx = aa.DeviceIndex(vecs, 0)
    highest = [1 if v < FLOOR else 0 for v in range(n)]
    idx.compact_graph(highest, [[[]] * n, [[]] * n], UPPER, BOTTOM)
    info = idx.info()
    assert info["top"] == 1 and info["entry"] < FLOOR and idx.export_graph().nbrs.size == 0
    with pytest.raises(aa.ArteaError, match="highest level"):
        idx.compact_graph([3] * n, [[[]] * n, [[]] * n], UPPER, BOTTOM)
    with pytest.raises(aa.ArteaError, match="permutation"):
        idx.compact_graph(highest, [[[]] * n, [[]] * n], UPPER, BOTTOM, order=[0] * n)
    with pytest.raises(aa.ArteaError, match="invalid neighbor|out of range"):
        idx.compact_graph(highest, [[[n + 3]] + [[]] * (n - 1), [[]] * n], UPPER, BOTTOM)

"""The product's HGRA v1 restore (artea_amd/csrc/hgra.cpp) on corrupt and
edge-case files, through the C-ABI: every rejection path of
HierarchicalGraphFileManager::restore
(persistence/hierarchical_graph_file_manager.hpp:152-260) with its message,
the empty graph (:196-206), and the round trip.  artea_graph_restore is
host code: these run without a GPU; the GPU case repeats them through
artea_index_set_graph_hgra."""
import ctypes as C
import struct

import numpy as np
import pytest

MAGIC = 0x48475241

# n = 4, top = 1, ul = bl = 4, entry 0.  vid 0: h 1, level 1 [1], level 0
# [1, 2, 3]; vid 1: h 1, level 1 [0], level 0 [0, 2]; vid 2: h 0, level 0
# [0, 1]; vid 3 unassigned.
HEADER = [MAGIC, 1, 4, 1, 4, 4, 0]
BODY = [1, 1, 1, 3, 1, 2, 3,
        1, 1, 0, 2, 0, 2,
        0, 2, 0, 1,
        0xFFFFFFFF]


def words(ws):
    return struct.pack("<%dI" % len(ws), *ws)


def good_file():
    return words(HEADER + BODY)


def restore(path):
    from artea_amd._lib import lib

    h = C.c_void_p()
    rc = lib().artea_graph_restore(str(path).encode(), C.byref(h))
    msg = lib().artea_last_error().decode()
    return rc, msg, h


def graph_levels(h, n, top):
    from artea_amd._lib import lib

    out = {}
    for v in range(n):
        for lvl in range(top + 1):
            cnt = C.c_uint32()
            p = lib().artea_graph_level_nbrs(h, v, lvl, C.byref(cnt))
            out[(v, lvl)] = [p[i] for i in range(cnt.value)]
    return out


def corrupt_cases():
    hdr_body = HEADER + BODY
    cases = {
        "header_short": (words(HEADER[:5]), "header read failed"),
        "magic": (words([0] + hdr_body[1:]), "bad magic 0"),
        "version": (words(HEADER[:1] + [2] + HEADER[2:] + BODY), "unsupported version 2"),
        "top_invalid": (words(HEADER[:3] + [0xFFFFFFFF] + HEADER[4:] + BODY), "invalid header"),
        "entry_range": (words(HEADER[:6] + [4] + BODY), "invalid header"),
        "top_device_bound": (words(HEADER[:3] + [70] + HEADER[4:] + BODY), "exceeds the supported maximum of 63"),
        "highest_above_top": (words(HEADER + [5] + BODY[1:]), "vid 0 has highest_level_id 5 > top_level_id 1"),
        "trunc_vid": (words(HEADER + BODY[:7]), "truncated at vid 1 in"),
        "trunc_level": (words(HEADER + BODY[:1]), "truncated at vid 0 level 1 in"),
        "capacity": (words(HEADER + BODY[:3] + [9] + BODY[4:]), "vid 0 level 0 count 9 > capacity 4"),
        "trunc_nbrs": (words(HEADER + BODY[:5]), "truncated reading nbrs of vid 0 level 0"),
        "neighbor": (words(HEADER + BODY[:4] + [77] + BODY[5:]), "invalid neighbor 77"),
        # a corrupt vertex count: no table of 2^32 entries is allocated, the
        # body runs out first
        "huge_n": (words(HEADER[:2] + [0xFFFFFFF0] + HEADER[3:] + BODY), "truncated at vid 4 in"),
    }
    return cases


@pytest.mark.parametrize("case", sorted(corrupt_cases()))
def test_restore_rejects(tmp_path, case):
    data, msg = corrupt_cases()[case]
    p = tmp_path / (case + ".graph")
    p.write_bytes(data)
    rc, err, _ = restore(p)
    assert rc == 2, (case, rc, err)  # ARTEA_EIO
    assert "HierarchicalGraphFileManager::restore" in err and msg in err, err
    assert str(p) in err


def test_restore_missing_file(tmp_path):
    rc, err, _ = restore(tmp_path / "nope.graph")
    assert rc == 2 and "failed to open" in err


def test_restore_good_file_and_round_trip(tmp_path):
    from artea_amd._lib import lib

    p = tmp_path / "g.graph"
    p.write_bytes(good_file())
    rc, err, h = restore(p)
    assert rc == 0, err
    try:
        vals = [C.c_uint32() for _ in range(5)]
        nn = C.c_uint64()
        assert lib().artea_graph_info(h, *[C.byref(x) for x in vals], C.byref(nn)) == 0
        assert [x.value for x in vals] == [4, 1, 4, 4, 0] and nn.value == 9
        lv = graph_levels(h, 4, 1)
        assert lv[(0, 1)] == [1] and lv[(0, 0)] == [1, 2, 3]
        assert lv[(1, 1)] == [0] and lv[(1, 0)] == [0, 2]
        assert lv[(2, 0)] == [0, 1] and lv[(2, 1)] == [] and lv[(3, 0)] == []
        q = tmp_path / "g2.graph"
        assert lib().artea_graph_snapshot(h, str(q).encode()) == 0
        assert q.read_bytes() == p.read_bytes()
    finally:
        lib().artea_graph_destroy(h)


def test_restore_empty_graph(tmp_path):
    # file_manager.hpp:196-206: n == 0 restores an empty graph with top + 1
    # empty groups and the header's entry point.
    from artea_amd._lib import lib

    p = tmp_path / "empty.graph"
    p.write_bytes(words([MAGIC, 1, 0, 3, 32, 64, 0xFFFFFFFF]))
    rc, err, h = restore(p)
    assert rc == 0, err
    try:
        vals = [C.c_uint32() for _ in range(5)]
        nn = C.c_uint64()
        assert lib().artea_graph_info(h, *[C.byref(x) for x in vals], C.byref(nn)) == 0
        n, top, ul, bl, entry = [x.value for x in vals]
        assert (n, ul, bl, entry, nn.value) == (0, 32, 64, 0xFFFFFFFF, 0)
        for g in range(4):
            cnt = C.c_uint64()
            off = lib().artea_graph_group_offsets(h, g, C.byref(cnt))
            assert cnt.value == 0 and off[0] == 0
    finally:
        lib().artea_graph_destroy(h)


@pytest.mark.gpu
def test_index_restore_rejects_and_accepts(tmp_path):
    """The same files through artea_index_set_graph_hgra (device layout)."""
    import artea_amd as aa

    vecs = np.zeros((4, 16), np.float32)
    vecs[:, 0] = np.arange(4)
    idx = aa.DeviceIndex(vecs, 0)
    for case, (data, msg) in sorted(corrupt_cases().items()):
        p = tmp_path / (case + ".graph")
        p.write_bytes(data)
        with pytest.raises(aa.ArteaError, match=msg):
            idx.load_hgra(str(p))
    p = tmp_path / "g.graph"
    p.write_bytes(good_file())
    idx.load_hgra(str(p))
    q = tmp_path / "g2.graph"
    idx.snapshot_hgra(str(q))
    assert q.read_bytes() == p.read_bytes()
    v, d = aa.HierarchicalGraphRouter(idx, None, 2, 4).batch_query(vecs[2:3])
    assert list(v[0]) == [2, 1] and list(d[0]) == [0.0, 1.0]


def test_restore_rejects_the_reference_one_vertex_corruptions(tmp_path):
    # test_hierarchical_graph_compactor.cpp:293-311: one L0 vertex with a
    # self edge; seven single-word corruptions and every truncation fail.
    valid = [MAGIC, 1, 1, 0, 3, 5, 0, 0, 1, 0]
    p = tmp_path / "valid.graph"
    p.write_bytes(words(valid))
    rc, err, h = restore(p)
    assert rc == 0, err
    from artea_amd._lib import lib

    lib().artea_graph_destroy(h)
    bad = []
    for word, repl in ((0, 0), (1, 2), (3, 0xFFFFFFFF), (6, 1), (7, 1), (8, 6), (9, 0xFFFFFFFF)):
        w = list(valid)
        w[word] = repl
        bad.append(w)
    bad += [valid[:k] for k in range(len(valid))]
    for i, w in enumerate(bad):
        q = tmp_path / ("bad%d.graph" % i)
        q.write_bytes(words(w))
        rc, err, _ = restore(q)
        assert rc == 2 and "HierarchicalGraphFileManager::restore" in err, (i, w, rc, err)


def test_unassigned_records_round_trip(tmp_path):
    # :276-291: vertices with no level restore as unassigned, no entry.
    from artea_amd._lib import lib

    for n in (0, 3):
        p = tmp_path / ("unassigned%d.graph" % n)
        p.write_bytes(words([MAGIC, 1, n, 0, 3, 5, 0xFFFFFFFF] + [0xFFFFFFFF] * n))
        rc, err, h = restore(p)
        assert rc == 0, err
        try:
            vals = [C.c_uint32() for _ in range(5)]
            nn = C.c_uint64()
            assert lib().artea_graph_info(h, *[C.byref(x) for x in vals], C.byref(nn)) == 0
            assert vals[0].value == n and vals[4].value == 0xFFFFFFFF and nn.value == 0
            vi = lib().artea_graph_vertex_info(h)
            for v in range(n):
                assert vi[2 * v] == 0xFFFFFFFF and vi[2 * v + 1] == 0xFFFFFFFF
            q = tmp_path / ("again%d.graph" % n)
            assert lib().artea_graph_snapshot(h, str(q).encode()) == 0
            assert q.read_bytes() == p.read_bytes()
        finally:
            lib().artea_graph_destroy(h)

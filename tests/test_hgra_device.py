"""HGRA v1 on the device (artea_amd/csrc/hgra_device.hip) against the
product's host reader and writer (artea_amd/csrc/hgra.cpp), which
test_hgra_product.py pins to HierarchicalGraphFileManager::restore's
rejection paths and messages (persistence/hierarchical_graph_file_manager.hpp
:152-324) and to the byte layout of snapshot (:88-145).

Bar: the device writer's file is byte-identical to the host writer's; the
device loader accepts exactly the files the host reader accepts, into the
same graph (its snapshot and its host CSR equal the file's), and rejects
every other file with the host reader's return code and text.  The
corrupted files are seeded single-word replacements and truncations of a
built graph's file; the segmented parse runs with its defaults and with
small segments whose first records are guessed badly on purpose (trial 0),
so the walk's re-parse path carries most of the records."""
import ctypes as C
import struct

import numpy as np
import pytest

import artea_amd as aa

pytestmark = pytest.mark.gpu

SETTINGS = [(0, 0), (16, 0), (16, 1), (64, 4)]


def lib():
    from artea_amd._lib import lib as L

    return L()


def host_restore(path):
    h = C.c_void_p()
    rc = lib().artea_graph_restore(str(path).encode(), C.byref(h))
    msg = lib().artea_last_error().decode()
    return rc, msg, h


def host_snapshot_bytes(h, path):
    assert lib().artea_graph_snapshot(h, str(path).encode()) == 0
    return path.read_bytes()


def device_load(idx, path, seg, trial):
    rc = lib().artea_index_set_graph_hgra_segmented(idx.h, str(path).encode(), seg, trial)
    return rc, lib().artea_last_error().decode()


def device_graph_bytes(idx, path):
    """Snapshot of the index's host CSR (built from the device layout when
    the graph came from the device loader), written by the host writer."""
    h = C.c_void_p()
    assert lib().artea_index_get_graph(idx.h, C.byref(h)) == 0
    try:
        return host_snapshot_bytes(h, path)
    finally:
        lib().artea_graph_destroy(h)


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    d = tmp_path_factory.mktemp("hgra_dev")
    vecs = aa.generate(2, 11, 0, 1500, 16)
    idx = aa.DeviceIndex(vecs, 0)
    aa.build_full(idx, max_level=3, build_loops=2, triu_iters=2, seed=3, level_fraction=0.25)
    dev = d / "device.graph"
    idx.snapshot_hgra(str(dev))
    return d, vecs, idx, dev.read_bytes()


def test_device_writer_equals_host_writer(built):
    d, vecs, idx, data = built
    assert data == device_graph_bytes(idx, d / "host.graph")
    n, top = struct.unpack("<2I", data[8:16])
    assert n == 1500 and top >= 1


@pytest.mark.parametrize("seg,trial", SETTINGS)
def test_device_loader_round_trip(built, seg, trial):
    d, vecs, idx, data = built
    p = d / "in.graph"
    p.write_bytes(data)
    idx2 = aa.DeviceIndex(vecs, 0)
    rc, msg = device_load(idx2, p, seg, trial)
    assert rc == 0, msg
    q = d / ("again_%d_%d.graph" % (seg, trial))
    idx2.snapshot_hgra(str(q))
    assert q.read_bytes() == data
    assert device_graph_bytes(idx2, d / "csr.graph") == data
    qs = aa.generate(2, 12, 1, 64, 16)
    for k, L in [(1, 1), (10, 40), (20, 120)]:
        v1, d1 = aa.HierarchicalGraphRouter(idx, None, k, L).batch_query(qs)
        v2, d2 = aa.HierarchicalGraphRouter(idx2, None, k, L).batch_query(qs)
        assert np.array_equal(v1, v2) and np.array_equal(d1.view(np.uint32), d2.view(np.uint32))


def corruptions(data, seed, count):
    """(name, bytes): single body words replaced, header fields replaced, and
    truncations, from a seeded stream."""
    rng = np.random.default_rng(seed)
    words = np.frombuffer(data, np.uint32).copy()
    n, top, ul, bl = (int(x) for x in words[2:6])
    nw = len(words)
    out = []
    for i in range(count):
        kind = i % 4
        if kind == 3:
            cut = int(rng.integers(28, len(data)))
            out.append(("cut%d" % cut, data[:cut]))
            continue
        w = words.copy()
        at = int(rng.integers(7, nw))
        choice = int(rng.integers(0, 6))
        val = [0xFFFFFFFF, int(rng.integers(0, top + 3)), int(rng.integers(0, max(ul, bl) + 3)),
               n + int(rng.integers(0, 3)), int(rng.integers(0, 2**32)), int(rng.integers(0, n))][choice]
        if kind == 2:  # a header field: n, ul_max, bl_max or entry
            at = int(rng.choice([2, 4, 5, 6]))
            val = [n - 1 - int(rng.integers(0, 5)), max(0, ul - 1), max(0, bl - 2), n][[2, 4, 5, 6].index(at)]
        w[at] = np.uint32(val)
        out.append(("w%d=%d" % (at, val), w.tobytes()))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_device_loader_matches_host_reader_on_corrupt_files(built, seed):
    d, vecs, idx, data = built
    idx2 = aa.DeviceIndex(vecs, 0)
    accepted = rejected = 0
    for name, blob in corruptions(data, seed, 40):
        p = d / "corrupt.graph"
        p.write_bytes(blob)
        hrc, hmsg, h = host_restore(p)
        try:
            hn = None
            if hrc == 0:
                vals = [C.c_uint32() for _ in range(5)]
                nn = C.c_uint64()
                assert lib().artea_graph_info(h, *[C.byref(x) for x in vals], C.byref(nn)) == 0
                hn = vals[0].value
                want = host_snapshot_bytes(h, d / "host_ok.graph")
        finally:
            if hrc == 0:
                lib().artea_graph_destroy(h)
        for seg, trial in SETTINGS:
            rc, msg = device_load(idx2, p, seg, trial)
            where = (name, seg, trial, hrc, hmsg, rc, msg)
            if hrc != 0:
                assert rc == hrc and msg == hmsg, where
            elif hn != len(vecs):
                assert rc == 1 and "graph has %d vertices" % hn in msg, where
            else:
                assert rc == 0, where
                assert device_graph_bytes(idx2, d / "dev_ok.graph") == want, where
        if hrc == 0:
            accepted += 1
        else:
            rejected += 1
    assert rejected >= 20


def test_device_loader_reference_corrupt_cases(tmp_path):
    """The host reader's named rejection cases (test_hgra_product.py) through
    the device loader at every setting: same code and text."""
    from test_hgra_product import corrupt_cases, good_file

    vecs = np.zeros((4, 16), np.float32)
    vecs[:, 0] = np.arange(4)
    idx = aa.DeviceIndex(vecs, 0)
    for case, (blob, _) in sorted(corrupt_cases().items()):
        p = tmp_path / (case + ".graph")
        p.write_bytes(blob)
        hrc, hmsg, _ = host_restore(p)
        for seg, trial in SETTINGS + [(1, 0), (2, 0), (3, 1)]:
            rc, msg = device_load(idx, p, seg, trial)
            assert (rc, msg) == (hrc, hmsg), (case, seg, trial)
    p = tmp_path / "good.graph"
    p.write_bytes(good_file())
    for seg, trial in SETTINGS + [(1, 0), (2, 0), (3, 1)]:
        rc, msg = device_load(idx, p, seg, trial)
        assert rc == 0, msg
        q = tmp_path / "good2.graph"
        idx.snapshot_hgra(str(q))
        assert q.read_bytes() == p.read_bytes()

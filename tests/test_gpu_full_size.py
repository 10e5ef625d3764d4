"""The search path at its full single-GPU size: 1,000,000 x 128 L2 (the
integer-valued 256-component mixture), the whole hierarchy built on the GPU,
2,000 queries of mixed topk {10, 100} at queue 200.

Size-independent properties checked at that size:
* batched (every kernel slot the autotuner times, and its pick) = plain
  router, ids and distances bit for bit; one-byte rows = FP32 rows;
* every returned distance is the exact integer squared distance (the data
  makes every FP32 partial sum exact), recomputed on the host in int64;
* each row ascends by (distance, vid) with distinct ids below n;
* the BF16 MFMA brute force equals the FP64 one, and recall@100 against it
  reaches 0.95 at queue 200;
* an HGRA snapshot restored into a new index answers identically, and the
  oracle (restated reference router) on that same file returns the same
  lists for a sample of the queries.
"""
import os

import numpy as np
import pytest

import artea_amd as aa

pytestmark = pytest.mark.gpu

N, D, NQ, QUEUE = 1_000_000, 128, 2000, 200


@pytest.fixture(scope="module")
def full():
    from artea_amd.device import DeviceBuffer, Stream

    seed = 20261204
    base = aa.generate(0, seed, 0, N, D)
    qs = aa.generate(0, seed, 1, NQ, D)
    topk = np.random.default_rng(seed).choice(np.array([10, 100], np.uint32), NQ)
    idx = aa.DeviceIndex(base, 0)
    assert idx.u8_rows
    aa.build_full(idx, ul_max=32, bl_max=64, build_loops=15, triu_iters=4, prefill=0.4, scale=1.10, shift=1.5,
                  l0_min_distance=1.0, seed=seed)
    s = Stream()
    q_d = DeviceBuffer.from_numpy(qs)
    tk_d = DeviceBuffer.from_numpy(topk)
    ql_d = DeviceBuffer.from_numpy(np.full(NQ, QUEUE, np.uint32))
    return dict(base=base, qs=qs, topk=topk, idx=idx, s=s, q_d=q_d, tk_d=tk_d, ql_d=ql_d)


def batched(f, idx, variant=None):
    from artea_amd.device import DeviceBuffer

    ov = DeviceBuffer.empty((NQ, 100), np.uint32)
    od = DeviceBuffer.empty((NQ, 100), np.float32)
    br = aa.BatchedRouter(idx, 100, QUEUE)
    if variant is not None:
        br.set_variant(variant)
    br.query_device(f["q_d"].ptr, NQ, f["tk_d"].ptr, f["ql_d"].ptr, ov.ptr, od.ptr, stream=f["s"].h)
    f["s"].synchronize()
    return ov.numpy(), od.numpy(), br


def plain(f, idx):
    out_v = np.zeros((NQ, 100), np.uint32)
    out_d = np.zeros((NQ, 100), np.float32)
    for k in (10, 100):
        sel = np.where(f["topk"] == k)[0]
        v, d = aa.HierarchicalGraphRouter(idx, None, k, QUEUE).batch_query(f["qs"][sel])
        out_v[sel, :k], out_d[sel, :k] = v, d
    return out_v, out_d


def same(f, a_v, a_d, b_v, b_d):
    for i in range(NQ):
        k = f["topk"][i]
        if not (np.array_equal(a_v[i, :k], b_v[i, :k]) and
                np.array_equal(a_d[i, :k].view(np.uint32), b_d[i, :k].view(np.uint32))):
            return i
    return -1


def test_full_size_batched_plain_and_rows_agree(full):
    f = full
    pv, pd = plain(f, f["idx"])
    bv, bd, br = batched(f, f["idx"])
    assert same(f, bv, bd, pv, pd) == -1
    from artea_amd.device import DeviceBuffer

    ov = DeviceBuffer.empty((NQ, 100), np.uint32)
    od = DeviceBuffer.empty((NQ, 100), np.float32)
    chosen, ms = br.autotune(f["q_d"].ptr, NQ, f["tk_d"].ptr, f["ql_d"].ptr, ov.ptr, od.ptr)
    assert len(ms) == 8 and 0 <= chosen < 8
    assert same(f, ov.numpy(), od.numpy(), pv, pd) == -1
    # The FP32 rows (slot 1 reads them even while the one-byte copy exists).
    fv, fd, _ = batched(f, f["idx"], variant=1)
    assert same(f, fv, fd, pv, pd) == -1
    f["result"] = (pv, pd)


def test_full_size_distances_exact_sorted_distinct(full):
    f = full
    pv, pd = full.get("result") or plain(f, f["idx"])
    base = f["base"].astype(np.int64)
    q = f["qs"].astype(np.int64)
    for i in range(0, NQ, 7):
        k = int(f["topk"][i])
        ids, ds = pv[i, :k].astype(np.int64), pd[i, :k]
        assert np.all(ids < N) and len(np.unique(ids)) == k
        exact = ((base[ids] - q[i]) ** 2).sum(1)
        assert np.array_equal(ds.astype(np.int64), exact) and np.all(ds == exact.astype(np.float32))
        d0, d1, i0, i1 = ds[:-1], ds[1:], ids[:-1], ids[1:]
        assert np.all((d0 < d1) | ((d0 == d1) & (i0 < i1)))


def test_full_size_recall_and_ground_truth(full):
    from artea_amd.device import DeviceBuffer, bruteforce_knn

    f = full
    pv, _ = full.get("result") or plain(f, f["idx"])
    view = DeviceBuffer.borrow(f["idx"].device_vectors(), f["base"].shape, np.float32)
    gi, gd = bruteforce_knn(aa.DistanceMetrics.EUCLIDEAN_SQR, view, N, f["q_d"], NQ, D, 100, f["s"])
    mi, md = bruteforce_knn(aa.DistanceMetrics.EUCLIDEAN_SQR, view, N, f["q_d"], NQ, D, 100, f["s"], method="mfma")
    assert np.array_equal(mi, gi) and np.array_equal(md, gd)
    sel = np.where(f["topk"] == 100)[0]
    hits = sum(len(set(pv[i, :100].tolist()) & set(gi[i, :100].tolist())) for i in sel)
    assert hits / (100.0 * len(sel)) >= 0.95


def test_full_size_snapshot_round_trip_and_oracle(full, orc, tmp_path):
    f = full
    pv, pd = full.get("result") or plain(f, f["idx"])
    path = os.path.join(str(tmp_path), "c1.graph")
    f["idx"].snapshot_hgra(path)
    idx2 = aa.DeviceIndex(f["base"], 0)
    idx2.load_hgra(path)
    rv, rd, _ = batched(f, idx2)
    assert same(f, rv, rd, pv, pd) == -1
    g = orc.Graph.restore(path)
    sample = np.where(f["topk"] == 100)[0][:60]
    ov, od, _ = orc.batch_query(g, f["base"], orc.L2SQR, f["qs"][sample], 100, QUEUE, nthreads=16)
    assert np.array_equal(ov, pv[sample, :100]) and np.array_equal(od, pd[sample, :100])


def test_full_size_cosine(orc, tmp_path):
    """1,000,000 x 100 (padded 112) unit-norm cosine, sparse hierarchy
    (level fraction 1/32), 1,000 queries, topk 10 at queue 400: batched =
    plain bit for bit, each distance within 1e-6 of the FP64 cosine of the
    same pair, rows ascending, recall@10 >= 0.95 against the FP64 ground
    truth, and the oracle on the index's HGRA file equal on 40 queries."""
    from artea_amd.device import DeviceBuffer, Stream, bruteforce_knn

    seed, nq, k, L = 20261205, 1000, 10, 400
    base = aa.generate(1, seed, 0, N, 100)
    qs = aa.generate(1, seed, 1, nq, 100)
    dp = base.shape[1]
    idx = aa.DeviceIndex(base, aa.DistanceMetrics.COSINE)
    assert not idx.u8_rows
    aa.build_full(idx, ul_max=32, bl_max=64, build_loops=15, triu_iters=4, prefill=0.4, scale=1.10, shift=1.5,
                  l0_min_distance=1.0, seed=seed, level_fraction=1.0 / 32)
    pv, pd = aa.HierarchicalGraphRouter(idx, None, k, L).batch_query(qs)
    s = Stream()
    q_d = DeviceBuffer.from_numpy(qs)
    tk = DeviceBuffer.from_numpy(np.full(nq, k, np.uint32))
    tl = DeviceBuffer.from_numpy(np.full(nq, L, np.uint32))
    ov = DeviceBuffer.empty((nq, k), np.uint32)
    od = DeviceBuffer.empty((nq, k), np.float32)
    br = aa.BatchedRouter(idx, k, L)
    br.autotune(q_d.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr)
    br.query_device(q_d.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr, stream=s.h)
    s.synchronize()
    assert np.array_equal(ov.numpy(), pv) and np.array_equal(od.numpy().view(np.uint32), pd.view(np.uint32))
    x = base[pv[::10].astype(np.int64)].astype(np.float64)
    q = qs[::10].astype(np.float64)
    cos64 = 1.0 - (x * q[:, None, :]).sum(-1) / np.sqrt((x * x).sum(-1) * (q * q).sum(-1)[:, None])
    assert np.max(np.abs(pd[::10].astype(np.float64) - cos64)) < 1e-6
    assert np.all((pd[:, :-1] < pd[:, 1:]) | ((pd[:, :-1] == pd[:, 1:]) & (pv[:, :-1] < pv[:, 1:])))
    view = DeviceBuffer.borrow(idx.device_vectors(), base.shape, np.float32)
    gi, _ = bruteforce_knn(aa.DistanceMetrics.COSINE, view, N, q_d, nq, dp, k, s, method="mfma")
    hits = sum(len(set(pv[i].tolist()) & set(gi[i].tolist())) for i in range(nq))
    assert hits / float(nq * k) >= 0.95
    path = os.path.join(str(tmp_path), "c2.graph")
    idx.snapshot_hgra(path)
    g = orc.Graph.restore(path)
    cv, cd, _ = orc.batch_query(g, base, orc.COSINE, qs[:40], k, L, nthreads=16)
    assert np.array_equal(cv, pv[:40]) and np.array_equal(cd, pd[:40])

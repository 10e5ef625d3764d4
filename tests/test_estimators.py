"""Recall@K, ADR@1 and the dataset prober (utils/recall_estimator.hpp,
utils/adr_estimator.hpp, utils/dataset_prober.hpp).

CPU: the oracle against the reference's recall known answers
(unit_tests/test_recall_estimator.cpp:50-212), a set-based restatement of
the intersection, closed forms on one-coordinate data for ADR, the k-NN
table and the farthest distance, and a full sort of the distance matrix for
the heap selection; the argument checks of the C-ABI entry points that fail
before any device work.  GPU: the product's kernels against the oracle, bit
for bit (the distance chains are dist_func_t's; ADR sums in query order)."""
import ctypes as C

import numpy as np
import pytest

INV = 0xFFFFFFFF


def recall_kats():
    """The reference's seven known answers: (pred, gt, k, expected)."""
    out = []
    nq, k = 100, 10
    base = np.arange(nq * k, dtype=np.uint32).reshape(nq, k)
    out.append(("perfect", base, base, k, 1.0))
    out.append(("zero", base, base + 10000, k, 0.0))
    half = base.copy()
    half[:, k // 2:] += 10000
    out.append(("partial", half, base, k, 0.5))
    nq, k = 10, 5
    b5 = np.arange(nq * k, dtype=np.uint32).reshape(nq, k)
    out.append(("unordered", b5[:, ::-1].copy(), b5, k, 1.0))
    nq, k = 10, 10
    b10 = np.arange(nq * k, dtype=np.uint32).reshape(nq, k)
    dup = np.repeat(b10[:, :1], k, axis=1)
    out.append(("duplicates", dup, b10, k, 0.1))
    out.append(("single", np.array([[0, 1, 2, 3, 4]], np.uint32), np.array([[0, 1, 2, 10, 11]], np.uint32), 5, 0.6))
    nq, k = 50, 100
    b100 = np.arange(nq * k, dtype=np.uint32).reshape(nq, k)
    big = b100.copy()
    big[:, k * 3 // 4:] += 10000
    out.append(("large_k", big, b100, k, 0.75))
    return out


def set_matches(pred, gt, k):
    return sum(len(set(p[:k].tolist()) & set(g[:k].tolist())) for p, g in zip(pred, gt))


def ragged_ids(rng, nq, k, pstride, gstride, universe):
    pred = rng.integers(0, universe, (nq, pstride), dtype=np.uint64).astype(np.uint32)
    gt = rng.integers(0, universe, (nq, gstride), dtype=np.uint64).astype(np.uint32)
    pred[rng.random((nq, pstride)) < 0.05] = INV  # padded slots
    pred[rng.random((nq, pstride)) < 0.02] = 0x7FFFFFFF
    return pred, gt


# ----------------------------------------------------------------- oracle


@pytest.mark.parametrize("case", recall_kats(), ids=lambda c: c[0])
def test_oracle_recall_reference_kats(orc, case):
    _, pred, gt, k, want = case
    m = orc.recall_matches(pred, gt, k)
    assert m / float(pred.shape[0] * k) == want


def test_oracle_recall_matches_sets(orc):
    rng = np.random.default_rng(5)
    for k, universe in ((1, 3), (7, 12), (64, 200), (300, 5000)):
        pred, gt = ragged_ids(rng, 40, k, k + 3, k + 9, universe)
        assert orc.recall_matches(pred, gt, k) == set_matches(pred, gt, k), k


def line(values, dim=16):
    v = np.zeros((len(values), dim), np.float32)
    v[:, 0] = values
    return v


def test_oracle_adr_closed_form(orc):
    # base x_j = j; query q_i = i + 0.25; the true NN is round(q) = i, the
    # prediction i + 2: ratio (1.75^2) / (0.25^2) = 49 under L2SQR and
    # 1.75 / 0.25 = 7 under L2 (exact in FP32); query 5 sits on base 5
    # (ground-truth distance 0: skipped); query 6's top-1 is the invalid id.
    n = 40
    base = line(np.arange(n))
    qx = np.arange(10) + 0.25
    qx[5] = 5.0
    queries = line(qx)
    gt = np.arange(10, dtype=np.uint32)[:, None].repeat(3, 1)
    pred = (np.arange(10, dtype=np.uint32) + 2)[:, None].repeat(2, 1)
    pred[6, 0] = 0x7FFFFFFF
    for metric, term in ((orc.L2SQR, 49.0), (orc.L2, 7.0)):
        adr, counted = orc.adr_at_1(metric, base, queries, pred, gt)
        assert counted == 8 and adr == term, (metric, adr, counted)
    # every query skipped -> 0
    adr, counted = orc.adr_at_1(orc.L2SQR, base, line(np.arange(3)), pred[:3], gt[:3])
    assert (adr, counted) == (0.0, 0)


def test_oracle_probe_knn_and_farthest_closed_form(orc):
    # points at 0..n-1 on one axis: sample s's sorted distances are the
    # smallest of {(s - j)^2 : j != s}; its farthest is max(s, n-1-s)^2.
    n = 300
    base = line(np.arange(n))
    samples = np.array([0, 1, 150, 299, 150, 64], np.uint32)
    knn = orc.probe_knn(orc.L2SQR, base, samples)
    far = orc.probe_farthest(orc.L2SQR, base, samples)
    for i, s in enumerate(samples.tolist()):
        want = sorted((s - j) ** 2 for j in range(n) if j != s)[:128]
        assert knn[i].tolist() == want
        assert far[i] == max(s, n - 1 - s) ** 2


def test_oracle_probe_knn_matches_full_sort(orc):
    rng = np.random.default_rng(11)
    base = rng.standard_normal((700, 32)).astype(np.float32)
    samples = rng.integers(0, 700, 20).astype(np.uint32)
    for metric in (orc.L2SQR, orc.DOT, orc.COSINE, orc.L2):
        dm = orc.distance_matrix(metric, base[samples], base)
        knn = orc.probe_knn(metric, base, samples)
        far = orc.probe_farthest(metric, base, samples)
        for i, s in enumerate(samples.tolist()):
            row = np.delete(dm[i], s)
            assert np.array_equal(knn[i], np.sort(row)[:128]), metric
            assert far[i] == row.max()


def test_quantile_index_is_float_arithmetic():
    # dataset_prober.hpp:283-287: q * column.size() is a float product; 0.7f
    # * 10 rounds to 7.0f (in doubles it is 6.99999988 -> 6).
    from artea_amd.estimators import quantile_columns

    col = np.arange(10, dtype=np.float32)[::-1].reshape(10, 1)
    assert quantile_columns(col, [0.7, 0.0, 0.5, 1.0, 2.0]) == [[7.0, 0.0, 5.0, 9.0, 9.0]]


# ------------------------------------------ C-ABI checks before device work


def test_entry_point_argument_errors():
    from artea_amd._lib import lib

    L = lib()
    fake = C.c_void_p(256)  # never dereferenced: each call fails its checks first
    m = C.c_uint64()
    assert L.artea_recall_at_k_device(fake, 10, fake, 5, 10, 10, C.byref(m), None) != 0
    assert "Ground truth dimension 5 is less than k=10" in L.artea_last_error().decode()
    knn = np.zeros(128, np.float32)
    s = np.array([3], np.uint32)
    sp = s.ctypes.data_as(C.c_void_p)
    assert L.artea_probe_knn_device(0, fake, 128, 16, sp, 1, knn.ctypes.data_as(C.c_void_p), None, None) != 0
    assert "Dataset must contain more than 128 vectors, got 128" in L.artea_last_error().decode()
    assert L.artea_probe_farthest_device(0, fake, 1, 16, sp, 1, knn.ctypes.data_as(C.c_void_p), None) != 0
    assert "Dataset must contain at least 2 vectors, got 1" in L.artea_last_error().decode()
    big = np.array([500], np.uint32)
    assert L.artea_probe_knn_device(0, fake, 200, 16, big.ctypes.data_as(C.c_void_p), 1,
                                    knn.ctypes.data_as(C.c_void_p), None, None) != 0
    assert "sample id 500 is not below 200" in L.artea_last_error().decode()
    assert L.artea_probe_knn_device(0, fake, 200, 16, sp, 0, knn.ctypes.data_as(C.c_void_p), None, None) != 0
    assert "num_samples must be at least 1" in L.artea_last_error().decode()
    a = C.c_double()
    assert L.artea_adr_at_1_device(0, fake, 10, fake, 1, 16, fake, 0, fake, 1, C.byref(a), None, None) != 0
    assert "dimensions must be >= 1" in L.artea_last_error().decode()


# -------------------------------------------------------------------- GPU


def dev(a):
    from artea_amd.device import DeviceBuffer

    return DeviceBuffer.from_numpy(np.ascontiguousarray(a))


@pytest.mark.gpu
def test_gpu_recall_reference_kats_and_ragged_rows(orc):
    from artea_amd import estimators as E
    from artea_amd._lib import check, lib

    for name, pred, gt, k, want in recall_kats():
        r = E.recall_at_k(dev(pred), pred.shape[1], dev(gt), gt.shape[1], pred.shape[0], k)
        assert r == want, name
    rng = np.random.default_rng(7)
    for k, universe in ((1, 4), (10, 30), (100, 400), (1000, 3000), (1024, 1 << 31)):
        pred, gt = ragged_ids(rng, 300, k, k + 5, k + 2, universe)
        dp, dg = dev(pred), dev(gt)  # held for the call
        m = C.c_uint64()
        check(lib().artea_recall_at_k_device(C.c_void_p(dp.ptr), pred.shape[1], C.c_void_p(dg.ptr), gt.shape[1], 300,
                                             k, C.byref(m), None))
        assert m.value == orc.recall_matches(pred, gt, k) == set_matches(pred, gt, k), k


@pytest.mark.gpu
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_gpu_adr_matches_oracle(orc, metric):
    from artea_amd import estimators as E

    rng = np.random.default_rng(100 + metric)
    n, nq, D = 4000, 300, 96
    base = rng.standard_normal((n, D)).astype(np.float32)
    queries = rng.standard_normal((nq, D)).astype(np.float32)
    if metric == 1:
        # DOT distances are -<q, x>: with every product negative the ground
        # truth distances are positive and the queries are counted.
        base, queries = np.abs(base), -np.abs(queries)
    queries[:7] = base[:7]  # ground-truth distance 0 under the L2 metrics
    gt = orc.bruteforce_knn(metric, base, queries, 4)[0].astype(np.uint32)
    pred = gt[:, [1, 0, 2]].copy()
    pred[::3, 0] = gt[::3, 0]
    pred[10, 0] = INV
    adr, counted = E.adr_at_1(metric, dev(base), n, dev(queries), nq, D, dev(pred), 3, dev(gt), 4)
    want = orc.adr_at_1(metric, base, queries, pred, gt)
    assert (adr, counted) == want
    assert counted > 0


@pytest.mark.gpu
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_gpu_probe_knn_and_farthest_match_oracle(orc, metric):
    from artea_amd import estimators as E

    rng = np.random.default_rng(200 + metric)
    n, D = 3000, 96
    base = rng.standard_normal((n, D)).astype(np.float32)
    samples = rng.integers(0, n, 48).astype(np.uint32)
    samples[:3] = [0, n - 1, samples[5]]  # ends and a repeat
    d = dev(base)
    knn, mn = E.knn_table(metric, d, n, D, samples)
    want = orc.probe_knn(metric, base, samples)
    assert np.array_equal(knn, want)
    assert mn == want[:, 0].min()
    far = E.farthest(metric, d, n, D, samples)
    assert np.array_equal(far, orc.probe_farthest(metric, base, samples))


@pytest.mark.gpu
def test_gpu_probe_knn_every_row_a_new_best_and_ties(orc):
    """Base rows in descending distance from the sample: every row enters the
    candidate buffer, so the selection is compacted over and over; integer
    coordinates make many equal distances; n = 129 is the smallest base."""
    from artea_amd import estimators as E

    n = 20000
    base = line(np.concatenate([np.arange(n - 1, 0, -1), [0]]).astype(np.float32))
    samples = np.array([n - 1, 0, n // 2], np.uint32)
    knn, _ = E.knn_table(0, dev(base), n, 16, samples)
    assert np.array_equal(knn, orc.probe_knn(0, base, samples))
    grid = line((np.arange(500) % 7).astype(np.float32))
    s = np.arange(0, 500, 37, dtype=np.uint32)
    knn, _ = E.knn_table(0, dev(grid), 500, 16, s)
    assert np.array_equal(knn, orc.probe_knn(0, grid, s))
    small = np.random.default_rng(3).standard_normal((129, 16)).astype(np.float32)
    s = np.arange(129, dtype=np.uint32)
    knn, _ = E.knn_table(3, dev(small), 129, 16, s)
    assert np.array_equal(knn, orc.probe_knn(3, small, s))


@pytest.mark.gpu
def test_gpu_probe_tables_and_query_probe(orc):
    """probe / l0_min_distance / probe_query through the host module against
    the same statistics computed from the oracle's rows."""
    from artea_amd import estimators as E

    rng = np.random.default_rng(9)
    n, nq, D, k = 5000, 200, 32, 20
    base = rng.standard_normal((n, D)).astype(np.float32)
    queries = rng.standard_normal((nq, D)).astype(np.float32)
    samples = rng.integers(0, n, 100).astype(np.uint32)
    qs = [0.0001, 0.01, 0.5, 0.7, 0.9999]
    d = dev(base)
    res = E.probe(0, d, n, D, qs, samples)
    want = orc.probe_knn(0, base, samples)
    assert res.table == E.quantile_columns(want, qs) and res.min_nearest == want[:, 0].min()
    assert all(res.table[r - 1][i] <= res.table[r][i] for r in range(1, 128) for i in range(len(qs)))
    assert E.l0_min_distance(0, d, n, D, 0.01, samples) == res.table[0][1]
    gt = orc.bruteforce_knn(0, base, queries, k)[0].astype(np.uint32)
    qr = E.probe_query(0, d, n, dev(queries), nq, D, dev(gt), k, qs)
    rows = np.sort(orc.distance_matrix(0, queries, base)[np.arange(nq)[:, None], gt], axis=1)
    assert qr.table == E.quantile_columns(rows, qs) and qr.num_queries == nq and qr.nn_ranks == list(range(1, k + 1))


@pytest.mark.gpu
def test_gpu_distance_rows_invalid_ids(orc):
    from artea_amd import estimators as E
    from artea_amd.device import DeviceBuffer

    rng = np.random.default_rng(4)
    n, na, D, k = 1000, 50, 48, 9
    a = rng.standard_normal((na, D)).astype(np.float32)
    b = rng.standard_normal((n, D)).astype(np.float32)
    ids = rng.integers(0, n, (na, k + 2)).astype(np.uint32)
    ids[3, 4] = INV
    ids[7, 0] = n
    out = DeviceBuffer.empty((na, k), np.float32)
    for metric in (0, 1, 2, 3):
        E.distance_rows(metric, dev(a), na, dev(b), n, dev(ids), k + 2, k, D, out)
        got = out.numpy()
        dm = orc.distance_matrix(metric, a, b)
        for i in range(na):
            for j in range(k):
                want = np.float32(np.inf) if ids[i, j] >= n else dm[i, ids[i, j]]
                assert got[i, j] == want, (metric, i, j)

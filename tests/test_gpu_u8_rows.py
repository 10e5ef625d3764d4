"""One-byte rows: an index whose values are all whole numbers in [0, 255]
keeps a one-byte copy of its rows, and the router computes the distances of
queries of the same kind from it in exact integer arithmetic.

Bar: bit-identical ids and FP32 distances to the FP32-row kernels (which are
themselves pinned against the oracle in test_gpu_router.py), for every
metric, the plain and the batched router and every batched kernel slot;
queries with other values in the same batch take the FP32 fallback.
"""
import numpy as np
import pytest

import artea_amd as aa

pytestmark = pytest.mark.gpu


def test_u8_rows_detection():
    ok = aa.generate(0, 3, 0, 2000, 128)
    assert np.all(ok == np.round(ok)) and ok.min() >= 0 and ok.max() <= 255
    assert aa.DeviceIndex(ok, 0).u8_rows
    for r, c, val in [(5, 7, 0.5), (0, 0, 256.0), (1, 1, -1.0), (3, 3, np.nan)]:
        bad = ok.copy()
        bad[r, c] = val
        assert not aa.DeviceIndex(bad, 0).u8_rows, "value %r must disable one-byte rows" % val
    negzero = ok.copy()
    negzero[2, 2] = -0.0  # either zero is a whole number in [0, 255]
    assert aa.DeviceIndex(negzero, 0).u8_rows
    wide = np.zeros((100, 272), np.float32)  # more than 256 values per row
    assert not aa.DeviceIndex(wide, 0).u8_rows
    assert aa.DeviceIndex(np.zeros((100, 256), np.float32), 0).u8_rows
    idx = aa.DeviceIndex(ok, 0)
    assert idx.set_u8_rows(False) is False and not idx.u8_rows
    assert idx.set_u8_rows(True) is True and idx.u8_rows


def mixed_queries(dim, nq, seed):
    """Whole-number queries, a few of them with one value that is not."""
    q = aa.generate(0, seed, 1, nq, dim)
    q[3, 5] += 0.25     # fraction
    q[7, 0] = 300.0     # above 255
    q[11, 9] = -0.0     # negative zero (still one-byte: exact either way)
    return q


@pytest.mark.parametrize("metric,dim", [(0, 128), (3, 128), (1, 96), (2, 112), (0, 112)])
def test_u8_rows_equal_f32_rows(metric, dim):
    from artea_amd.device import DeviceBuffer, Stream

    vecs = aa.generate(0, 17, 0, 4000, dim)
    qs = mixed_queries(dim, 64, 17)
    idx = aa.DeviceIndex(vecs, metric)
    assert idx.u8_rows
    aa.build_full(idx, max_level=2, build_loops=3, triu_iters=2, seed=5)
    nq = qs.shape[0]
    s = Stream()
    tq = DeviceBuffer.from_numpy(qs)
    for k, L in [(1, 1), (10, 64), (100, 200), (50, 800)]:
        idx.set_u8_rows(False)
        fv, fd = aa.HierarchicalGraphRouter(idx, None, k, L).batch_query(qs)
        idx.set_u8_rows(True)
        uv, ud = aa.HierarchicalGraphRouter(idx, None, k, L).batch_query(qs)
        assert np.array_equal(uv, fv) and np.array_equal(ud.view(np.uint32), fd.view(np.uint32)), (metric, k, L)
        tk = DeviceBuffer.from_numpy(np.full(nq, k, np.uint32))
        tl = DeviceBuffer.from_numpy(np.full(nq, L, np.uint32))
        ov = DeviceBuffer.empty((nq, k), np.uint32)
        od = DeviceBuffer.empty((nq, k), np.float32)
        br = aa.BatchedRouter(idx, k, L)
        assert len(br.variant_names) == 8
        for v in range(8):
            br.set_variant(v)
            assert br.variant() == v
            br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr, stream=s.h)
            s.synchronize()
            assert np.array_equal(ov.numpy(), fv), (metric, k, L, v)
            assert np.array_equal(od.numpy().view(np.uint32), fd.view(np.uint32)), (metric, k, L, v)
        with pytest.raises(aa.ArteaError):
            br.set_variant(8)
        # A pinned one-byte slot reads as its FP32 counterpart once the rows are dropped.
        br.set_variant(5)
        idx.set_u8_rows(False)
        assert br.variant() == 1 and len(br.variant_names) == 4
        br.query_device(tq.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr, stream=s.h)
        s.synchronize()
        assert np.array_equal(ov.numpy(), fv)
        idx.set_u8_rows(True)

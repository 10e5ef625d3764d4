"""Scoring and dataset probes on the GPU (utils/recall_estimator.hpp,
utils/adr_estimator.hpp, utils/dataset_prober.hpp), through libartea_hip.so.

Ids and vectors are device buffers (artea_amd.device.DeviceBuffer); the
probes take their sample ids from the caller and return host arrays."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from ._lib import check, lib

MAX_K = 128  # DatasetProber::MAX_K


def recall_at_k(d_pred, pred_stride: int, d_gt, gt_stride: int, nq: int, k: int, stream=None) -> float:
    """RecallEstimator::calculate_recall_at_k: matches / (nq * k)."""
    m = C.c_uint64()
    check(lib().artea_recall_at_k_device(C.c_void_p(d_pred.ptr), pred_stride, C.c_void_p(d_gt.ptr), gt_stride, nq, k,
                                         C.byref(m), C.c_void_p(stream.h if stream else None)))
    # matches / (nq * k) in doubles: NaN for an empty batch, as there
    return m.value / float(nq * k) if nq * k else float("nan")


def adr_at_1(metric: int, d_base, n: int, d_queries, nq: int, dim_pad: int, d_pred, pred_stride: int, d_gt,
             gt_stride: int, stream=None):
    """ADREstimator::calculate_adr_at_1: (ADR, queries counted)."""
    a, c = C.c_double(), C.c_uint64()
    check(lib().artea_adr_at_1_device(metric, C.c_void_p(d_base.ptr), n, C.c_void_p(d_queries.ptr), nq, dim_pad,
                                      C.c_void_p(d_pred.ptr), pred_stride, C.c_void_p(d_gt.ptr), gt_stride,
                                      C.byref(a), C.byref(c), C.c_void_p(stream.h if stream else None)))
    return a.value, c.value


def knn_table(metric: int, d_base, n: int, dim_pad: int, samples, stream=None):
    """DatasetProber::_compute_knn_table for the given sample ids:
    (num_samples x 128 ascending distances, min_nearest)."""
    s = np.ascontiguousarray(samples, np.uint32)
    out = np.empty((len(s), MAX_K), np.float32)
    mn = C.c_float()
    check(lib().artea_probe_knn_device(metric, C.c_void_p(d_base.ptr), n, dim_pad, s.ctypes.data_as(C.c_void_p),
                                       len(s), out.ctypes.data_as(C.c_void_p), C.byref(mn),
                                       C.c_void_p(stream.h if stream else None)))
    return out, mn.value


def farthest(metric: int, d_base, n: int, dim_pad: int, samples, stream=None) -> np.ndarray:
    """DatasetProber::_compute_farthest for the given sample ids."""
    s = np.ascontiguousarray(samples, np.uint32)
    out = np.empty(len(s), np.float32)
    check(lib().artea_probe_farthest_device(metric, C.c_void_p(d_base.ptr), n, dim_pad, s.ctypes.data_as(C.c_void_p),
                                            len(s), out.ctypes.data_as(C.c_void_p),
                                            C.c_void_p(stream.h if stream else None)))
    return out


def distance_rows(metric: int, d_a, na: int, d_b, nb: int, d_ids, ids_stride: int, k: int, dim_pad: int, d_out,
                  stream=None):
    """out[i, j] = dist(a_i, b[ids[i, j]]) (device buffers)."""
    check(lib().artea_distance_rows_device(metric, C.c_void_p(d_a.ptr), na, C.c_void_p(d_b.ptr), nb,
                                           C.c_void_p(d_ids.ptr), ids_stride, k, dim_pad, C.c_void_p(d_out.ptr),
                                           C.c_void_p(stream.h if stream else None)))


def quantile_columns(rows: np.ndarray, quantiles: Sequence[float]) -> List[List[float]]:
    """Per column (nn rank), the sorted column's entry at index
    min(int(q * len), len - 1) for each quantile (dataset_prober.hpp:268-290)."""
    out = []
    for col in range(rows.shape[1]):
        c = np.sort(rows[:, col])
        row = []
        for q in quantiles:
            idx = int(np.float32(q) * np.float32(len(c)))
            row.append(float(c[min(idx, len(c) - 1)]))
        out.append(row)
    return out


@dataclass
class ProbeResult:
    nn_ranks: List[int]
    quantiles: List[float]
    table: List[List[float]]
    num_samples: int
    min_nearest: float


def probe(metric: int, d_base, n: int, dim_pad: int, quantiles: Sequence[float], samples, stream=None) -> ProbeResult:
    """DatasetProber::probe (dataset_prober.hpp:253-306) over the given
    samples: the 128 x len(quantiles) table and the minimum 1-NN distance."""
    knn, mn = knn_table(metric, d_base, n, dim_pad, samples, stream)
    return ProbeResult(list(range(1, MAX_K + 1)), list(quantiles), quantile_columns(knn, quantiles), len(samples), mn)


def l0_min_distance(metric: int, d_base, n: int, dim_pad: int, quantile: float, samples, stream=None) -> float:
    """The auto-probed L0 minimum distance of a build: probe({quantile})
    .table[0][0] (unit_tests/test_hierarchical_graph_persistence.cpp:125-137)."""
    return probe(metric, d_base, n, dim_pad, [quantile], samples, stream).table[0][0]


@dataclass
class QueryProbeResult:
    nn_ranks: List[int]
    quantiles: List[float]
    table: List[List[float]]
    num_queries: int


def probe_query(metric: int, d_base, n: int, d_queries, nq: int, dim_pad: int, d_gt, k: int,
                quantiles: Sequence[float], stream=None) -> QueryProbeResult:
    """DatasetProber::probe_query (dataset_prober.hpp:327-432): per query the
    distances to its k ground-truth ids (GPU), each row sorted ascending,
    then the quantiles of each rank's column."""
    from .device import DeviceBuffer

    if nq == 0:
        raise ValueError("probe_query: ground truth has zero queries")
    if k == 0:
        raise ValueError("probe_query: ground truth vec_dim (k) is zero")
    out = DeviceBuffer.empty((nq, k), np.float32)
    distance_rows(metric, d_queries, nq, d_base, n, d_gt, k, k, dim_pad, out, stream)
    rows = np.sort(out.numpy(stream), axis=1)
    return QueryProbeResult(list(range(1, k + 1)), list(quantiles), quantile_columns(rows, quantiles), nq)

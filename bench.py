"""Benchmark: QPS at recall@100 >= 0.95 on 1,000,000 x 128 L2 (10,000 queries).

Workload (one "step" = one pass of the router over the whole 10,000-query
batch, inputs and index resident in HBM): the integer-rounded 256-component
Gaussian mixture (seed 20261204), full-hierarchy index built on this GPU
(build_loops 15, triu_iters 4, prefill 0.4, scale 1.10, shift 1.5,
l0_min_distance 1), queries of mixed topk {10, 100} (seeded 50/50 draw).  The
queue size sweep {100, 200, 400, 800} picks the smallest queue with
recall@100 >= 0.95 against an FP64 brute-force ground truth; the timed steps
replay a HIP graph of the batched persistent router at that queue, with the
kernel variant its autotuner picked on the same batch, alternating over
`--streams` streams so consecutive batches overlap (one batch alone is timed
too, and is the roofline's launch duration).  The plain router (one
launch per topk) is timed beside it, and the oracle (C/OpenMP restatement of
the reference router) is timed on the host cores for the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_numa_policy():
    import glob
    import shutil

    nodes = len(glob.glob("/sys/devices/system/node/node[0-9]*"))
    tool = "numactl present but not used" if shutil.which("numactl") else "numactl not installed"
    return "%d NUMA node(s); default local (first-touch) allocation, no interleave (%s)" % (max(nodes, 1), tool)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def recall(pred, gt, k):
    """RecallEstimator (utils/recall_estimator.hpp:53-150): |pred ∩ GT| / k."""
    hits = 0
    for p, g in zip(pred, gt):
        hits += len(set(p[:k].tolist()) & set(g[:k].tolist()))
    return hits / (len(pred) * k)


def gpu_build(idx, args, seed, join=0):
    """The index build with the C0 parameters (build_loops 15, triu_iters 4,
    prefill 0.4, scale 1.10, shift 1.5, l0_min_distance 1): the reference's
    pipeline (r-net insertion with beta 2, tau_k 0, search queue 64, select
    queues 100, then refine_layer per level and compaction), or with
    --hierarchy geometric the seeded geometric level draw.  Returns seconds
    per phase or per level."""
    from artea_amd.refine import build_full, build_full_rnet, rnet_params

    l0 = resolve_l0(idx, args, seed)
    if args.hierarchy == "rnet":
        rs, total = build_full_rnet(idx, rnet_params(seed=seed, l0_min_distance=l0), build_loops=15, triu_iters=4,
                                    prefill=0.4, shift=1.5, join=join, routing_loops=args.routing_loops)
        return np.array([rs, total - rs])
    return build_full(idx, ul_max=32, bl_max=64, build_loops=15, triu_iters=4, prefill=0.4, scale=1.10, shift=1.5,
                      l0_min_distance=l0, seed=seed, level_fraction=args.level_fraction, join=join,
                      routing_loops=args.routing_loops)


def resolve_l0(idx, args, seed):
    """l0_min_distance: the given value, or with --l0-min-distance probe the
    reference's auto-probe (probe's first column at quantile 0.9 over 500
    seeded samples, unit_tests/test_hierarchical_graph_persistence.cpp
    :125-137) on the GPU, in the build metric."""
    if args.l0_min_distance != "probe":
        return float(args.l0_min_distance)
    if getattr(args, "_l0_probed", None) is None or args._l0_probed[0] != idx.n:
        from artea_amd import estimators as E
        from artea_amd.device import DeviceBuffer

        view = DeviceBuffer.borrow(idx.device_vectors(), (idx.n, idx.dim_pad), np.float32)
        samples = np.random.default_rng(seed).integers(0, idx.n, min(500, idx.n)).astype(np.uint32)
        bmetric = 3 if idx.metric == 0 else idx.metric
        t0 = time.time()
        val = E.l0_min_distance(bmetric, view, idx.n, idx.dim_pad, 0.9, samples)
        log("auto-probed l0_min_distance %.6f in %.2fs" % (val, time.time() - t0))
        args._l0_probed = (idx.n, val)
    return args._l0_probed[1]


def cpu_build(orc, base, args, seed, cores, l0):
    """The same build by the oracle on the host cores (l0_min_distance as the
    GPU build used)."""
    if args.hierarchy == "rnet":
        return orc.build_full_rnet(base, orc.rnet_params(seed=seed, l0_min_distance=l0), metric=orc.L2,
                                   build_loops=15, triu_iters=4, prefill=0.4, shift=1.5,
                                   routing_loops=args.routing_loops, nthreads=cores)
    return orc.build_full(base, metric=orc.L2, ul_max=32, bl_max=64, build_loops=15, triu_iters=4, prefill=0.4,
                          scale=1.10, shift=1.5, l0_min_distance=l0, seed=seed, level_fraction=args.level_fraction,
                          routing_loops=args.routing_loops, nthreads=cores)


def sweep_queues(aa, idx, q_d, nq, gt_i, topk, queues, stream, target):
    """Batched router at each queue size (autotuned, HIP-graph replay);
    returns per-queue rows and the smallest queue reaching the target."""
    from artea_amd.device import DeviceBuffer, Event

    tk = DeviceBuffer.from_numpy(np.full(nq, topk, np.uint32))
    ov = DeviceBuffer.empty((nq, topk), np.uint32)
    od = DeviceBuffer.empty((nq, topk), np.float32)
    rows = []
    for L in queues:
        ql = DeviceBuffer.from_numpy(np.full(nq, L, np.uint32))
        br = aa.BatchedRouter(idx, topk, L)
        variant, _ = br.autotune(q_d.ptr, nq, tk.ptr, ql.ptr, ov.ptr, od.ptr)
        graph = br.capture(q_d.ptr, nq, tk.ptr, ql.ptr, ov.ptr, od.ptr)
        graph.launch(stream.h)
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(3):
            graph.launch(stream.h)
        e1.record(stream)
        stream.synchronize()
        t = e0.elapsed_ms(e1) / 3 / 1e3
        pred = ov.numpy(stream)
        r = recall(pred, gt_i, topk)
        st = br.last_stats()
        rows.append({"queue": L, "recall@%d" % topk: round(r, 4), "qps": round(nq / t), "ms": round(t * 1e3, 3),
                     "ndc_per_query": round(st["ndc"] / nq, 1), "variant": br.variant_names[variant]})
        log("queue %d: recall@%d %.4f, %.0f q/s (%s)" % (L, topk, r, nq / t, br.variant_names[variant]))
    ok = [x for x in rows if x["recall@%d" % topk] >= target]
    return rows, (ok[0] if ok else rows[-1])


def workload_cosine(args):
    """1,000,000 x 100 (padded 112) unit-norm power-law Gaussian, cosine
    metric, topk 10 and 100, queue sweep 64..800, FP64 ground truth."""
    import artea_amd as aa
    from artea_amd.device import DeviceBuffer, Stream, bruteforce_knn
    from artea_amd.refine import build_full

    n, nq, dim, seed = args.n, args.nq, 100, 20261203 + 2
    base = aa.generate(1, seed, 0, n, dim)
    queries = aa.generate(1, seed, 1, nq, dim)
    D = base.shape[1]
    idx = aa.DeviceIndex(base, aa.DistanceMetrics.COSINE)
    t0 = time.time()
    secs = gpu_build(idx, args, seed)
    build_s = time.time() - t0
    log("cosine build %.1fs (levels %s, top %d)" % (build_s, np.round(secs, 2).tolist(), idx.info()["top"]))
    stream = Stream()
    q_d = DeviceBuffer.from_numpy(queries)
    base_view = DeviceBuffer.borrow(idx.device_vectors(), base.shape, np.float32)
    t0 = time.time()
    gt_i, gt_d = bruteforce_knn(aa.DistanceMetrics.COSINE, base_view, n, q_d, nq, D, 100, stream, method="mfma")
    log("ground truth (BF16 MFMA filter + FP64 recheck) in %.2fs" % (time.time() - t0))
    out = {"metric": "QPS @ recall>=0.95 (1M x 100 cosine)", "unit": "queries/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f32", "data": "synthetic (power-law-spectrum Gaussian, unit-norm)",
           "config": {"workload": "cosine search, 1,000,000 x 100 (pad 112), %d queries" % nq, "n": n, "nq": nq},
           "build_seconds": round(build_s, 2), "hierarchy": args.hierarchy,
           "routing_loops": args.routing_loops, "level_fraction": args.level_fraction}
    for topk in (10, 100):
        queues = [L for L in (64, 100, 200, 400, 800) if L >= topk]
        rows, best = sweep_queues(aa, idx, q_d, nq, gt_i[:, :topk], topk, queues, stream, args.target_recall)
        out["topk%d" % topk] = {"sweep": rows, "chosen": best}
    # FP32 cosine distances of returned ids against FP64 (norms in FP64).
    r = aa.HierarchicalGraphRouter(idx, None, 10, 200)
    pv, pd = r.batch_query(queries[:200])
    xb = base[pv.astype(np.int64)].astype(np.float64)
    qq = queries[:200].astype(np.float64)
    cos64 = 1.0 - (xb * qq[:, None, :]).sum(-1) / np.sqrt((xb * xb).sum(-1) * (qq * qq).sum(-1)[:, None])
    out["distance_max_abs_err_vs_fp64"] = float(np.max(np.abs(pd.astype(np.float64) - cos64)))
    out["value"] = out["topk100"]["chosen"]["qps"]
    if not args.no_cpu:
        # The oracle router on the same graph (HGRA file), host cores: top-10
        # at the queue chosen for recall@10, on a sample of the queries.
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        orc.build()
        cores = len(os.sched_getaffinity(0))
        L10 = out["topk10"]["chosen"]["queue"]
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "c2.graph")
            idx.snapshot_hgra(p)
            g = orc.Graph.restore(p)
        sample = queries[:1000]
        gv, gd = aa.HierarchicalGraphRouter(idx, None, 10, L10).batch_query(sample)
        reps, t0 = 0, time.time()
        while time.time() - t0 < args.cpu_seconds or reps == 0:
            cv, cd, _ = orc.batch_query(g, base, orc.COSINE, sample, 10, L10, nthreads=cores)
            reps += 1
        cpu_qps = sample.shape[0] * reps / (time.time() - t0)
        same = bool(np.array_equal(cv, gv) and np.array_equal(cd, gd))
        out["cpu_baseline"] = {"value": round(cpu_qps, 1), "unit": "queries/s", "cores": cores, "kind": "port",
                               "cpu_model": host_cpu_model(), "numa": host_numa_policy(),
                               "sample": "1000 queries, top-10 at queue %d, %d repetitions, oracle restatement of "
                                         "batch_query on the GPU-built graph" % (L10, reps),
                               "results_equal_gpu": same}
        log("cosine CPU oracle: %.0f q/s on %d cores (top-10, queue %d), results equal %s" % (cpu_qps, cores, L10,
                                                                                           same))
    print(json.dumps(out))


def workload_c0(args):
    """100,000 x 128 L2 (the integer-valued 256-component mixture, seed
    20261203), 1,000 queries, k = 100 at queue 200: the reference's own
    CPU-runnable configuration.  The hierarchy is built on the GPU and by the
    oracle on the host cores (the two HGRA files compared byte for byte);
    then the GPU routers (plain, and batched as a HIP graph) and the oracle
    router search the same graph (results compared slot for slot), with
    recall@100 against the FP64 ground truth."""
    import artea_amd as aa
    from artea_amd.device import DeviceBuffer, Event, Stream, bruteforce_knn

    n, nq, D, seed, k, L = 100_000, 1000, 128, 20261203, 100, 200
    base = aa.generate(0, seed, 0, n, D)
    queries = aa.generate(0, seed, 1, nq, D)
    idx = aa.DeviceIndex(base, aa.DistanceMetrics.EUCLIDEAN_SQR)
    t0 = time.time()
    gpu_build(idx, args, seed)
    gpu_build_s = time.time() - t0
    log("C0 GPU build %.2fs (%s hierarchy)" % (gpu_build_s, args.hierarchy))
    stream = Stream()
    q_d = DeviceBuffer.from_numpy(queries)
    view = DeviceBuffer.borrow(idx.device_vectors(), base.shape, np.float32)
    gt_i, _ = bruteforce_knn(aa.DistanceMetrics.EUCLIDEAN_SQR, view, n, q_d, nq, D, k, stream)
    # plain router (one launch) and batched router (autotuned, HIP graph)
    ov = DeviceBuffer.empty((nq, k), np.uint32)
    od = DeviceBuffer.empty((nq, k), np.float32)
    plain = aa.HierarchicalGraphRouter(idx, None, k, L)
    tk = DeviceBuffer.from_numpy(np.full(nq, k, np.uint32))
    tl = DeviceBuffer.from_numpy(np.full(nq, L, np.uint32))
    br = aa.BatchedRouter(idx, k, L)
    br.autotune(q_d.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr)
    graph = br.capture(q_d.ptr, nq, tk.ptr, tl.ptr, ov.ptr, od.ptr)
    pv_d, pd_d = DeviceBuffer.empty((nq, k), np.uint32), DeviceBuffer.empty((nq, k), np.float32)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        e0, e1 = Event(), Event()
        stream.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / args.steps / 1e3

    t_plain = timed(lambda: plain.batch_query_device(q_d.ptr, nq, pv_d.ptr, pd_d.ptr, stream=stream.h))
    t_batched = timed(lambda: graph.launch(stream.h))
    gv, gd = ov.numpy(stream), od.numpy(stream)
    same_plain = bool(np.array_equal(gv, pv_d.numpy(stream)) and np.array_equal(gd, pd_d.numpy(stream)))
    rec = recall(gv, gt_i, k)
    log("C0 GPU search: plain %.0f q/s, batched %.0f q/s, recall@100 %.4f, batched = plain %s" % (
        nq / t_plain, nq / t_batched, rec, same_plain))
    out = {"metric": "C0 build seconds and QPS at k=100, queue 200 (100K x 128 L2), GPU vs CPU oracle",
           "value": round(nq / t_batched, 1), "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "u8" if idx.u8_rows else "f32",
           "data": "synthetic (seeded 256-component integer Gaussian mixture)",
           "config": {"workload": "C0: 100,000 x 128 L2, 1,000 queries, k 100, queue 200", "n": n, "nq": nq},
           "gpu_build_s": round(gpu_build_s, 3), "hierarchy": args.hierarchy,
           "routing_loops": args.routing_loops, "recall@100": round(rec, 4),
           "plain_qps": round(nq / t_plain, 1),
           "batched_qps": round(nq / t_batched, 1), "batched_equals_plain": same_plain}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        orc.build()
        cores = len(os.sched_getaffinity(0))
        t0 = time.time()
        g, _ = cpu_build(orc, base, args, seed, cores, resolve_l0(idx, args, seed))
        cpu_build_s = time.time() - t0
        with tempfile.TemporaryDirectory() as d:
            a, b = os.path.join(d, "gpu.graph"), os.path.join(d, "cpu.graph")
            idx.snapshot_hgra(a)
            g.snapshot(b)
            hgra_same = open(a, "rb").read() == open(b, "rb").read()
        cv, cd, _ = orc.batch_query(g, base, orc.L2SQR, queries, k, L, nthreads=cores)
        reps, t0 = 0, time.time()
        while time.time() - t0 < min(args.cpu_seconds, 5.0) or reps == 0:
            orc.batch_query(g, base, orc.L2SQR, queries, k, L, nthreads=cores)
            reps += 1
        cpu_qps = nq * reps / (time.time() - t0)
        same_cpu = bool(np.array_equal(cv, gv) and np.array_equal(cd, gd))
        log("C0 CPU oracle on %d cores: build %.1fs (HGRA identical to the GPU's: %s), search %.0f q/s, results equal "
            "%s" % (cores, cpu_build_s, hgra_same, cpu_qps, same_cpu))
        out["cpu_baseline"] = {"build_s": round(cpu_build_s, 2), "value": round(cpu_qps, 1), "unit": "queries/s",
                               "cores": cores, "kind": "port", "hgra_identical_to_gpu": hgra_same,
                               "cpu_model": host_cpu_model(), "numa": host_numa_policy(),
                               "results_equal_gpu": same_cpu,
                               "sample": "the whole C0 workload: oracle build (same pipeline) and batch_query, "
                                         "OpenMP"}
    print(json.dumps(out))


def workload_probe(args):
    """The L0 minimum-distance auto-probe that precedes a build
    (DatasetProber::probe with the reference CLI's defaults, 500 samples and
    quantile 0.9: each sample's 128 nearest distances over the whole base) on
    the 10,000,000 x 96 unit-norm mixture, L2.  One step = one probe call
    (sample upload, the sliced scan, the merge, the 256 KB download); the
    roofline counts 3 FP32 operations per coordinate (subtract, fused
    multiply-add) against the 157.3 TFLOP/s FP32 peak."""
    import artea_amd as aa
    from artea_amd import estimators as E
    from artea_amd.device import DeviceBuffer, Event, Stream

    n, dim, seed, ns, quantile = args.n, 96, 20261203 + 3, 500, 0.9
    t0 = time.time()
    base = aa.generate(2, seed, 0, n, dim)
    log("generated %d x %d in %.1fs" % (n, dim, time.time() - t0))
    stream = Stream()
    d = DeviceBuffer.from_numpy(base)
    samples = np.random.default_rng(seed).integers(0, n, ns).astype(np.uint32)
    for _ in range(max(args.warmup, 1)):
        knn, mn = E.knn_table(aa.DistanceMetrics.EUCLIDEAN_SQR, d, n, dim, samples, stream)
    e0, e1 = Event(), Event()
    stream.synchronize()
    t0 = time.time()
    e0.record(stream)
    for _ in range(args.steps):
        knn, mn = E.knn_table(aa.DistanceMetrics.EUCLIDEAN_SQR, d, n, dim, samples, stream)
    e1.record(stream)
    stream.synchronize()
    wall = (time.time() - t0) / args.steps
    ev = e0.elapsed_ms(e1) / args.steps / 1e3
    table = E.quantile_columns(knn, [quantile])
    flops = float(ns) * (n - 1) * dim * 3
    achieved = flops / ev / 1e12
    out = {"metric": "L0 minimum-distance probe (500 samples x 10M x 96 L2, top-128 each)", "value": round(ns / wall, 1),
           "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (1024-component mixture, unit-norm)",
           "config": {"workload": "L0 auto-probe, %d x 96 L2, %d samples, quantile %.1f" % (n, ns, quantile), "n": n,
                      "samples": ns},
           "l0_min_distance": table[0][0], "min_nearest": mn,
           "roofline": {"bound": "fp32 valu", "achieved": round(achieved, 2), "peak": 157.3, "unit": "TFLOP/s",
                        "frac": round(achieved / 157.3, 4), "traffic": None,
                        "span": "HIP events around each call (kernels plus its small copies and allocation gaps)"}}
    log("probe: %.2f ms per call (events %.2f ms), %.1f TFLOP/s, l0_min_distance %.6f" % (wall * 1e3, ev * 1e3,
                                                                                       achieved, table[0][0]))
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        orc.build()
        cores = len(os.sched_getaffinity(0))
        done, same, t0 = 0, True, time.time()
        while time.time() - t0 < args.cpu_seconds or done == 0:
            sl = samples[done % ns:done % ns + cores]
            ck = orc.probe_knn(orc.L2SQR, base, sl, nthreads=cores)
            same = same and bool(np.array_equal(ck, knn[done % ns:done % ns + len(sl)]))
            done += len(sl)
        rate = done / (time.time() - t0)
        out["cpu_baseline"] = {"value": round(rate, 2), "unit": "samples/s", "cores": cores, "kind": "port",
                               "cpu_model": host_cpu_model(), "numa": host_numa_policy(),
                               "sample": "%d of the same samples (%d per batch, one per thread), oracle restatement "
                                         "of _compute_knn_table" % (done, cores),
                               "results_equal_gpu": same}
        log("probe CPU oracle: %.2f samples/s on %d cores, rows equal %s" % (rate, cores, same))
    print(json.dumps(out))


def workload_rnet(args):
    """R-net insertion of the 1,000,000 x 128 L2 workload (beta 2, tau_k 0,
    l0_min_distance 1, search queue 64, select queues 100, max 32 / 64, scale
    1.10, waves of inserted/8 up to 32768 after 10 serial vertices).  One step
    = the whole insertion.  The roofline counts 512 bytes (one FP32 row) per
    distance evaluated."""
    import artea_amd as aa
    from artea_amd.refine import RNet, rnet_params

    n, dim, seed = args.n, 128, 20261204
    base = aa.generate(0, seed, 0, n, dim)
    idx = aa.DeviceIndex(base, aa.DistanceMetrics.EUCLIDEAN_SQR)
    p = rnet_params(seed=seed)
    for _ in range(args.warmup):
        RNet(idx, p)
    secs, info, levels = [], None, None
    for _ in range(args.steps):
        t0 = time.time()
        r = RNet(idx, p)
        secs.append(time.time() - t0)
        info, levels = r.info(), np.bincount(r.highest()).tolist()
        del r
    step = float(np.median(secs))
    achieved = info["distances"] * 512 / step / 1e9
    out = {"metric": "r-net insertion vertices/s (1M x 128 L2)", "value": round(n / step, 1), "unit": "vertices/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step * 1e3, 1),
           "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (integer-rounded 256-component mixture)",
           "config": {"workload": "r-net insertion, %d x 128 L2, waves of inserted/8 up to 32768" % n, "n": n},
           "levels": levels, "waves": info["waves"], "distances_per_vertex": round(info["distances"] / n, 1),
           "unresolved": info["unresolved"],
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "span": "host clock around the whole insertion (every wave's plan, apply and host sync)"}}
    log("r-net: %.2fs, %.0f vertices/s, %d waves, %.0f distances per vertex, levels %s" % (
        step, n / step, info["waves"], info["distances"] / n, levels))
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        orc.build()
        cores = len(os.sched_getaffinity(0))
        m = min(n, 100_000)
        sub = np.ascontiguousarray(base[:m])
        op = orc.rnet_params(seed=seed)
        t0 = time.time()
        R = orc.RNet(sub, orc.L2, op, nthreads=cores)
        cpu_s = time.time() - t0
        sidx = aa.DeviceIndex(sub, aa.DistanceMetrics.EUCLIDEAN_SQR)
        t0 = time.time()
        g = RNet(sidx, p)
        gpu_s = time.time() - t0
        same = bool(np.array_equal(g.highest(), R.highest()))
        for lvl in range(R.max_level + 1):
            gc, ge = g.level_rows(lvl)
            oc, oe = R.level_rows(lvl)
            same = same and np.array_equal(gc, oc) and np.array_equal(ge, oe.view(np.uint32).reshape(-1, 2))
        out["cpu_baseline"] = {"value": round(m / cpu_s, 1), "unit": "vertices/s", "cores": cores, "kind": "port",
                               "cpu_model": host_cpu_model(), "numa": host_numa_policy(),
                               "sample": "insertion of the first %d vectors (the oracle's wave restatement, OpenMP "
                                         "over each wave)" % m,
                               "gpu_seconds_same_sample": round(gpu_s, 2), "levels_and_rows_equal_gpu": same}
        log("r-net CPU oracle on the first %d: %.1fs on %d cores (GPU %.2fs), levels and rows equal %s" % (
            m, cpu_s, cores, gpu_s, same))
    print(json.dumps(out))


def workload_hgra(args):
    """HGRA v1 snapshot and restore of the search workload's graph (1M x 128
    L2, built on the GPU with the C0 parameters) through the device path
    (record layout / segmented parse, checks and ELL fill in HBM), against the
    host path (hgra.cpp: parse into the grouped CSR, ELL conversion, copy) and
    the oracle's reader.  One step = one restore of the file into the device
    layout.  `value` is the device-resident part (the parse, its walk, the id
    checks and the ELL fill, file bytes over their host-clock span); the file
    read and the copy to the device are reported beside it."""
    import ctypes as C
    import tempfile

    import artea_amd as aa
    from artea_amd._lib import lib

    n, seed = args.n, 20261203 + 1
    base = aa.generate(0, seed, 0, n, 128)
    idx = aa.DeviceIndex(base, aa.DistanceMetrics.EUCLIDEAN)
    t0 = time.time()
    gpu_build(idx, args, seed)
    log("graph built in %.1fs" % (time.time() - t0))
    tmp = tempfile.mkdtemp(prefix="hgra_", dir=os.environ.get("TMPDIR", "/tmp"))
    path = os.path.join(tmp, "c1.graph")
    host_path = os.path.join(tmp, "c1_host.graph")
    # snapshot: device writer vs the host writer over the same graph
    dev_w = []
    for _ in range(max(1, args.warmup) + args.steps):
        t0 = time.time()
        idx.snapshot_hgra(path)
        dev_w.append((time.time() - t0, idx.hgra_seconds()[1]))
    dev_w = dev_w[max(1, args.warmup):]
    h = C.c_void_p()
    check_rc(lib().artea_index_get_graph(idx.h, C.byref(h)))
    t0 = time.time()
    check_rc(lib().artea_graph_snapshot(h, host_path.encode()))
    host_w = time.time() - t0
    lib().artea_graph_destroy(h)
    data = open(path, "rb").read()
    size = len(data)
    same_w = data == open(host_path, "rb").read()
    # restore: device loader (timed steps) vs host reader + attach
    idx2 = aa.DeviceIndex(base, aa.DistanceMetrics.EUCLIDEAN)
    for _ in range(args.warmup):
        idx2.load_hgra(path)
    loads = []
    for _ in range(args.steps):
        t0 = time.time()
        idx2.load_hgra(path)
        loads.append((time.time() - t0, idx2.hgra_seconds()[0]))
    idx2.snapshot_hgra(host_path)
    same_r = open(host_path, "rb").read() == data
    t0 = time.time()
    h = C.c_void_p()
    check_rc(lib().artea_graph_restore(path.encode(), C.byref(h)))
    host_parse = time.time() - t0
    idx3 = aa.DeviceIndex(base, aa.DistanceMetrics.EUCLIDEAN)
    t0 = time.time()
    check_rc(lib().artea_index_set_graph(idx3.h, h))
    host_attach = time.time() - t0
    lib().artea_graph_destroy(h)
    med = lambda xs: float(np.median(xs))
    total = med([x[0] for x in loads])
    read_copy = med([x[1][0] for x in loads])
    parse = med([x[1][1] for x in loads])
    fill = med([x[1][2] for x in loads])
    dev_part = parse + fill
    w_total = med([x[0] for x in dev_w])
    w_layout = med([x[1][0] for x in dev_w])
    # algorithmic bytes of the device part: the body read by the checks and by
    # the fill, and the ELL arrays written
    lay = idx2.layout()
    ell_bytes = (n * lay["stride0"] + lay["rows_upper"] * lay["stride_upper"]) * 4
    alg = 2 * (size - 28) + ell_bytes
    achieved = alg / dev_part / 1e9
    out = {"metric": "HGRA restore into the device layout, GB/s of file (1M x 128 L2 graph)",
           "value": round(size / dev_part / 1e9, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(dev_part * 1e3, 2), "higher_is_better": True,
           "scaling": "strong", "vs_baseline": None, "dtype": "u32",
           "data": "synthetic (the search workload's graph, built on the GPU)",
           "config": {"workload": "HGRA v1 restore / snapshot, %d-vertex graph, %.1f MB file" % (n, size / 1e6),
                      "n": n, "file_bytes": size},
           "restore_seconds": {"device_total": round(total, 4), "file_read_and_copy": round(read_copy, 4),
                               "parse_and_walk": round(parse, 4), "checks_and_fill": round(fill, 4),
                               "host_parse": round(host_parse, 4), "host_ell_and_copy": round(host_attach, 4)},
           "snapshot_seconds": {"device_total": round(w_total, 4), "device_layout": round(w_layout, 4),
                                "host_writer": round(host_w, 4)},
           "device_file_equals_host_file": same_w, "restored_snapshot_equals_file": same_r,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "span": "host clock over the parse, its walk, the checks and the fill (with their syncs)"}}
    log("HGRA %.1f MB: device restore %.3fs (read+copy %.3f, parse %.4f, checks+fill %.4f) vs host %.3f + %.3f; "
        "device snapshot %.3fs (layout %.4f) vs host writer %.3f; files equal %s/%s" % (
            size / 1e6, total, read_copy, parse, fill, host_parse, host_attach, w_total, w_layout, host_w, same_w,
            same_r))
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        orc.build()
        t0 = time.time()
        g = orc.Graph.restore(path)
        cpu_s = time.time() - t0
        q = os.path.join(tmp, "c1_oracle.graph")
        g.snapshot(q)
        same_o = open(q, "rb").read() == data
        del g
        out["cpu_baseline"] = {"value": round(size / cpu_s / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                               "cpu_model": host_cpu_model(), "numa": host_numa_policy(),
                               "sample": "the whole file restored by the oracle's reader (one thread)",
                               "oracle_snapshot_equals_file": same_o}
        log("oracle reader: %.3fs (%.3f GB/s), its snapshot equals the file %s" % (cpu_s, size / cpu_s / 1e9, same_o))
    for f in (path, host_path, os.path.join(tmp, "c1_oracle.graph")):
        if os.path.exists(f):
            os.remove(f)
    os.rmdir(tmp)
    print(json.dumps(out))


def make_server(Server, idx, max_queue, args):
    """The serving leg's server: admission by request count (--serve-budget),
    or with --serve-hbm-mb by device memory and resident wavefronts."""
    if args.serve_hbm_mb > 0:
        return Server(idx, max_topk=100, max_queue=max_queue, lanes=args.serve_lanes,
                      hbm_bytes=int(args.serve_hbm_mb * 2**20), lds_rounds=args.serve_lds_rounds)
    return Server(idx, max_topk=100, max_queue=max_queue, lanes=args.serve_lanes, budget=args.serve_budget)


def check_rc(rc):
    from artea_amd._lib import check

    check(rc)


def workload_routing(args):
    """One routing loop of refine_layer (RoutingUpdater topk 64 / queue 96,
    merge, PruningUpdater scale 1.10) on the L0 layer of the 1,000,000 x 128
    L2 workload: cap 96 (bl_max 64 x 1.5), rows from prefill 0.4 and the 15 x 4
    build loops (shift 1.5).  One step = one routing loop over every row,
    restored to the same rows before each step.  The roofline counts 512 bytes
    (one FP32 row) per distance evaluated (router and pruning pairs)."""
    import artea_amd as aa
    from artea_amd.device import Event, Stream
    from artea_amd.refine import PropagateEngine

    n, dim, seed, cap, topk, queue = args.n, 128, 20261204, 96, 64, 96
    vecs = aa.generate(0, seed, 0, n, dim)
    idx = aa.DeviceIndex(vecs, aa.DistanceMetrics.EUCLIDEAN_SQR)
    eng = PropagateEngine(idx, cap=cap, metric=aa.DistanceMetrics.EUCLIDEAN, seed=seed)
    t0 = time.time()
    eng.next_prefill(int(np.float32(cap) * np.float32(0.4)), 0, n, stream=0)
    eng.schedule(15, 4, 1.10, 1.5)
    log("L0 build loops: %.1fs" % (time.time() - t0))
    counts, ents = eng.download()
    for _ in range(max(args.warmup, 1)):
        eng.upload(counts, ents)
        eng.routing_loops(1, topk, queue, 1.10)
    eng.routing_stats()
    secs = []
    for _ in range(args.steps):
        eng.upload(counts, ents)
        t0 = time.time()
        eng.routing_loops(1, topk, queue, 1.10)
        secs.append(time.time() - t0)
    st = eng.routing_stats()
    out_c, out_e = eng.download()
    step = float(np.median(secs))
    dist_per_step = st["distances"] / args.steps
    achieved = dist_per_step * 512 / step / 1e9
    out = {"metric": "routing loop (RoutingUpdater + PruningUpdater) rows/s, L0 of 1M x 128 L2",
           "value": round(n / step, 1), "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(step * 1e3, 3), "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (integer-rounded 256-component mixture)",
           "config": {"workload": "one routing loop, L0 of %d x 128 L2, cap 96, topk 64, queue 96" % n, "n": n},
           "distances_per_row": round(dist_per_step / n, 1), "expansions_per_row": round(st["expansions"] /
                                                                                         args.steps / n, 1),
           "unresolved_rows": st["unresolved"], "mean_row_after": round(float(out_c.mean()), 2),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "span": "host clock around each synchronous call (routing pass, rerun check, row copy)"}}
    log("routing loop: %.1f ms, %.0f rows/s, %.0f distances/row, %.1f GB/s by 512 B per distance" % (
        step * 1e3, n / step, dist_per_step / n, achieved))
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        orc.build()
        cores = len(os.sched_getaffinity(0))
        # Every sampled pivot must route over the rows as they were before
        # the pass: a pilot on one copy sizes the sample, one call on a
        # fresh copy is timed.
        entries = ents.view(orc.NBR_DTYPE).reshape(-1)
        L = orc.Layer(vecs, orc.L2, cap)
        L.set_rows_raw(counts, entries)
        t0 = time.time()
        L.routing_range(topk, queue, 0, 4 * cores, nthreads=cores)
        pilot = 4 * cores / max(time.time() - t0, 1e-6)
        rows = int(min(n, max(4 * cores, pilot * args.cpu_seconds)))
        L = orc.Layer(vecs, orc.L2, cap)
        L.set_rows_raw(counts, entries)
        t0 = time.time()
        L.routing_range(topk, queue, 0, rows, nthreads=cores)
        rate = rows / (time.time() - t0)
        # the pruned rows of the sampled pivots against the GPU's
        L.pruning(1.10)
        same = True
        oc, oe = L.export()
        off = np.concatenate([[0], np.cumsum(oc)]).astype(np.int64)
        goff = np.concatenate([[0], np.cumsum(out_c)]).astype(np.int64)
        for r in range(0, rows, 97):
            want = oe[off[r]:off[r] + min(int(oc[r]), cap)].view(np.uint32).reshape(-1, 2)
            same = same and np.array_equal(want, out_e[goff[r]:goff[r + 1]])
        out["cpu_baseline"] = {"value": round(rate, 1), "unit": "rows/s", "cores": cores, "kind": "port",
                               "cpu_model": host_cpu_model(), "numa": host_numa_policy(),
                               "sample": "routing of the first %d pivot rows (oracle restatement, OpenMP), then the "
                                         "pruning of every row" % rows,
                               "results_equal_gpu": bool(same)}
        log("routing CPU oracle: %.0f rows/s on %d cores, sampled rows equal %s" % (rate, cores, same))
    print(json.dumps(out))


def workload_build(args):
    """10,000,000 x 96 unit-norm 1024-component mixture, L2: the hierarchy
    build (15 x 4 schedule, prefill 0.4, scale 1.10, shift 1.5) with sparse
    upper layers (seeded geometric levels, --level-fraction), with the FP32
    join and with the BF16 MFMA join; edge-set equality, recall@100 of the
    resulting index, then the serving leg on the same resident index
    (open-loop Poisson arrivals, topk {10,50,100}, queue {64,128,256,512,800}).
    The hierarchy comes from r-net insertion (--hierarchy rnet, the
    reference's pipeline): on unit-norm data its covering radii (1, 2, ...)
    absorb almost every vertex, so the upper levels are small; layer 0 alone
    does not route here (the 1.5 shift exceeds every distance, leaving a
    plain k-NN graph whose clusters do not link up)."""
    import hashlib

    import artea_amd as aa
    from artea_amd.device import DeviceBuffer, Stream, bruteforce_knn
    from artea_amd.refine import JOIN_BF16, JOIN_FP32, build_full

    n, nq, dim, seed = args.n, args.nq, 96, 20261203 + 3
    t0 = time.time()
    base = aa.generate(2, seed, 0, n, dim)
    queries = aa.generate(2, seed, 1, nq, dim)
    log("generated %d x %d in %.1fs" % (n, dim, time.time() - t0))
    idx = aa.DeviceIndex(base, aa.DistanceMetrics.EUCLIDEAN_SQR)
    builds = {}
    graphs = {}
    for name, join in (("fp32", JOIN_FP32), ("bf16", JOIN_BF16)):
        t0 = time.time()
        lv = gpu_build(idx, args, seed, join=join)
        secs = time.time() - t0
        g = idx.export_graph()
        digest = hashlib.sha256(g.nbrs.tobytes() + b"".join(o.tobytes() for o in g.group_offsets)).hexdigest()[:16]
        builds[name] = {"seconds": round(secs, 2),
                        ("rnet_then_refine_seconds" if args.hierarchy == "rnet" else "layer_seconds"):
                            np.round(lv, 2).tolist(),
                        "edges": int(g.nbrs.size), "digest": digest}
        graphs[name] = g
        log("%s build: %.1fs (layers %s), %d edges, digest %s" % (name, secs, np.round(lv, 2).tolist(), g.nbrs.size,
                                                                digest))
    same = bool(np.array_equal(graphs["fp32"].nbrs, graphs["bf16"].nbrs) and
                all(np.array_equal(a, b) for a, b in zip(graphs["fp32"].group_offsets, graphs["bf16"].group_offsets)))
    del graphs
    stream = Stream()
    q_d = DeviceBuffer.from_numpy(queries)
    base_view = DeviceBuffer.borrow(idx.device_vectors(), base.shape, np.float32)
    t0 = time.time()
    gt_i, gt_d = bruteforce_knn(aa.DistanceMetrics.EUCLIDEAN_SQR, base_view, n, q_d, nq, dim, 100, stream,
                                method="mfma")
    gt_s = time.time() - t0
    log("ground truth (BF16 MFMA filter + FP64 recheck) in %.2fs" % gt_s)
    rows, best = sweep_queues(aa, idx, q_d, nq, gt_i, 100, [100, 200, 400, 800], stream, args.target_recall)
    out = {"metric": "10M x 96 propagate build seconds (every level), BF16 vs FP32 join; serving p50/p99 ms",
           "value": builds["bf16"]["seconds"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
           "dtype": "bf16 MFMA join tiles (FP32 accumulate, FP32 rechecks) / f32",
           "data": "synthetic (1024-component mixture, unit-norm)",
           "config": {"workload": "hierarchy build (every level) + serving, %d x 96 L2" % n, "n": n, "nq": nq},
           "builds": builds, "bf16_speedup": round(builds["fp32"]["seconds"] / builds["bf16"]["seconds"], 3),
           "edge_sets_identical": same, "top": idx.info()["top"], "hierarchy": args.hierarchy,
           "routing_loops": args.routing_loops,
           "l0_min_distance": resolve_l0(idx, args, seed),
           "level_fraction": args.level_fraction,
           "ground_truth_s": round(gt_s, 2), "search_sweep": rows, "search_chosen": best}
    if args.serve_requests > 0:
        from artea_amd.serving import Server, percentile_nearest_rank, poisson_arrivals

        rng = np.random.default_rng(seed)
        ns = args.serve_requests
        pick = rng.integers(0, nq, ns)
        stopk = rng.choice(np.array([10, 50, 100], np.uint32), ns)
        squeue = rng.choice(np.array([64, 128, 256, 512, 800], np.uint32), ns)
        squeue = np.maximum(squeue, stopk)
        srv = make_server(Server, idx, 800, args)
        # untimed warm-up (first launches of every class router), then the
        # saturation: one closed batch of the same mix
        nw = min(ns, args.serve_lanes * srv.budget)
        srv.run(queries[pick[:nw]], stopk[:nw], squeue[:nw], np.zeros(nw))
        v, d, lat, wall = srv.run(queries[pick], stopk, squeue, np.zeros(ns))
        sat = ns / wall
        hits = sum(len(set(v[i, :stopk[i]].tolist()) & set(gt_i[pick[i], :stopk[i]].tolist())) for i in range(ns))
        rec = hits / float(stopk.sum())
        loads = []
        for i, frac in enumerate((0.25, 0.5, 0.75, 0.9)):
            rate = frac * sat
            v, d, lat, wall = srv.run(queries[pick], stopk, squeue, poisson_arrivals(ns, rate, seed=seed + i))
            loads.append({"offered_qps": round(rate), "achieved_qps": round(ns / wall),
                          "p50_ms": round(percentile_nearest_rank(lat, 50) * 1e3, 3),
                          "p99_ms": round(percentile_nearest_rank(lat, 99) * 1e3, 3),
                          "p99_ms_queue64": round(percentile_nearest_rank(lat[squeue == 64], 99) * 1e3, 3),
                          "p99_ms_queue800": round(percentile_nearest_rank(lat[squeue == 800], 99) * 1e3, 3)})
            log("serving at %.0f q/s offered: %.0f achieved, p50 %.2f ms p99 %.2f ms (queue 64: %.2f, queue 800: "
                "%.2f)" % (rate, ns / wall, loads[-1]["p50_ms"], loads[-1]["p99_ms"], loads[-1]["p99_ms_queue64"],
                           loads[-1]["p99_ms_queue800"]))
        out["serving"] = {"requests_per_load": ns, "lanes": args.serve_lanes, "lane_budget": srv.budget,
                          "admission": srv.admission(),
                          "saturation_qps": round(sat), "recall_at_topk": round(rec, 4), "loads": loads}
        if not args.no_cpu:
            # CPU closed loop on a sample of the same mix: the oracle router on
            # this index's HGRA file, one batch per (topk, queue) group, every
            # request present at the start; a request's latency is the time
            # its group's batch completes.  Its lists are compared with the
            # server's for the same requests.
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as orc

            orc.build()
            cores = len(os.sched_getaffinity(0))
            with tempfile.TemporaryDirectory() as tmpd:
                gp = os.path.join(tmpd, "c3.graph")
                idx.snapshot_hgra(gp)
                g = orc.Graph.restore(gp)
            smp = np.arange(min(ns, 2000))
            clat = np.zeros(len(smp))
            same_cpu = True
            t0 = time.time()
            for kk in np.unique(stopk[smp]):
                for LL in np.unique(squeue[smp][stopk[smp] == kk]):
                    sel = smp[(stopk[smp] == kk) & (squeue[smp] == LL)]
                    cv, cd, _ = orc.batch_query(g, base, orc.L2SQR, queries[pick[sel]], int(kk), int(LL),
                                                nthreads=cores)
                    clat[sel] = time.time() - t0
                    same_cpu &= bool(np.array_equal(cv, v[sel, :kk]) and np.array_equal(cd, d[sel, :kk]))
            cwall = time.time() - t0
            out["serving"]["cpu_closed_loop"] = {
                "requests": len(smp), "qps": round(len(smp) / cwall, 1), "cores": cores,
                "p50_ms": round(percentile_nearest_rank(clat, 50) * 1e3, 1),
                "p99_ms": round(percentile_nearest_rank(clat, 99) * 1e3, 1), "results_equal_gpu": same_cpu}
            log("CPU oracle closed loop on %d requests of the mix: %.0f q/s on %d cores, p50 %.0f ms, p99 %.0f ms, "
                "results equal %s" % (len(smp), len(smp) / cwall, cores, out["serving"]["cpu_closed_loop"]["p50_ms"],
                                       out["serving"]["cpu_closed_loop"]["p99_ms"], same_cpu))
    if not args.no_cpu:
        # CPU baseline on a bounded sample: the whole build (same parameters,
        # same level draw) of the first 500,000 vectors, by the GPU (FP32
        # join) and by the oracle on the host cores; HGRA files compared.
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        orc.build()
        cores = len(os.sched_getaffinity(0))
        m = min(n, 500_000)
        sub = np.ascontiguousarray(base[:m])
        sidx = aa.DeviceIndex(sub, aa.DistanceMetrics.EUCLIDEAN_SQR)
        t0 = time.time()
        gpu_build(sidx, args, seed, join=JOIN_FP32)
        gpu_s = time.time() - t0
        t0 = time.time()
        g, _ = cpu_build(orc, sub, args, seed, cores, resolve_l0(sidx, args, seed))
        cpu_s = time.time() - t0
        with tempfile.TemporaryDirectory() as d:
            a, b = os.path.join(d, "gpu.graph"), os.path.join(d, "cpu.graph")
            sidx.snapshot_hgra(a)
            g.snapshot(b)
            hgra_same = open(a, "rb").read() == open(b, "rb").read()
        out["cpu_baseline"] = {"value": round(cpu_s, 2), "unit": "s", "cores": cores, "kind": "port",
                               "cpu_model": host_cpu_model(), "numa": host_numa_policy(),
                               "sample": "whole build of the first %d vectors (FP32 join), the oracle's build "
                                         "(same pipeline) with OpenMP" % m,
                               "gpu_seconds_same_sample": round(gpu_s, 2), "hgra_identical_to_gpu": hgra_same}
        log("CPU oracle build of the first %d vectors: %.1fs on %d cores (GPU %.2fs), HGRA identical %s" % (
            m, cpu_s, cores, gpu_s, hgra_same))
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--join", choices=("fp32", "bf16"), default="bf16",
                    help="triangle-join pair distances of the search workload's build: BF16 MFMA tiles with FP32 "
                         "rechecks (default) or the FP32 chain; the graph is the same")
    ap.add_argument("--hierarchy", choices=("rnet", "geometric"), default="rnet",
                    help="levels and upper-level rows from r-net insertion (the reference's pipeline) or from the "
                         "seeded geometric draw with a random prefill")
    ap.add_argument("--routing-loops", type=int, default=0,
                    help="routing loops (RoutingUpdater topk 64, queue 96) after each level's build loops; the "
                         "configs name none (PropagateConfig's default 0), the reference recommends 1")
    ap.add_argument("--l0-min-distance", default="1.0",
                    help="a value, or 'probe' for the reference's auto-probe (quantile 0.9, 500 samples)")
    ap.add_argument("--level-fraction", type=float, default=None,
                    help="hierarchy level fraction (default: 1 for search, 1/32 for cosine and build)")
    ap.add_argument("--workload", choices=("search", "cosine", "build", "c0", "probe", "routing", "rnet", "hgra"),
                    default="search",
                    help="search: 1M x 128 L2 batched vs plain (the headline); cosine: 1M x 100 cosine sweep; "
                         "build: 10M x 96 hierarchy build (every level) FP32 vs BF16 join, then serving on that index; "
                         "c0: 100K x 128 build and search, GPU vs the CPU oracle")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--queues", type=str, default="100,200,400,800")
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bounded CPU-baseline sample")
    ap.add_argument("--streams", type=int, default=2,
                    help="streams the timed steps go round (consecutive batches overlap); 1 = one batch at a time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--f32-rows", action="store_true",
                    help="search the FP32 rows only (drop the exact one-byte copy of whole-number data)")
    ap.add_argument("--serve-requests", type=int, default=100000,
                    help="requests per offered-load level of the serving leg (0 = skip)")
    ap.add_argument("--serve-lanes", type=int, default=None,
                    help="serving: overlapping launch lanes (streams); default 2 for search, 4 for build")
    ap.add_argument("--serve-budget", type=int, default=None,
                    help="serving: most requests one lane launches at once; default 6144 for search, 2048 for build")
    ap.add_argument("--serve-hbm-mb", type=float, default=0.0,
                    help="serving: admit by device memory instead of --serve-budget: the lanes' buffers share this "
                         "many MiB, and a launch takes at most --serve-lds-rounds times its kernel's resident waves")
    ap.add_argument("--serve-lds-rounds", type=int, default=2,
                    help="serving with --serve-hbm-mb: resident-wavefront passes one launch may take")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after timing, run one plain copy kernel of the base vectors (counter calibration)")
    ap.add_argument("--graph-cache", type=str, default="",
                    help="HGRA path: load the index graph from it if present, else build and save it there")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X and this path uses no collectives; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.level_fraction is None:
        args.level_fraction = 1.0 if args.workload == "search" else 1.0 / 32
    if args.serve_lanes is None:
        args.serve_lanes = 2 if args.workload == "search" else 4
    if args.serve_budget is None:
        args.serve_budget = 6144 if args.workload == "search" else 2048
    if args.workload == "cosine":
        return workload_cosine(args)
    if args.workload == "c0":
        return workload_c0(args)
    if args.workload == "routing":
        return workload_routing(args)
    if args.workload == "rnet":
        return workload_rnet(args)
    if args.workload == "hgra":
        return workload_hgra(args)
    if args.workload == "probe":
        if args.n == 1_000_000:
            args.n = 10_000_000
        return workload_probe(args)
    if args.workload == "build":
        if args.n == 1_000_000:
            args.n = 10_000_000
        return workload_build(args)

    import artea_amd as aa
    from artea_amd.device import DeviceBuffer, Event, Stream, bruteforce_knn, device_info
    from artea_amd.refine import build_full

    D, DIM = 128, 128
    seed = 20261203 + 1
    t0 = time.time()
    base = aa.generate(0, seed, 0, args.n, DIM)
    queries = aa.generate(0, seed, 1, args.nq, DIM)
    log("generated %d x %d base, %d queries in %.1fs" % (args.n, D, args.nq, time.time() - t0))
    rng = np.random.default_rng(seed)
    topk_q = rng.choice(np.array([10, 100], np.uint32), args.nq)
    dev = device_info()
    log("device %s, %d CUs" % (dev["arch"], dev["compute_units"]))

    idx = aa.DeviceIndex(base, aa.DistanceMetrics.EUCLIDEAN_SQR)
    if args.f32_rows:
        idx.set_u8_rows(False)
    log("one-byte rows: %s" % idx.u8_rows)
    t0 = time.time()
    if args.graph_cache and os.path.exists(args.graph_cache):
        idx.load_hgra(args.graph_cache)
        secs = np.zeros(0)
        log("loaded the index graph from %s (build not timed in this run)" % args.graph_cache)
    else:
        from artea_amd.refine import JOIN_BF16, JOIN_FP32

        secs = gpu_build(idx, args, seed, join=JOIN_BF16 if args.join == "bf16" else JOIN_FP32)
        if args.graph_cache:
            idx.snapshot_hgra(args.graph_cache)
    build_s = time.time() - t0
    info = idx.info()
    log("GPU build %.1fs (levels %s), top %d entry %d" % (build_s, np.round(secs, 2).tolist(), info["top"],
                                                          info["entry"]))

    stream = Stream()
    nq = args.nq
    q_d = DeviceBuffer.from_numpy(queries)
    base_view = DeviceBuffer.borrow(idx.device_vectors(), base.shape, np.float32)
    t0 = time.time()
    gt_i, gt_d = bruteforce_knn(aa.DistanceMetrics.EUCLIDEAN_SQR, base_view, args.n, q_d, nq, D, 100, stream)
    gt_fp64_s = time.time() - t0
    t0 = time.time()
    gst = {}
    mi, md = bruteforce_knn(aa.DistanceMetrics.EUCLIDEAN_SQR, base_view, args.n, q_d, nq, D, 100, stream,
                            method="mfma", stats=gst)
    gt_mfma_s = time.time() - t0
    gt_same = bool(np.array_equal(mi, gt_i) and np.array_equal(md, gt_d))
    log("ground truth: FP64 kernel %.2fs, BF16 MFMA filter + FP64 recheck %.2fs, identical %s, kept %.0f pairs/query" % (
        gt_fp64_s, gt_mfma_s, gt_same, gst.get("kept_pairs", 0) / nq))

    tk_d = DeviceBuffer.from_numpy(topk_q)
    ov = DeviceBuffer.empty((nq, 100), np.uint32)
    od = DeviceBuffer.empty((nq, 100), np.float32)
    sel10 = np.where(topk_q == 10)[0]
    sel100 = np.where(topk_q == 100)[0]
    q10 = DeviceBuffer.from_numpy(queries[sel10])
    q100 = DeviceBuffer.from_numpy(queries[sel100])
    o10v = DeviceBuffer.empty((len(sel10), 10), np.uint32)
    o10d = DeviceBuffer.empty((len(sel10), 10), np.float32)
    o100v = DeviceBuffer.empty((len(sel100), 100), np.uint32)
    o100d = DeviceBuffer.empty((len(sel100), 100), np.float32)
    sh = stream.h

    def time_it(fn, reps):
        e0, e1 = Event(), Event()
        stream.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / reps / 1e3

    sweep = []
    routers = {}
    for L in [int(x) for x in args.queues.split(",")]:
        ql_d = DeviceBuffer.from_numpy(np.full(nq, L, np.uint32))
        br = aa.BatchedRouter(idx, 100, L)
        p10 = aa.HierarchicalGraphRouter(idx, None, 10, L)
        p100 = aa.HierarchicalGraphRouter(idx, None, 100, L)

        def run_plain(p10=p10, p100=p100):
            p10.batch_query_device(q10.ptr, len(sel10), o10v.ptr, o10d.ptr, stream=sh)
            p100.batch_query_device(q100.ptr, len(sel100), o100v.ptr, o100d.ptr, stream=sh)

        # Autotune the persistent kernel variant on this batch, then capture
        # the batched query (counter reset + router) as a HIP graph.
        variant, var_ms = br.autotune(q_d.ptr, nq, tk_d.ptr, ql_d.ptr, ov.ptr, od.ptr)
        graph = br.capture(q_d.ptr, nq, tk_d.ptr, ql_d.ptr, ov.ptr, od.ptr)

        def run_batched(graph=graph):
            graph.launch(sh)

        run_batched()
        stream.synchronize()
        tb = time_it(run_batched, 1)
        st = br.last_stats()
        pred = ov.numpy(stream)
        r100 = recall(pred[sel100], gt_i[sel100], 100)
        r10 = recall(pred[sel10], gt_i[sel10], 10)
        run_plain()
        tp = time_it(run_plain, 1)
        pv = o100v.numpy(stream)
        same = bool(np.array_equal(pv, pred[sel100]) and np.array_equal(o10v.numpy(stream), pred[sel10, :10]))
        sweep.append({"queue": L, "recall@100": round(r100, 4), "recall@10": round(r10, 4),
                      "batched_qps": round(nq / tb), "plain_qps": round(nq / tp), "ndc_per_query":
                      round(st["ndc"] / nq, 1), "batched_equals_plain": same, "tail_overflows": st["tail_overflows"],
                      "variant": br.variant_names[variant],
                      "variant_ms": {br.variant_names[i]: round(float(x), 3) for i, x in enumerate(var_ms)}})
        log("queue %d: recall@100 %.4f recall@10 %.4f batched %.0f q/s plain %.0f q/s ndc %.0f same %s variant %s" % (
            L, r100, r10, nq / tb, nq / tp, st["ndc"] / nq, same, br.variant_names[variant]))
        routers[L] = (br, p10, p100, run_batched, run_plain, st, graph, var_ms, ql_d)
    ok = [s for s in sweep if s["recall@100"] >= args.target_recall]
    chosen = ok[0]["queue"] if ok else sweep[-1]["queue"]
    br, p10, p100, run_batched, run_plain, st, graph, var_ms, ql_d = routers[chosen]

    # One batch on one stream: the kernel's launch duration (the roofline's
    # time) and the batch latency.
    for _ in range(args.warmup):
        run_batched()
    t_batch = time_it(run_batched, args.steps)
    st = br.last_stats()
    for _ in range(args.warmup):
        run_plain()
    t_plain_batch = time_it(run_plain, args.steps)
    pred = ov.numpy(stream)
    pdist = od.numpy(stream)
    slot = br.variant()

    # Timed steps: consecutive batches are independent, so the steps go round
    # `--streams` streams, each with its own routers (work counters), outputs
    # and launch graph, and one batch's tail overlaps the next batch's start,
    # as a serving loop's lanes do.  The plain routers are timed the same way.
    lanes = [(stream, run_batched, run_plain, ov, od, br)]
    for _ in range(1, max(1, args.streams)):
        s_j = Stream()
        br_j = aa.BatchedRouter(idx, 100, chosen)
        br_j.set_variant(slot)
        ov_j, od_j = DeviceBuffer.empty((nq, 100), np.uint32), DeviceBuffer.empty((nq, 100), np.float32)
        g_j = br_j.capture(q_d.ptr, nq, tk_d.ptr, ql_d.ptr, ov_j.ptr, od_j.ptr)
        p10_j = aa.HierarchicalGraphRouter(idx, None, 10, chosen)
        p100_j = aa.HierarchicalGraphRouter(idx, None, 100, chosen)
        b10 = (DeviceBuffer.empty((len(sel10), 10), np.uint32), DeviceBuffer.empty((len(sel10), 10), np.float32))
        b100 = (DeviceBuffer.empty((len(sel100), 100), np.uint32), DeviceBuffer.empty((len(sel100), 100), np.float32))

        def run_b(g_j=g_j, s_j=s_j, br_j=br_j):
            g_j.launch(s_j.h)

        def run_p(p10_j=p10_j, p100_j=p100_j, s_j=s_j, b10=b10, b100=b100):
            p10_j.batch_query_device(q10.ptr, len(sel10), b10[0].ptr, b10[1].ptr, stream=s_j.h)
            p100_j.batch_query_device(q100.ptr, len(sel100), b100[0].ptr, b100[1].ptr, stream=s_j.h)

        lanes.append((s_j, run_b, run_p, ov_j, od_j, br_j))

    def time_lanes(fns, reps, launch_ms=None):
        """Seconds per step over `reps` steps going round the lanes.  With
        launch_ms, every launch of the timed region is bracketed by events on
        its own stream and their durations are appended (the kernel's
        average launch duration for the roofline, overlap included)."""
        for fn in fns:
            for _ in range(args.warmup):
                fn()
        for ln in lanes:
            ln[0].synchronize()
        evs = [(Event(), Event()) for _ in range(reps)] if launch_ms is not None else None
        t0 = time.perf_counter()
        for i in range(reps):
            j = i % len(fns)
            if evs:
                evs[i][0].record(lanes[j][0])
            fns[j]()
            if evs:
                evs[i][1].record(lanes[j][0])
        for ln in lanes:
            ln[0].synchronize()
        dt = (time.perf_counter() - t0) / reps
        if evs:
            launch_ms.extend(e0.elapsed_ms(e1) for e0, e1 in evs)
        return dt

    def written_equal(a, b):
        # A top-k query writes the first k slots of its row; the rest are untouched.
        return bool(np.array_equal(a[sel100, :100], b[sel100, :100]) and np.array_equal(a[sel10, :10], b[sel10, :10]))

    launch_ms = []
    t_step = time_lanes([ln[1] for ln in lanes], args.steps, launch_ms)
    t_launch = float(np.mean(launch_ms)) / 1e3
    t_plain = time_lanes([ln[2] for ln in lanes], args.steps)
    lanes_equal = all(written_equal(ln[3].numpy(ln[0]), pred) and
                      written_equal(ln[4].numpy(ln[0]).view(np.uint32), pdist.view(np.uint32)) for ln in lanes)
    qps = nq / t_step
    qps_plain = nq / t_plain
    log("chosen queue %d: batched %.0f q/s (%.3f ms/step over %d streams; one batch alone %.3f ms), plain %.0f q/s "
        "(alone %.3f ms)" % (chosen, qps, t_step * 1e3, len(lanes), t_batch * 1e3, qps_plain, t_plain_batch * 1e3))
    u8 = idx.u8_rows and slot >= len(aa.BatchedRouter.VARIANTS)
    f32_rows = None
    if u8:
        # The same batch on the FP32 rows (the fastest FP32 slot of the
        # autotune): equal results, its throughput reported beside.
        f32_slot = int(np.argmin(np.asarray(var_ms[: len(aa.BatchedRouter.VARIANTS)])))
        f32_fns = []
        for s_j, _, _, ov_j, od_j, br_j in lanes:
            br_j.set_variant(f32_slot)
            gf = br_j.capture(q_d.ptr, nq, tk_d.ptr, ql_d.ptr, ov_j.ptr, od_j.ptr)
            f32_fns.append(lambda gf=gf, s_j=s_j: gf.launch(s_j.h))
        for _ in range(args.warmup):
            f32_fns[0]()
        t_f32_batch = time_it(f32_fns[0], args.steps)
        st_f32 = br.last_stats()  # the FP32 kernel's own counters (its visited cache differs)
        f32_launch_ms = []
        t_f32 = time_lanes(f32_fns, args.steps, f32_launch_ms)
        same_f32 = all(written_equal(ln[3].numpy(ln[0]), pred) and
                       written_equal(ln[4].numpy(ln[0]).view(np.uint32), pdist.view(np.uint32)) for ln in lanes)
        for ln in lanes:
            ln[5].set_variant(slot)
        f32_rows = {"qps": round(nq / t_f32, 1), "ms_per_step": round(t_f32 * 1e3, 4),
                    "batch_ms": round(t_f32_batch * 1e3, 4),
                    "launch_ms": round(float(np.mean(f32_launch_ms)), 4),
                    "ndc_per_query": round(st_f32["ndc"] / nq, 1),
                    "variant": br.variant_names[f32_slot], "results_equal": same_f32}
        log("FP32 rows: %.0f q/s (%.3f ms/step; one batch alone %.3f ms, %s), results equal %s" % (
            nq / t_f32, t_f32 * 1e3, t_f32_batch * 1e3, br.variant_names[f32_slot], same_f32))
    # FP32 router distances against FP64 recomputation (same ids), on the
    # top-100 queries: the integer-valued data makes every FP32 partial sum
    # exact, so the expected error is zero.
    smp = sel100[:200]
    x = base[pred[smp, :100].astype(np.int64)].astype(np.float64)
    d64 = ((x - queries[smp][:, None, :].astype(np.float64)) ** 2).sum(-1)
    dist_err = float(np.max(np.abs(pdist[smp, :100].astype(np.float64) - d64) / np.maximum(d64, 1e-300)))
    log("router FP32 distance max relative error vs FP64: %.3g" % dist_err)
    if args.calibrate_copy:
        # One plain copy of the base vectors (B bytes in, B bytes out): the
        # counter calibration dispatch for the HBM traffic figure.
        from artea_amd._lib import check, lib

        dst = DeviceBuffer(base.nbytes)
        check(lib().artea_copy_device(dst.ptr, idx.device_vectors(), base.nbytes, stream.h))
        stream.synchronize()
        log("calibration copy of %d bytes done" % base.nbytes)

    # Algorithmic bytes of one launch: per query the query row and the result
    # rows, per distance evaluation one base row (4*D bytes, or D with
    # one-byte rows), per beam expansion one neighbour row (4 bytes per
    # stored neighbour id).
    row_bytes = 4 * 64
    vec_bytes = D if u8 else 4 * D
    alg_bytes = nq * 4 * D + 8 * int(topk_q.sum()) + st["ndc"] * vec_bytes + st["expansions"] * row_bytes
    # the roofline's time: the kernel's average launch duration over the
    # timed steps (two launches overlap at a time); alone = one batch alone
    achieved = alg_bytes / t_launch / 1e9
    if f32_rows is not None:
        fb = nq * 4 * D + 8 * int(topk_q.sum()) + st_f32["ndc"] * 4 * D + st_f32["expansions"] * row_bytes
        f32_rows["algorithmic_bytes_per_launch"] = int(fb)
        f32_rows["roofline_frac"] = round(fb / (f32_rows["launch_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
        f32_rows["roofline_frac_alone"] = round(fb / (f32_rows["batch_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
    # this workload (calibrated on a copy kernel), committed under profiles/.
    def measured_traffic(rows):
        name = "router_c1_traffic.json" if rows == "u8" else "router_c1_traffic_f32.json"
        tpath = os.path.join(ROOT, "profiles", name)
        if os.path.exists(tpath) and args.n == 1_000_000 and nq == 10_000 and chosen == 200:
            tj = json.load(open(tpath))
            if tj.get("rows") == rows:
                return tj["traffic_bytes_per_launch"]
        return None

    traffic = measured_traffic("u8" if u8 else "f32")
    if f32_rows is not None:
        f32_rows["traffic"] = measured_traffic("f32")
        if f32_rows["traffic"]:
            # counter-based: bytes the kernel fetched from and wrote to memory
            # beyond L2 (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel)
            f32_rows["traffic_frac"] = round(f32_rows["traffic"] / (f32_rows["launch_ms"] / 1e3) / 1e9 /
                                             HBM_PEAK_GBS, 4)
    log("roofline: %.1f GB algorithmic per launch, %.1f GB/s (%.3f of peak)" % (alg_bytes / 1e9, achieved,
                                                                          achieved / HBM_PEAK_GBS))

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        orc.build()
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "c1.graph")
            idx.snapshot_hgra(p)
            g = orc.Graph.restore(p)
        cores = len(os.sched_getaffinity(0))
        # the first 1000 queries of the batch with their own topk (the same
        # mix of 10 and 100), one oracle batch_query per topk
        smp = np.arange(min(1000, nq))
        s10, s100 = smp[topk_q[smp] == 10], smp[topk_q[smp] == 100]
        t0 = time.time()
        reps = 0
        cv10 = cv100 = None
        ndcs = []
        while time.time() - t0 < args.cpu_seconds or reps == 0:
            cv10, _, n10 = orc.batch_query(g, base, orc.L2SQR, queries[s10], 10, chosen, nthreads=cores)
            cv100, _, n100 = orc.batch_query(g, base, orc.L2SQR, queries[s100], 100, chosen, nthreads=cores)
            ndcs = [n10, n100]
            reps += 1
        cpu_t = (time.time() - t0) / reps
        # parity of the sample against the batched GPU router, slot for slot
        match = float((np.sum(cv10 == pred[s10, :10]) + np.sum(cv100 == pred[s100, :100])) /
                      (cv10.size + cv100.size))
        cpu = {"value": round(smp.size / cpu_t, 1), "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": "the first %d queries of the same batch with their own topk (%d top-10, %d top-100) at "
                         "queue %d, %d repetitions (%.1fs), oracle restatement of "
                         "HierarchicalGraphRouter::batch_query with OpenMP over queries" % (
                             smp.size, s10.size, s100.size, chosen, reps, time.time() - t0),
               "cpu_model": host_cpu_model(), "numa": host_numa_policy(),
               "slot_match_vs_gpu": round(match, 6),
               "ndc_per_query": round(float(np.concatenate(ndcs).mean()), 1)}
        log("cpu baseline %.0f q/s on %d cores (%s), GPU slot match %.6f" % (cpu["value"], cores, cpu["cpu_model"],
                                                                           match))

    serving = None
    if args.serve_requests > 0:
        # Open-loop Poisson arrivals of the same queries at fractions of the
        # batched peak; continuous batching over overlapping lanes.
        from artea_amd.serving import Server, percentile_nearest_rank, poisson_arrivals

        srv = make_server(Server, idx, chosen, args)
        ns = args.serve_requests
        pick = np.arange(ns) % nq
        sq, stopk, squeue = queries[pick], topk_q[pick], np.full(ns, chosen, np.uint32)
        # Untimed warm-up: every lane's routers launch once (first launches
        # load kernels and size the persistent grids).
        nw = min(ns, args.serve_lanes * srv.budget)
        srv.run(sq[:nw], stopk[:nw], squeue[:nw], np.zeros(nw))
        # Closed-loop saturation: every request present at the start.
        _, _, _, wall = srv.run(sq, stopk, squeue, np.zeros(ns))
        sat = ns / wall
        log("serving saturation (all %d requests at once): %.0f q/s" % (ns, sat))
        loads = []
        same = True
        for i, frac in enumerate((0.25, 0.5, 0.75, 0.9)):
            rate = frac * qps
            v, d, lat, wall = srv.run(sq, stopk, squeue, poisson_arrivals(ns, rate, seed=seed + i))
            same &= bool(np.array_equal(v[:, :100][stopk == 100], pred[pick][stopk == 100]))
            loads.append({"offered_qps": round(rate), "achieved_qps": round(ns / wall),
                          "p50_ms": round(percentile_nearest_rank(lat, 50) * 1e3, 3),
                          "p99_ms": round(percentile_nearest_rank(lat, 99) * 1e3, 3)})
            log("serving at %.0f q/s offered: %.0f q/s achieved, p50 %.2f ms, p99 %.2f ms" % (
                rate, ns / wall, loads[-1]["p50_ms"], loads[-1]["p99_ms"]))
        serving = {"requests_per_load": ns, "lanes": args.serve_lanes, "lane_budget": srv.budget,
                   "admission": srv.admission(),
                   "queue": chosen, "saturation_qps": round(sat), "results_equal_batch": same, "loads": loads}

    out = {
        "metric": "QPS @ recall@100>=0.95 (1M x 128 L2), batched vs plain GPU; p99 ms",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "streams": len(lanes),
        "batch_ms": round(t_batch * 1e3, 4),
        "plain_batch_ms": round(t_plain_batch * 1e3, 4),
        "streams_results_equal": lanes_equal,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8" if u8 else "f32",
        "data": "synthetic (seeded 256-component integer Gaussian mixture; index built on the GPU)",
        "config": {"workload": "batched search, 1,000,000 x 128 base, 10,000 queries, L2, topk mixed {10,100}",
                   "n": args.n, "dim": D, "nq": nq, "queue": chosen, "parallelism": "single GPU"},
        "plain_qps": round(qps_plain, 1),
        "batched_over_plain": round(qps / qps_plain, 3),
        "recall@100": next(s["recall@100"] for s in sweep if s["queue"] == chosen),
        "build_seconds": round(build_s, 2), "build_join": args.join,
        "sweep": sweep,
        "ground_truth": {"fp64_kernel_s": round(gt_fp64_s, 3), "bf16_mfma_s": round(gt_mfma_s, 3),
                         "identical": gt_same, "kept_pairs_per_query": round(gst.get("kept_pairs", 0) / nq, 1)},
        "router_variant": br.variant_names[br.variant()],
        "rows": ("one-byte rows (every base and query value a whole number in [0, 255]; integer distances equal "
                 "the FP32 chains bit for bit)" if u8 else "FP32 rows"),
        "f32_rows": f32_rows,
        "launch": ("HIP graph replay (counter-reset kernel + persistent router kernel); timed steps alternate over "
                   "%d streams, each with its own router and outputs; batch_ms = one batch alone" % len(lanes)),
        "distance_max_rel_err_vs_fp64": dist_err,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_frac": round(traffic / t_launch / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "kernel": "router_persistent_kernel (batched router), one launch per step",
                     "algorithmic_bytes_per_launch": int(alg_bytes), "launch_ms": round(t_launch * 1e3, 4),
                     "launch_ms_source": "HIP events around each launch of the timed steps on its own stream "
                                         "(%d launches, %d streams)" % (len(launch_ms), len(lanes)),
                     "frac_alone": round(alg_bytes / t_batch / 1e9 / HBM_PEAK_GBS, 4),
                     "frac_per_step": round(alg_bytes / t_step / 1e9 / HBM_PEAK_GBS, 4),
                     "note": ("one-byte rows: a quarter of the FP32 rows' bytes per distance, so the kernel is bound "
                              "by its dependent row-id -> vector -> merge chain and instruction issue, not HBM; the "
                              "FP32-row kernel on the same batch is HBM-bound (f32_rows.roofline_frac)")
                     if u8 else "FP32 rows: gather-bound on HBM"},
        "cpu_baseline": cpu,
        "p99_ms": serving["loads"][2]["p99_ms"] if serving else None,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

// Hierarchical graph router on MI355X: one wavefront per query.
//
// Semantics follow HierarchicalGraphRouter::_query
// (router/hierarchical_graph_router.hpp:195-248): greedy best-improvement
// walks on levels top..1 (router/detail/beam_loop.hpp:129-166), then the L0
// beam search with StdCandidateQueue (router/data_structures/
// std_candidate_queue.hpp:358-513) and results by (distance, vid).
//
// This file is the host side: kernel choice by metric family, list size and
// row width (pick_kernels), launch sizing (LDS per wave, occupancy-sized
// persistent grids), the batched router's variant slots, autotuner and
// launch graphs, and the C-ABI entry points.  The device design is described
// in router_device.hpp.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <memory>
#include <string>

#include "router_host.hpp"
#include "router_kernels.hpp"

using namespace artea;
using namespace artea::rk;

namespace {

// Clears the launch counters with a kernel rather than a memset node, so a
// captured launch sequence replays correctly.
__global__ void reset_kernel(unsigned long long* stats, uint32_t* counter) {
    if (threadIdx.x < 4) stats[threadIdx.x] = 0;
    if (threadIdx.x == 4) *reinterpret_cast<uint32_t*>(stats + 4) = 0;  // overflow list count
    if (threadIdx.x == 0 && counter) *counter = 0;
}

// Empty-graph launch: every slot gets the invalid entry.
__global__ void fill_invalid_kernel(uint32_t* vids, float* dists, size_t count) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) {
        vids[i] = kInvalidVid;
        dists[i] = FLT_MAX;
    }
}

constexpr uint32_t kMinTail = 16;

int metric_family(int metric) { return metric == ARTEA_DOT ? 1 : (metric == ARTEA_COSINE ? 2 : 0); }

// Kernel table for queue size L (candidate-list blocks K: L plus a tie tail
// of at least kMinTail entries), row chunks C and the index metric.
int pick_kernels(uint32_t L, uint32_t C, int metric, KernelSet* ks, uint32_t* K) {
    const uint32_t need = L + kMinTail;
    uint32_t k = 0;
    if (need <= 64) k = 1;
    else if (need <= 128) k = 2;
    else if (need <= 256) k = 4;
    else if (need <= 512) k = 8;
    else if (need <= 832) k = 13;
    else if (need <= 1024) k = 16;
    else {
        set_error("candidate_queue_size " + std::to_string(L) + " exceeds the device maximum of 1008");
        return ARTEA_EINVAL;
    }
    if (K) *K = k;
    if (!ks) return ARTEA_OK;
    switch (metric_family(metric) * 100 + int(k)) {
#define ARTEA_RK_CASE(MF, KK) \
    case MF * 100 + KK: *ks = kernels_mf##MF##_k##KK(C); break;
        ARTEA_RK_LIST(ARTEA_RK_CASE)
#undef ARTEA_RK_CASE
    default: set_error("router: no kernel for this queue size"); return ARTEA_EINVAL;
    }
    return ARTEA_OK;
}

// Persistent variant used until an autotune picks one: measured at
// 1,000,000 x 128 L2 on MI355X, the unbounded 32-row kernel is fastest up
// to 4 blocks and the 3-waves-per-SIMD budget above; with one-byte rows the
// unbounded kernel.
int default_variant(uint32_t K, bool u8) { return u8 ? kNumVariants : (K <= 4 ? kVarWide : kVarWide3); }

// Variant slots of the batched router: 0..3 the FP32-row kernels, 4..7 the
// one-byte-row kernels (only while the index keeps one-byte rows).
int num_slots(const artea_index* idx) { return idx->d_u8 ? 2 * kNumVariants : kNumVariants; }

size_t lds_bytes(uint32_t D, uint32_t K, bool u8) { return kWavesPerBlock * wave_lds_bytes(D, K, u8); }

}  // namespace

struct artea_batched_router {
    artea_router base;
    uint32_t max_topk = 0, max_queue = 0;
};

// A captured launch sequence keeps its own copy of the early-stop
// thresholds; it shares the router's work counter and counters, so its
// replays are ordered with the router's other launches.
struct artea_launch_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    artea_router* r = nullptr;
    float* d_thr = nullptr;
};

namespace {

int router_init(artea_router* r, artea_index* idx, uint32_t topk, uint32_t queue) {
    r->idx = idx;
    r->topk = topk;
    r->queue = queue;
    ARTEA_HIP_TRY(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    ARTEA_HIP_TRY(hipMalloc(&r->d_stats, kStatsBytes));
    ARTEA_HIP_TRY(hipMemset(r->d_stats, 0, kStatsBytes));  // counters read before any launch are zero
    ARTEA_HIP_TRY(hipMalloc(&r->d_thr, 64 * sizeof(float)));
    ARTEA_HIP_TRY(hipMalloc(&r->d_counter, 64));
    ARTEA_HIP_TRY(hipEventCreateWithFlags(&r->last, hipEventDisableTiming));
    return ARTEA_OK;
}

void router_free(artea_router* r) {
    if (r->last_stream) (void)hipEventSynchronize(r->last);
    if (r->last) (void)hipEventDestroy(r->last);
    if (r->d_stats) (void)hipFree(r->d_stats);
    if (r->d_thr) (void)hipFree(r->d_thr);
    if (r->d_counter) (void)hipFree(r->d_counter);
    if (r->stream) (void)hipStreamDestroy(r->stream);
}

// Orders a launch on stream s after the router's previous one.
int order_after_last(artea_router* r, hipStream_t s) {
    if (r->last_stream && r->last_stream != s) ARTEA_HIP_TRY(hipStreamWaitEvent(s, r->last, 0));
    return ARTEA_OK;
}

int record_last(artea_router* r, hipStream_t s) {
    ARTEA_HIP_TRY(hipEventRecord(r->last, s));
    r->last_stream = s;
    return ARTEA_OK;
}

// Common launch path of both routers.  The graph must be attached.
// thr_mode: 0 = enqueue the thresholds' upload on s, 1 = upload them
// synchronously to d_thr (before a capture), 2 = they are already in d_thr
// (inside a capture: no ordering event either, the capture is preceded by a
// synchronisation).  d_thr = nullptr: the router's own buffer.
int launch_body(artea_router* r, const float* d_q, uint32_t nq, uint32_t topk, uint32_t queue, const uint32_t* d_topk,
                const uint32_t* d_queue, uint32_t out_stride, uint32_t flags, const float* thresholds,
                uint32_t* d_ov, float* d_od, hipStream_t s, bool persistent, int thr_mode, float* d_thr);

// The kernel a launch of queue size `queue` runs (the pinned or default slot
// of the batched router, or the plain kernel) and its LDS bytes.
int select_kernel(const artea_router* r, uint32_t queue, bool persistent, KernelFn* fn, size_t* lds) {
    const artea_index* idx = r->idx;
    KernelSet ks;
    uint32_t K;
    int rc = pick_kernels(queue, idx->dim_pad / 16, idx->metric, &ks, &K);
    if (rc) return rc;
    const bool has_u8 = idx->d_u8 != nullptr;
    int slot = r->variant >= 0 ? r->variant : default_variant(K, has_u8);
    if (!has_u8 && slot >= kNumVariants) slot -= kNumVariants;  // the one-byte rows were dropped
    const bool u8 = slot >= kNumVariants;
    const int variant = slot % kNumVariants;
    *fn = persistent ? (u8 ? ks.persistent_u8[variant] : ks.persistent[variant]) : (has_u8 ? ks.plain_u8 : ks.plain);
    const bool fn_u8 = (persistent ? u8 : has_u8) && ks.u8_kernels;  // the kernel's LDS layout
    *lds = lds_bytes(idx->dim_pad, K, fn_u8);
    return ARTEA_OK;
}

int launch(artea_router* r, const float* d_q, uint32_t nq, uint32_t topk, uint32_t queue, const uint32_t* d_topk,
           const uint32_t* d_queue, uint32_t out_stride, uint32_t flags, const float* thresholds, uint32_t* d_ov,
           float* d_od, hipStream_t s, bool persistent, int thr_mode = 0, float* d_thr = nullptr) {
    if (thr_mode == 2)
        return launch_body(r, d_q, nq, topk, queue, d_topk, d_queue, out_stride, flags, thresholds, d_ov, d_od, s,
                           persistent, thr_mode, d_thr);
    int rc = order_after_last(r, s);
    if (rc) return rc;
    rc = launch_body(r, d_q, nq, topk, queue, d_topk, d_queue, out_stride, flags, thresholds, d_ov, d_od, s,
                     persistent, thr_mode, d_thr);
    const int rc2 = record_last(r, s);  // also after a failed launch: whatever was enqueued stays ordered
    return rc ? rc : rc2;
}

int launch_body(artea_router* r, const float* d_q, uint32_t nq, uint32_t topk, uint32_t queue, const uint32_t* d_topk,
                const uint32_t* d_queue, uint32_t out_stride, uint32_t flags, const float* thresholds,
                uint32_t* d_ov, float* d_od, hipStream_t s, bool persistent, int thr_mode, float* d_thr) {
    artea_index* idx = r->idx;
    if (!d_thr) d_thr = r->d_thr;
    ARTEA_CHECK_ARG(idx->dg.attached, "router: no graph attached to the index");
    ARTEA_CHECK_ARG((flags & ~uint32_t(ARTEA_EARLY_STOP | ARTEA_L0_ONLY)) == 0, "router: unknown query flags");
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, r->d_stats, persistent ? r->d_counter : nullptr);
    ARTEA_HIP_TRY(hipGetLastError());
    if (nq == 0) return ARTEA_OK;
    if (idx->n == 0) {
        const size_t cnt = size_t(nq) * out_stride;
        fill_invalid_kernel<<<dim3(uint32_t((cnt + 255) / 256)), dim3(256), 0, s>>>(d_ov, d_od, cnt);
        ARTEA_HIP_TRY(hipGetLastError());
        return ARTEA_OK;
    }
    ARTEA_CHECK_ARG(idx->dg.entry < idx->n, "router: graph has no valid entry point");
    const DeviceGraph& g = idx->dg;
    if (flags & ARTEA_EARLY_STOP) {
        ARTEA_CHECK_ARG(thresholds != nullptr, "router: early stop needs per-level thresholds");
        ARTEA_CHECK_ARG(g.top < 64, "router: too many levels");
        for (uint32_t h = 1; h <= g.top; ++h) {
            const float t = thresholds[h];
            if (!(t > 0.0f) || !(t <= FLT_MAX)) {
                set_error("Early-stop threshold is not representable in query distance units");
                return ARTEA_EOVERFLOW;
            }
        }
        if (thr_mode == 1) {
            ARTEA_HIP_TRY(hipStreamSynchronize(s));
            ARTEA_HIP_TRY(hipMemcpy(d_thr, thresholds, sizeof(float) * (g.top + 1), hipMemcpyHostToDevice));
        } else if (thr_mode == 0) {
            ARTEA_HIP_TRY(hipMemcpyAsync(d_thr, thresholds, sizeof(float) * (g.top + 1), hipMemcpyHostToDevice, s));
        }
    }
    RouterArgs a{};
    a.vecs = idx->d_vecs;
    a.vecs_u8 = reinterpret_cast<const uint32_t*>(idx->d_u8);
    a.n = idx->n;
    a.D = idx->dim_pad;
    a.C = idx->dim_pad / 16;
    a.metric = idx->metric;
    a.ell0 = g.ell0;
    a.stride0 = g.stride0;
    a.vinfo = g.vinfo;
    a.ellU = g.ellU;
    a.strideU = g.strideU;
    a.top = g.top;
    a.entry = g.entry;
    a.queries = d_q;
    a.nq = nq;
    a.topk = topk;
    a.queue = queue;
    a.topk_arr = d_topk;
    a.queue_arr = d_queue;
    a.out_stride = out_stride;
    a.flags = flags;
    a.thresholds = d_thr;
    a.out_vids = d_ov;
    a.out_dists = d_od;
    a.stats = r->d_stats;
    a.work_counter = nullptr;
    {
        uint32_t B = 13;
        while (B < 32 && (uint64_t(1) << B) < idx->n) ++B;
        ARTEA_CHECK_ARG(B <= 28, "router: more than 2^28 vectors");
        a.hmask = uint32_t((uint64_t(1) << B) - 1);
        a.vis_mode = B <= 20 ? 0u : (B <= 27 ? 1u : 2u);
    }
    KernelFn fn = nullptr;
    size_t lds = 0;
    int rc = select_kernel(r, queue, persistent, &fn, &lds);
    if (rc) return rc;
    if (lds > 65536) {
        // above the default dynamic-LDS limit (large queues, wide rows)
        ARTEA_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          int(lds)));
    }
    if (persistent) {
        a.work_counter = r->d_counter;
        if (r->occ_fn != reinterpret_cast<const void*>(fn) || r->occ_lds != lds) {
            int dev = 0, cus = 0, per_cu = 0;
            ARTEA_HIP_TRY(hipGetDevice(&dev));
            ARTEA_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            ARTEA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWavesPerBlock * 64, lds));
            r->occ_fn = reinterpret_cast<const void*>(fn);
            r->occ_lds = lds;
            r->occ_blocks = uint32_t(std::max(1, cus * std::max(per_cu, 1)));
        }
        uint32_t blocks = r->occ_blocks;
        blocks = std::min(blocks, (nq + kWavesPerBlock - 1) / kWavesPerBlock);
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(kWavesPerBlock * 64), lds, s, a);
    } else {
        const uint32_t blocks = (nq + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(kWavesPerBlock * 64), lds, s, a);
    }
    ARTEA_HIP_TRY(hipGetLastError());
    return ARTEA_OK;
}

int read_stats(artea_router* r, uint64_t* ndc, uint64_t* exps, uint64_t* ovf) {
    if (r->last_stream) ARTEA_HIP_TRY(hipEventSynchronize(r->last));
    unsigned long long h[4];
    ARTEA_HIP_TRY(hipMemcpy(h, r->d_stats, sizeof h, hipMemcpyDeviceToHost));
    if (ndc) *ndc = h[0];
    if (exps) *exps = h[1];
    if (ovf) *ovf = h[2];
    return ARTEA_OK;
}

int read_overflows(artea_router* r, uint32_t max, uint32_t* qidx, uint32_t* count) {
    ARTEA_CHECK_ARG(count && (max == 0 || qidx), "last_overflows: null argument");
    if (r->last_stream) ARTEA_HIP_TRY(hipEventSynchronize(r->last));
    uint32_t list[1 + kOverflowList];
    ARTEA_HIP_TRY(hipMemcpy(list, r->d_stats + 4, sizeof list, hipMemcpyDeviceToHost));
    *count = list[0];
    const uint32_t listed = std::min(std::min(list[0], kOverflowList), max);
    std::copy(list + 1, list + 1 + listed, qidx);
    std::sort(qidx, qidx + listed);
    return ARTEA_OK;
}

}  // namespace

namespace artea {

int batched_router_resident_waves(artea_batched_router* br, uint32_t* waves) {
    KernelFn fn = nullptr;
    size_t lds = 0;
    int rc = select_kernel(&br->base, br->base.queue, true, &fn, &lds);
    if (rc) return rc;
    if (lds > 65536)
        ARTEA_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          int(lds)));
    int dev = 0, cus = 0, per_cu = 0;
    ARTEA_HIP_TRY(hipGetDevice(&dev));
    ARTEA_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    ARTEA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWavesPerBlock * 64, lds));
    *waves = uint32_t(std::max(1, cus * std::max(per_cu, 1)) * kWavesPerBlock);
    return ARTEA_OK;
}

int router_launch(artea_router* r, const float* d_q, uint32_t nq, uint32_t topk, uint32_t queue,
                  const uint32_t* d_topk, const uint32_t* d_queue, uint32_t out_stride, uint32_t flags,
                  const float* thresholds, uint32_t* d_ov, float* d_od, hipStream_t s, bool persistent, int thr_mode,
                  float* d_thr) {
    return launch_body(r, d_q, nq, topk, queue, d_topk, d_queue, out_stride, flags, thresholds, d_ov, d_od, s,
                       persistent, thr_mode, d_thr);
}

int router_check_queue(uint32_t queue, uint32_t dim_pad, int metric) {
    return pick_kernels(queue, dim_pad / 16, metric, nullptr, nullptr);
}

uint32_t rk_overflow_list() { return kOverflowList; }

int router_overflow_list_async(artea_batched_router* r, uint32_t* host_dst, hipStream_t stream) {
    int rc = order_after_last(&r->base, stream);
    if (rc) return rc;
    ARTEA_HIP_TRY(hipMemcpyAsync(host_dst, r->base.d_stats + 4, (1 + kOverflowList) * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, stream));
    return ARTEA_OK;
}

}  // namespace artea

extern "C" {

int artea_router_create(artea_index* idx, uint32_t topk, uint32_t queue, artea_router** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(idx && out, "artea_router_create: null argument");
    if (queue < topk) {
        set_error("candidate_queue_size (" + std::to_string(queue) + ") must be >= topk (" + std::to_string(topk) + ")");
        return ARTEA_EINVAL;
    }
    int rc = pick_kernels(queue, idx->dim_pad / 16, idx->metric, nullptr, nullptr);
    if (rc) return rc;
    auto r = std::make_unique<artea_router>();
    rc = router_init(r.get(), idx, topk, queue);
    if (rc) {
        router_free(r.get());
        return rc;
    }
    *out = r.release();
    return ARTEA_OK;
}

int artea_router_destroy(artea_router* r) {
    if (!r) return ARTEA_OK;
    router_free(r);
    delete r;
    return ARTEA_OK;
}

int artea_router_batch_query_device(artea_router* r, const float* d_q, uint32_t nq, uint32_t flags,
                                    const float* thresholds, uint32_t* d_ov, float* d_od, void* stream) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_router_batch_query_device: null router");
    ARTEA_CHECK_ARG(nq == 0 || (d_q && d_ov && d_od), "artea_router_batch_query_device: null buffer");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : r->stream;
    return launch(r, d_q, nq, r->topk, r->queue, nullptr, nullptr, r->topk, flags, thresholds, d_ov, d_od, s, false);
}

int artea_router_batch_query(artea_router* r, const float* q, uint32_t nq, uint32_t flags, const float* thresholds,
                             uint32_t* ov, float* od) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_router_batch_query: null router");
    ARTEA_CHECK_ARG(nq == 0 || (q && ov && od), "artea_router_batch_query: null buffer");
    const size_t qb = size_t(nq) * r->idx->dim_pad * 4, rb = size_t(nq) * r->topk;
    float* dq = nullptr;
    uint32_t* dv = nullptr;
    float* dd = nullptr;
    int rc = ARTEA_OK;
    auto cleanup = [&]() {
        if (dq) (void)hipFree(dq);
        if (dv) (void)hipFree(dv);
        if (dd) (void)hipFree(dd);
    };
    if (nq == 0) return ARTEA_OK;
    if (hipMalloc(&dq, qb) != hipSuccess || hipMalloc(&dv, rb * 4 + 4) != hipSuccess ||
        hipMalloc(&dd, rb * 4 + 4) != hipSuccess) {
        cleanup();
        set_error("artea_router_batch_query: device allocation failed");
        return ARTEA_ENOMEM;
    }
    if (hipMemcpyAsync(dq, q, qb, hipMemcpyHostToDevice, r->stream) != hipSuccess) {
        cleanup();
        set_error("artea_router_batch_query: query upload failed");
        return ARTEA_EHIP;
    }
    rc = launch(r, dq, nq, r->topk, r->queue, nullptr, nullptr, r->topk, flags, thresholds, dv, dd, r->stream, false);
    if (rc == ARTEA_OK) {
        if (hipMemcpyAsync(ov, dv, rb * 4, hipMemcpyDeviceToHost, r->stream) != hipSuccess ||
            hipMemcpyAsync(od, dd, rb * 4, hipMemcpyDeviceToHost, r->stream) != hipSuccess ||
            hipStreamSynchronize(r->stream) != hipSuccess) {
            set_error("artea_router_batch_query: result download failed");
            rc = ARTEA_EHIP;
        }
    } else {
        (void)hipStreamSynchronize(r->stream);
    }
    cleanup();
    return rc;
}

int artea_router_last_stats(artea_router* r, uint64_t* ndc, uint64_t* exps, uint64_t* ovf) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_router_last_stats: null router");
    return read_stats(r, ndc, exps, ovf);
}

int artea_batched_router_create(artea_index* idx, uint32_t max_topk, uint32_t max_queue, artea_batched_router** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(idx && out, "artea_batched_router_create: null argument");
    if (max_queue < max_topk) {
        set_error("candidate_queue_size (" + std::to_string(max_queue) + ") must be >= topk (" +
                  std::to_string(max_topk) + ")");
        return ARTEA_EINVAL;
    }
    int rc = pick_kernels(max_queue, idx->dim_pad / 16, idx->metric, nullptr, nullptr);
    if (rc) return rc;
    auto r = std::make_unique<artea_batched_router>();
    rc = router_init(&r->base, idx, max_topk, max_queue);
    if (rc) {
        router_free(&r->base);
        return rc;
    }
    r->max_topk = max_topk;
    r->max_queue = max_queue;
    *out = r.release();
    return ARTEA_OK;
}

int artea_batched_router_destroy(artea_batched_router* r) {
    if (!r) return ARTEA_OK;
    router_free(&r->base);
    delete r;
    return ARTEA_OK;
}

int artea_batched_router_query_device(artea_batched_router* r, const float* d_q, uint32_t nq, const uint32_t* d_topk,
                                      const uint32_t* d_queue, uint32_t flags, const float* thresholds,
                                      uint32_t* d_ov, float* d_od, void* stream) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_batched_router_query_device: null router");
    ARTEA_CHECK_ARG(nq == 0 || (d_q && d_ov && d_od && d_topk && d_queue),
                    "artea_batched_router_query_device: null buffer");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : r->base.stream;
    return launch(&r->base, d_q, nq, r->max_topk, r->max_queue, d_topk, d_queue, r->max_topk, flags, thresholds, d_ov,
                  d_od, s, true);
}

int artea_batched_router_set_variant(artea_batched_router* r, int variant) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_batched_router_set_variant: null router");
    ARTEA_CHECK_ARG(variant >= -1 && variant < num_slots(r->base.idx), "artea_batched_router_set_variant: unknown variant");
    r->base.variant = variant;
    return ARTEA_OK;
}

int artea_batched_router_get_variant(artea_batched_router* r, int* variant) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r && variant, "artea_batched_router_get_variant: null argument");
    uint32_t K = 0;
    int rc = pick_kernels(r->max_queue, r->base.idx->dim_pad / 16, r->base.idx->metric, nullptr, &K);
    if (rc) return rc;
    const bool has_u8 = r->base.idx->d_u8 != nullptr;
    int slot = r->base.variant >= 0 ? r->base.variant : default_variant(K, has_u8);
    if (!has_u8 && slot >= kNumVariants) slot -= kNumVariants;
    *variant = slot;
    return ARTEA_OK;
}

int artea_batched_router_autotune(artea_batched_router* r, const float* d_q, uint32_t nq, const uint32_t* d_topk,
                                  const uint32_t* d_queue, uint32_t flags, const float* thresholds, uint32_t* d_ov,
                                  float* d_od, float* ms, int* chosen) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_batched_router_autotune: null router");
    ARTEA_CHECK_ARG(nq > 0 && d_q && d_ov && d_od && d_topk && d_queue, "artea_batched_router_autotune: empty or null batch");
    hipStream_t s = r->base.stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ARTEA_HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("artea_batched_router_autotune: hipEventCreate failed");
        return ARTEA_EHIP;
    }
    const int saved = r->base.variant;
    int best = -1, rc = ARTEA_OK;
    float best_ms = 0.f;
    const int slots = num_slots(r->base.idx);
    for (int v = 0; v < slots && rc == ARTEA_OK; ++v) {
        r->base.variant = v;
        // One untimed launch (occupancy query, caches), then one timed launch.
        rc = launch(&r->base, d_q, nq, r->max_topk, r->max_queue, d_topk, d_queue, r->max_topk, flags, thresholds, d_ov,
                    d_od, s, true);
        if (rc) break;
        float t = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess ||
            (rc = launch(&r->base, d_q, nq, r->max_topk, r->max_queue, d_topk, d_queue, r->max_topk, flags,
                         thresholds, d_ov, d_od, s, true)) != ARTEA_OK ||
            hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            if (rc == ARTEA_OK) {
                set_error("artea_batched_router_autotune: timing failed");
                rc = ARTEA_EHIP;
            }
            break;
        }
        if (ms) ms[v] = t;
        if (best < 0 || t < best_ms) {
            best = v;
            best_ms = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        r->base.variant = saved;
        return rc;
    }
    r->base.variant = best;
    if (chosen) *chosen = best;
    return ARTEA_OK;
}

int artea_batched_router_capture(artea_batched_router* r, const float* d_q, uint32_t nq, const uint32_t* d_topk,
                                 const uint32_t* d_queue, uint32_t flags, const float* thresholds, uint32_t* d_ov,
                                 float* d_od, artea_launch_graph** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r && out, "artea_batched_router_capture: null argument");
    ARTEA_CHECK_ARG(nq == 0 || (d_q && d_ov && d_od && d_topk && d_queue), "artea_batched_router_capture: null buffer");
    auto g = std::unique_ptr<artea_launch_graph, int (*)(artea_launch_graph*)>(new artea_launch_graph{},
                                                                                 artea_launch_graph_destroy);
    g->r = &r->base;
    ARTEA_HIP_TRY(hipMalloc(&g->d_thr, 64 * sizeof(float)));
    hipStream_t s = r->base.stream;
    // One eager launch sizes the persistent grid (occupancy query) and puts
    // the early-stop thresholds into the graph's own buffer, so the captured
    // sequence holds kernel nodes only: the counter reset and the router.
    int rc = launch(&r->base, d_q, nq, r->max_topk, r->max_queue, d_topk, d_queue, r->max_topk, flags, thresholds, d_ov,
                    d_od, s, true, 1, g->d_thr);
    if (rc) return rc;
    ARTEA_HIP_TRY(hipStreamSynchronize(s));
    ARTEA_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = launch(&r->base, d_q, nq, r->max_topk, r->max_queue, d_topk, d_queue, r->max_topk, flags, thresholds, d_ov,
                d_od, s, true, 2, g->d_thr);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) {
        set_error(std::string("hipStreamEndCapture failed: ") + hipGetErrorString(ec));
        return ARTEA_EHIP;
    }
    g->graph = graph;
    const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        set_error(std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ei));
        return ARTEA_EHIP;
    }
    *out = g.release();
    return ARTEA_OK;
}

int artea_launch_graph_launch(artea_launch_graph* g, void* stream) {
    artea::clear_error();
    ARTEA_CHECK_ARG(g && g->exec, "artea_launch_graph_launch: null graph");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = order_after_last(g->r, s);
    if (rc) return rc;
    ARTEA_HIP_TRY(hipGraphLaunch(g->exec, s));
    return record_last(g->r, s);
}

int artea_launch_graph_destroy(artea_launch_graph* g) {
    if (!g) return ARTEA_OK;
    if (g->r && g->r->last_stream) (void)hipEventSynchronize(g->r->last);  // replays may still read d_thr
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_thr) (void)hipFree(g->d_thr);
    delete g;
    return ARTEA_OK;
}

int artea_batched_router_last_stats(artea_batched_router* r, uint64_t* ndc, uint64_t* exps, uint64_t* ovf) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_batched_router_last_stats: null router");
    return read_stats(&r->base, ndc, exps, ovf);
}

int artea_router_last_overflows(artea_router* r, uint32_t max, uint32_t* qidx, uint32_t* count) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_router_last_overflows: null router");
    return read_overflows(r, max, qidx, count);
}

int artea_batched_router_last_overflows(artea_batched_router* r, uint32_t max, uint32_t* qidx, uint32_t* count) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_batched_router_last_overflows: null router");
    return read_overflows(&r->base, max, qidx, count);
}


}  // extern "C"

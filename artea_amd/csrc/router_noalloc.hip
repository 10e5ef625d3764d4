// Allocation-free batched query: artea_batched_query_noalloc runs the
// batched router's persistent kernels (router.hip, router_device.hpp) for
// callers that cannot allocate device memory, keeping a launch's counters,
// work counter and early-stop thresholds in a fixed array of this library
// instead of a router's buffers.
// Layout: [0, kStatsBytes) the counter block (counters, then the overflow
// list), kScratchCounter the work counter, kScratchThr 64 thresholds.  In
// host code `scratch` only names the symbol: its device address, one per
// GPU, comes from hipGetSymbolAddress (noalloc_scratch).
#include <algorithm>
#include <mutex>
#include <string>

#include "router_host.hpp"
#include "router_kernels.hpp"

using namespace artea;
using namespace artea::rk;

namespace {

constexpr size_t kScratchCounter = (kStatsBytes + 63) / 64 * 64;
constexpr size_t kScratchThr = kScratchCounter + 64;
constexpr size_t kScratchBytes = 2048;
constexpr uint32_t kScratchLevels = 64;
static_assert(kScratchThr + kScratchLevels * sizeof(float) <= kScratchBytes, "scratch layout");

}  // namespace

// One copy per GPU, held until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int noalloc_scratch(unsigned char** out) {
    int dev = 0;
    ARTEA_HIP_TRY(hipGetDevice(&dev));
    ARTEA_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "artea_batched_query_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        ARTEA_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        ARTEA_CHECK_ARG(ptr != nullptr, "artea_batched_query_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return ARTEA_OK;
}

struct ThrPack {
    float t[kScratchLevels];
};

// Writes the thresholds from the kernel's arguments (no host-to-device copy).
__global__ void set_thresholds_kernel(ThrPack p, float* dst, uint32_t count) {
    if (threadIdx.x < count) dst[threadIdx.x] = p.t[threadIdx.x];
}

}  // namespace


extern "C" {

int artea_batched_query_noalloc(artea_index* idx, uint32_t max_topk, uint32_t max_queue, const float* d_q, uint32_t nq,
                                const uint32_t* d_topk, const uint32_t* d_queue, uint32_t flags,
                                const float* thresholds, uint32_t* d_ov, float* d_od, void* stream) {
    artea::clear_error();
    ARTEA_CHECK_ARG(idx, "artea_batched_query_noalloc: null index");
    if (max_queue < max_topk) {
        set_error("candidate_queue_size (" + std::to_string(max_queue) + ") must be >= topk (" +
                  std::to_string(max_topk) + ")");
        return ARTEA_EINVAL;
    }
    int rc = router_check_queue(max_queue, idx->dim_pad, idx->metric);
    if (rc) return rc;
    ARTEA_CHECK_ARG(nq == 0 || (d_q && d_ov && d_od && d_topk && d_queue), "artea_batched_query_noalloc: null buffer");
    if ((flags & ARTEA_EARLY_STOP) && idx->dg.attached && idx->dg.top >= kScratchLevels) {
        set_error("artea_batched_query_noalloc: " + std::to_string(idx->dg.top + 1) +
                  " levels of early-stop thresholds exceed the scratch array's " + std::to_string(kScratchLevels));
        return ARTEA_EINVAL;
    }
    unsigned char* base = nullptr;
    rc = noalloc_scratch(&base);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // A router on the stack over the scratch array (it owns nothing; the
    // persistent grid is sized by an occupancy query on every call).
    artea_router r;
    r.idx = idx;
    r.topk = max_topk;
    r.queue = max_queue;
    r.stream = s;
    r.d_stats = reinterpret_cast<unsigned long long*>(base);
    r.d_counter = reinterpret_cast<uint32_t*>(base + kScratchCounter);
    r.d_thr = reinterpret_cast<float*>(base + kScratchThr);
    if ((flags & ARTEA_EARLY_STOP) && thresholds && idx->dg.attached && idx->n > 0) {
        ThrPack p{};
        const uint32_t count = idx->dg.top + 1;
        std::copy(thresholds, thresholds + count, p.t);
        hipLaunchKernelGGL(set_thresholds_kernel, dim3(1), dim3(kScratchLevels), 0, s, p, r.d_thr, count);
        ARTEA_HIP_TRY(hipGetLastError());
    }
    return router_launch(&r, d_q, nq, max_topk, max_queue, d_topk, d_queue, max_topk, flags, thresholds, d_ov, d_od,
                         s, true, 2, r.d_thr);
}

}  // extern "C"

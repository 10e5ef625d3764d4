// State of one layer under the propagate engine, shared by the refine
// kernels (refine.hip) and the routing loops (routing.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace artea {
namespace rf {

constexpr uint32_t kNew = 0x80000000u;
constexpr uint32_t kMask = 0x7FFFFFFFu;
constexpr uint32_t kMaxCap = 128;  // triangle / merge LDS sizing

struct Ctx {
    const float* vecs;
    uint32_t n_global, D, C;
    int metric;
    uint32_t n, cap;
    const uint32_t* l2g;  // null = identity
    const uint32_t* g2l;  // null = identity; 0xFFFFFFFF = not in layer
    uint2* rows;          // n * cap
    uint32_t* cnt;        // n
    uint32_t* e_t;        // n * cap emission targets
    uint2* e_vd;          // n * cap emission (vid, dist bits)
    uint32_t* e_cnt;      // n
    uint32_t* t_cnt;      // n + 1
    uint32_t* t_off;      // n + 1
    uint32_t* t_fill;     // n
    uint2* seg;           // n * cap
    unsigned long long* stats;  // [0] logs, [1] FP32 pair distances, [2] BF16 pair distances
    uint32_t list_cap;          // triangle recheck queue entries per row (<= kListCap)
    uint32_t rb, re;            // rows the triangle / reverse producers run on (next_range); [0, n) otherwise
};

__device__ inline uint32_t storage_of(const Ctx& c, uint32_t local) { return c.l2g ? c.l2g[local] : local; }
__device__ inline uint32_t local_of(const Ctx& c, uint32_t g) {
    return c.g2l ? c.g2l[g] : (g < c.n ? g : 0xFFFFFFFFu);
}


// Routing loops (routing.hip).  A routing pass writes, per local row, the
// row merged with the warm-start router's results and pruned
// (RoutingUpdater, merge_logs, PruningUpdater) into ext (ext_stride slots
// per row, up to cap + topk entries) and ext_cnt; it reads the rows only.
struct RoutingPass {
    uint32_t topk, queue;        // router results per pivot, candidate queue
    float inv_scale;             // PruningUpdater's 1 / scale
    uint2* ext;                  // n * ext_stride
    uint32_t* ext_cnt;           // n
    uint32_t ext_stride;         // >= cap + topk
    unsigned long long* stats;   // device: [0] distances, [1] expansions, [2] overflows, [3] list count,
                                 // then up to kRouteOverflowList row indices
};
constexpr uint32_t kRouteOverflowList = 1024;
constexpr size_t kRouteStatsBytes = 4 * sizeof(unsigned long long) + kRouteOverflowList * sizeof(uint32_t);

// Launches the routing pass over every row; returns a HIP error code.
// Rows whose candidate list dropped an entry tied with its bound are run
// again with a longer list; *unresolved counts those that still did.
hipError_t routing_pass(const Ctx& c, const RoutingPass& p, hipStream_t s, int dev_cus, uint32_t* unresolved);
// Copies the first cap entries of every ext row into the rows; with
// `reverse`, also counts the ReverseUpdater logs of the full ext rows into
// t_cnt (targets outside the layer go to the dead bucket n).
hipError_t routing_finish(const Ctx& c, const RoutingPass& p, bool reverse, hipStream_t s);
// Places those reverse logs into `seg` at t_off (t_fill cleared).
hipError_t routing_scatter(const Ctx& c, const RoutingPass& p, uint2* seg, hipStream_t s);
// PruningUpdater over the rows [rb, re) in place (no logs).
hipError_t prune_rows(const Ctx& c, float inv_scale, uint32_t rb, uint32_t re, hipStream_t s, int dev_cus);

}  // namespace rf
}  // namespace artea

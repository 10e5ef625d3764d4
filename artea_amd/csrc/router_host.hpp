// Host-side router state shared by the router's translation units
// (router.hip: the C-ABI routers; other units that launch the same kernels
// over their own buffers).
#pragma once

#include "common.hpp"

// A router's launches share its work counter and counter block, so they run
// one after another whichever streams they are enqueued on: every launch
// records `last` on its stream, and a launch on another stream first waits
// for it (use one router per stream for launches that overlap).
struct artea_router {
    artea_index* idx = nullptr;
    uint32_t topk = 0, queue = 0;
    hipStream_t stream = nullptr;
    unsigned long long* d_stats = nullptr;
    float* d_thr = nullptr;
    uint32_t* d_counter = nullptr;
    hipEvent_t last = nullptr;        // completion of the latest launch
    hipStream_t last_stream = nullptr;  // its stream (nullptr: none yet)
    bool persistent = false;
    const void* occ_fn = nullptr;  // cached persistent grid size for (kernel, LDS bytes)
    size_t occ_lds = 0;
    uint32_t occ_blocks = 0;
    int variant = -1;  // persistent kernel variant (-1: default_variant)
};

namespace artea {

// The routers' launch path: the counter reset, the early-stop thresholds,
// kernel choice by queue size and row kind, and the plain or persistent
// launch on stream s.  thr_mode: 0 = enqueue the thresholds' upload on s into
// d_thr, 1 = upload them synchronously, 2 = they are already in d_thr (no
// upload).  d_thr = nullptr: the router's own buffer.  Nothing is allocated.
int router_launch(artea_router* r, const float* d_q, uint32_t nq, uint32_t topk, uint32_t queue,
                  const uint32_t* d_topk, const uint32_t* d_queue, uint32_t out_stride, uint32_t flags,
                  const float* thresholds, uint32_t* d_ov, float* d_od, hipStream_t s, bool persistent, int thr_mode,
                  float* d_thr);

// ARTEA_OK when the kernels cover queue size `queue` at this row width and
// metric, else ARTEA_EINVAL with the router's message.
int router_check_queue(uint32_t queue, uint32_t dim_pad, int metric);

}  // namespace artea

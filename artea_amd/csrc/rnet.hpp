// R-net insertion state on the device (rnet.hip), used by the artea_graph
// build in refine.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "artea_hip.h"

struct artea_rnet {
    uint32_t n = 0, D = 0, C = 0, L = 0;
    int metric = 0;
    const float* d_vecs = nullptr;
    artea_rnet_params p{};
    std::vector<uint2*> rows;     // per level, n * cap (null: level 0 without insert_on_l0)
    std::vector<uint32_t*> cnt;   // per level, n
    std::vector<uint32_t> highest;  // host copy, ARTEA_INVALID_LEVEL until assigned
    uint32_t top = 0xFFFFFFFFu;
    uint32_t waves = 0;
    uint64_t distances = 0, expansions = 0, unresolved = 0;
    hipStream_t s = nullptr;
    bool own_stream = false;
};

namespace artea {
// Insertion of every vector of d_vecs (n x dim_pad rows) in waves; metric is
// the build metric.  Returns an ARTEA_* status (message set on failure).
int rnet_run(const float* d_vecs, uint32_t n, uint32_t dim_pad, int metric, const artea_rnet_params& p,
             hipStream_t s, artea_rnet** out);
void rnet_free(artea_rnet* r);
}  // namespace artea

// Shared host/device declarations of libartea_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include "host.hpp"

#define ARTEA_HIP_TRY(expr)                                                                   \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::artea::set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));      \
            return ARTEA_EHIP;                                                                \
        }                                                                                     \
    } while (0)


// Device layout of an attached graph (the MI355X-side form of the compact
// CSR): L0 rows as a padded ELL array indexed directly by vid, so a beam
// expansion is one coalesced row read with no pointer chasing; upper-level
// rows as a second ELL array addressed through vinfo[v] = {highest,
// first upper row}.  Row order inside every row equals the CSR order.
struct DeviceGraph {
    uint32_t top = 0, entry = 0, ul_max = 0, bl_max = 0;
    uint32_t stride0 = 16, strideU = 16;  // ELL row strides (multiples of 16)
    uint32_t* ell0 = nullptr;             // n * stride0
    uint2* vinfo = nullptr;               // n
    uint32_t* ellU = nullptr;             // rowsU * strideU
    uint64_t rowsU = 0;
    bool attached = false;
};

struct artea_index {
    uint32_t n = 0, dim_pad = 0;
    int metric = 0;
    float* d_vecs = nullptr;
    uint8_t* d_u8 = nullptr;  // 8-bit copy of d_vecs when every value is an integer in [0, 255]
    DeviceGraph dg;
    artea::HostCsr csr;        // host copy of a graph attached from a host CSR
    bool csr_current = false;  // csr describes dg (false after a device HGRA load)
    // Host-clock phases of the last device HGRA call: load = {file read and
    // copy to the device, segmented parse and walk, checks and ELL fill};
    // snapshot = {record layout on the device, copy back and file write, 0}.
    double hgra_load_s[3] = {0, 0, 0}, hgra_store_s[3] = {0, 0, 0};
    hipStream_t stream = nullptr;
};

namespace artea {
// Replaces the index's device graph (the previous one is freed).
void replace_device_graph(artea_index* idx, const DeviceGraph& g);
// HGRA v1 restore straight into the device layout / snapshot from it
// (hgra_device.hip).  segment_words = 0 and trial_records = 0 pick the defaults.
int hgra_load_device(artea_index* idx, const char* path, uint32_t segment_words, uint32_t trial_records);
int hgra_store_device(artea_index* idx, const char* path);
// Host grouped CSR of the device graph (local ids in vid order per group).
int graph_from_device(const artea_index* idx, HostCsr* out);
// Checks a grouped CSR, uploads its device form and keeps it as the index's
// host graph (moved in, no copy).
int attach_csr(artea_index* idx, HostCsr&& c);
// Router internals used by the server: the length of a launch's overflow
// list, and an asynchronous copy of that list (count, then batch indices)
// into host memory, ordered after the router's latest launch on `stream`.
uint32_t rk_overflow_list();
// Wavefronts the batched router's persistent kernel keeps resident on the
// current GPU (its LDS per wave and registers decide it).
int batched_router_resident_waves(artea_batched_router* r, uint32_t* waves);
int router_overflow_list_async(artea_batched_router* r, uint32_t* host_dst, hipStream_t stream);
}  // namespace artea

// Device plumbing of the C-ABI: allocations, copies, streams, events, the
// copy kernel that calibrates HBM byte counters, device information.
#include <cfloat>
#include <cstring>
#include <string>

#include "common.hpp"

using namespace artea;

namespace {

// Plain vectorised copy (16 B per lane per step): the reference kernel that
// calibrates HBM byte counters (a copy of B bytes moves exactly B in, B out).
__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

extern "C" {

int artea_copy_device(void* d_dst, const void* d_src, size_t bytes, void* stream) {
    clear_error();
    ARTEA_CHECK_ARG(bytes % 16 == 0, "artea_copy_device: bytes must be a multiple of 16");
    if (!bytes) return ARTEA_OK;
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint4*>(d_src), static_cast<uint4*>(d_dst), bytes / 16);
    ARTEA_HIP_TRY(hipGetLastError());
    return ARTEA_OK;
}

int artea_device_alloc(size_t bytes, void** d_ptr) {
    clear_error();
    ARTEA_CHECK_ARG(d_ptr, "artea_device_alloc: null output");
    ARTEA_HIP_TRY(hipMalloc(d_ptr, bytes ? bytes : 1));
    return ARTEA_OK;
}
int artea_device_free(void* d_ptr) {
    clear_error();
    if (d_ptr) ARTEA_HIP_TRY(hipFree(d_ptr));
    return ARTEA_OK;
}
int artea_memcpy_htod(void* d_dst, const void* src, size_t bytes, void* stream) {
    clear_error();
    if (!bytes) return ARTEA_OK;
    if (stream) {
        ARTEA_HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
    } else {
        ARTEA_HIP_TRY(hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice));
    }
    return ARTEA_OK;
}
int artea_memcpy_dtoh(void* dst, const void* d_src, size_t bytes, void* stream) {
    clear_error();
    if (!bytes) return ARTEA_OK;
    if (stream) {
        ARTEA_HIP_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
    } else {
        ARTEA_HIP_TRY(hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost));
    }
    return ARTEA_OK;
}
int artea_memset_device(void* d_dst, int value, size_t bytes, void* stream) {
    clear_error();
    if (!bytes) return ARTEA_OK;
    ARTEA_HIP_TRY(hipMemsetAsync(d_dst, value, bytes, static_cast<hipStream_t>(stream)));
    if (!stream) ARTEA_HIP_TRY(hipStreamSynchronize(nullptr));
    return ARTEA_OK;
}
int artea_stream_create(void** stream) {
    clear_error();
    ARTEA_CHECK_ARG(stream, "artea_stream_create: null output");
    hipStream_t s;
    ARTEA_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return ARTEA_OK;
}
int artea_stream_destroy(void* stream) {
    clear_error();
    if (stream) ARTEA_HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return ARTEA_OK;
}
int artea_stream_synchronize(void* stream) {
    clear_error();
    ARTEA_HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return ARTEA_OK;
}
int artea_event_create(void** event) {
    clear_error();
    ARTEA_CHECK_ARG(event, "artea_event_create: null output");
    hipEvent_t e;
    ARTEA_HIP_TRY(hipEventCreate(&e));
    *event = e;
    return ARTEA_OK;
}
int artea_event_destroy(void* event) {
    clear_error();
    if (event) ARTEA_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return ARTEA_OK;
}
int artea_event_record(void* event, void* stream) {
    clear_error();
    ARTEA_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return ARTEA_OK;
}
int artea_event_elapsed_ms(void* start, void* end, float* ms) {
    clear_error();
    ARTEA_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(end)));
    ARTEA_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
    return ARTEA_OK;
}
int artea_event_query(void* event, int* done) {
    clear_error();
    ARTEA_CHECK_ARG(event && done, "artea_event_query: null argument");
    const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
    if (e == hipSuccess) {
        *done = 1;
        return ARTEA_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return ARTEA_OK;
    }
    set_error(std::string("hipEventQuery failed: ") + hipGetErrorString(e));
    return ARTEA_EHIP;
}
int artea_host_alloc(size_t bytes, void** h_ptr) {
    clear_error();
    ARTEA_CHECK_ARG(h_ptr, "artea_host_alloc: null output");
    ARTEA_HIP_TRY(hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return ARTEA_OK;
}
int artea_host_free(void* h_ptr) {
    clear_error();
    if (h_ptr) ARTEA_HIP_TRY(hipHostFree(h_ptr));
    return ARTEA_OK;
}
int artea_device_info(char* name, size_t len, int* cus, size_t* total) {
    clear_error();
    int dev = 0;
    ARTEA_HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t p;
    ARTEA_HIP_TRY(hipGetDeviceProperties(&p, dev));
    if (name && len) {
        strncpy(name, p.gcnArchName, len - 1);
        name[len - 1] = 0;
    }
    if (cus) *cus = p.multiProcessorCount;
    if (total) *total = p.totalGlobalMem;
    return ARTEA_OK;
}

}  // extern "C"

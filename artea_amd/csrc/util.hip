// Device plumbing of the C-ABI (allocations, copies, streams, events) and
// the FP64 brute-force k-NN used as the recall oracle.
#include <cfloat>
#include <cstring>
#include <string>

#include "common.hpp"

using namespace artea;

namespace {

constexpr int kQB = 16;   // queries per block
constexpr int kTB = 32;   // base rows per tile
constexpr int kMaxK2 = 288;

// One block scores kQB queries against the whole base in FP64, keeping per
// query the best k2 = k + 32 candidates sorted by (distance, vid) in LDS.
// Thread (q = tid / 16, t = tid % 16) scores base rows t and t + 16 of each
// tile.  Candidates below a query's current bound are staged, then one
// thread per query inserts them in order (tile scan order keeps vid ties
// ascending).
__global__ __launch_bounds__(256) void bruteforce_kernel(int metric, const float* __restrict__ base, uint32_t n,
                                                         const float* __restrict__ queries, uint32_t nq, uint32_t D,
                                                         uint32_t k, uint32_t k2, uint32_t* __restrict__ out_v,
                                                         double* __restrict__ out_d) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* qv = reinterpret_cast<float*>(smem);                   // kQB * D
    float* xt = qv + kQB * D;                                     // kTB * D
    double* ld = reinterpret_cast<double*>(xt + kTB * D);         // kQB * k2
    double* cd = ld + kQB * k2;                                   // kQB * kTB
    uint32_t* lv = reinterpret_cast<uint32_t*>(cd + kQB * kTB);   // kQB * k2
    uint32_t* cv = lv + kQB * k2;                                 // kQB * kTB
    uint32_t* ln = cv + kQB * kTB;                                // kQB
    uint32_t* cn = ln + kQB;                                      // kQB
    const uint32_t tid = threadIdx.x, q = tid / 16, t = tid % 16;
    const uint32_t q0 = blockIdx.x * kQB;
    for (uint32_t i = tid; i < kQB * D; i += 256) {
        const uint32_t qq = q0 + i / D;
        qv[i] = qq < nq ? queries[size_t(qq) * D + (i % D)] : 0.f;
    }
    if (tid < kQB) {
        ln[tid] = 0;
        cn[tid] = 0;
    }
    __syncthreads();
    double qn = 0;
    for (uint32_t d = 0; d < D; ++d) qn = fma(double(qv[q * D + d]), double(qv[q * D + d]), qn);
    for (uint32_t t0 = 0; t0 < n; t0 += kTB) {
        for (uint32_t i = tid; i < kTB * D; i += 256) {
            const uint32_t b = t0 + i / D;
            xt[i] = b < n ? base[size_t(b) * D + (i % D)] : 0.f;
        }
        __syncthreads();
        const double bound = ln[q] < k2 ? INFINITY : ld[q * k2 + k2 - 1];
        const uint32_t bound_v = ln[q] < k2 ? 0xFFFFFFFFu : lv[q * k2 + k2 - 1];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t r = t + 16 * j, b = t0 + r;
            if (b >= n || q0 + q >= nq) continue;
            double s = 0, dt = 0, xn = 0;
            const float* x = xt + r * D;
            const float* y = qv + q * D;
            if (metric == 0 || metric == 3) {
                for (uint32_t d = 0; d < D; ++d) {
                    const double df = double(y[d]) - double(x[d]);
                    s = fma(df, df, s);
                }
                if (metric == 3) s = sqrt(s);
            } else {
                for (uint32_t d = 0; d < D; ++d) {
                    dt = fma(double(y[d]), double(x[d]), dt);
                    xn = fma(double(x[d]), double(x[d]), xn);
                }
                if (metric == 1) {
                    s = -dt;
                } else {
                    const double den = sqrt(qn * xn);
                    s = den > 0 ? 1.0 - dt / den : 1.0;
                }
            }
            if (s < bound || (s == bound && b < bound_v)) {
                const uint32_t slot = atomicAdd(cn + q, 1u);
                cd[q * kTB + slot] = s;
                cv[q * kTB + slot] = b;
            }
        }
        __syncthreads();
        if (t == 0) {
            const uint32_t m = cn[q];
            // stage in vid order so equal distances keep ascending vids
            for (uint32_t a = 1; a < m; ++a)
                for (uint32_t c = a; c > 0 && cv[q * kTB + c] < cv[q * kTB + c - 1]; --c) {
                    const double td = cd[q * kTB + c];
                    cd[q * kTB + c] = cd[q * kTB + c - 1];
                    cd[q * kTB + c - 1] = td;
                    const uint32_t tv = cv[q * kTB + c];
                    cv[q * kTB + c] = cv[q * kTB + c - 1];
                    cv[q * kTB + c - 1] = tv;
                }
            uint32_t len = ln[q];
            for (uint32_t a = 0; a < m; ++a) {
                const double s = cd[q * kTB + a];
                const uint32_t b = cv[q * kTB + a];
                if (len == k2 && !(s < ld[q * k2 + k2 - 1] || (s == ld[q * k2 + k2 - 1] && b < lv[q * k2 + k2 - 1])))
                    continue;
                uint32_t pos = len < k2 ? len : k2 - 1;
                while (pos > 0 && (s < ld[q * k2 + pos - 1] || (s == ld[q * k2 + pos - 1] && b < lv[q * k2 + pos - 1]))) {
                    ld[q * k2 + pos] = ld[q * k2 + pos - 1];
                    lv[q * k2 + pos] = lv[q * k2 + pos - 1];
                    --pos;
                }
                ld[q * k2 + pos] = s;
                lv[q * k2 + pos] = b;
                if (len < k2) ++len;
            }
            ln[q] = len;
            cn[q] = 0;
        }
        __syncthreads();
    }
    if (q0 + q < nq)
        for (uint32_t i = t; i < k; i += 16) {
            out_v[size_t(q0 + q) * k + i] = i < ln[q] ? lv[q * k2 + i] : kInvalidVid;
            out_d[size_t(q0 + q) * k + i] = i < ln[q] ? ld[q * k2 + i] : INFINITY;
        }
}

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

int artea_bruteforce_knn_device(int metric, const float* d_base, uint32_t n, const float* d_queries, uint32_t nq,
                                uint32_t D, uint32_t k, uint32_t* d_out_vids, double* d_out_dists, void* stream) {
    clear_error();
    ARTEA_CHECK_ARG(metric >= 0 && metric <= 3, "artea_bruteforce_knn_device: metric must be 0..3");
    ARTEA_CHECK_ARG(D > 0 && D % 16 == 0 && D <= 1024, "artea_bruteforce_knn_device: dim_pad must be a multiple of 16, <= 1024");
    ARTEA_CHECK_ARG(k >= 1 && k <= 256, "artea_bruteforce_knn_device: k must be in [1, 256]");
    if (nq == 0) return ARTEA_OK;
    ARTEA_CHECK_ARG(d_base && d_queries && d_out_vids && d_out_dists, "artea_bruteforce_knn_device: null buffer");
    const uint32_t k2 = k + 32;
    const size_t lds = size_t(kQB) * D * 4 + size_t(kTB) * D * 4 + size_t(kQB) * k2 * 12 + size_t(kQB) * kTB * 12 +
                       kQB * 8;
    int dev = 0, max_lds = 65536;
    ARTEA_HIP_TRY(hipGetDevice(&dev));
    ARTEA_HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    ARTEA_CHECK_ARG(lds <= size_t(max_lds), "artea_bruteforce_knn_device: k and dim_pad exceed the LDS budget");
    if (lds > 65536)
        ARTEA_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(bruteforce_kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL(bruteforce_kernel, dim3((nq + kQB - 1) / kQB), dim3(256), lds, static_cast<hipStream_t>(stream),
                       metric, d_base, n, d_queries, nq, D, k, k2, d_out_vids, d_out_dists);
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

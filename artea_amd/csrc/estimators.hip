// Scoring and dataset probes on the GPU: Recall@K
// (utils/recall_estimator.hpp:24-230), ADR@1 (utils/adr_estimator.hpp),
// and the dataset prober's base k-NN distance table, query-to-ground-truth
// distances and farthest distances (utils/dataset_prober.hpp:253-587), which
// give the L0 minimum distance a build is parameterised with.
//
// Distances are dist_func_t's FP32 chains (device.hpp quad_distance, bit for
// bit); ratios are formed in FP64 as the reference does.  Samples are given
// by the caller (the reference draws them from MKL streams).
#include <algorithm>
#include <cfloat>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"

using namespace artea;
using namespace artea::dev;

namespace {

constexpr uint32_t kProbeK = 128;     // DatasetProber::MAX_K
constexpr uint32_t kRecallMaxK = 1024;

// Bitonic sort of s[0..P) ascending (P a power of two <= 2048), 256 threads.
template <typename T>
__device__ void block_bitonic(T* s, uint32_t P) {
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < P / 2; t += blockDim.x) {
                const uint32_t i = 2 * j * (t / j) + (t % j);
                const uint32_t l = i + j;
                const bool up = (i & k) == 0;
                const T a = s[i], b = s[l];
                if ((a > b) == up) {
                    s[i] = b;
                    s[l] = a;
                }
            }
            __syncthreads();
        }
    }
}

__host__ __device__ inline uint32_t pow2_at_least(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Recall@K: one block per query.  Both rows are sorted in LDS (padding
// sorts last, so the first k entries are the row's values); a prediction
// counts when it differs from its sorted predecessor (std::unique) and
// occurs among the k ground-truth ids (std::set_intersection against the
// sorted ground truth).
__global__ __launch_bounds__(256) void recall_kernel(const uint32_t* __restrict__ pred, uint32_t pstride,
                                                     const uint32_t* __restrict__ gt, uint32_t gstride, uint32_t k,
                                                     unsigned long long* __restrict__ matches) {
    __shared__ uint32_t sp[kRecallMaxK], sg[kRecallMaxK];
    __shared__ uint32_t s_cnt;
    const uint32_t q = blockIdx.x, P = pow2_at_least(k);
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        sp[i] = i < k ? pred[size_t(q) * pstride + i] : 0xFFFFFFFFu;
        sg[i] = i < k ? gt[size_t(q) * gstride + i] : 0xFFFFFFFFu;
    }
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    block_bitonic(sp, P);
    block_bitonic(sg, P);
    uint32_t mine = 0;
    for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
        const uint32_t v = sp[j];
        if (j > 0 && sp[j - 1] == v) continue;
        uint32_t lo = 0, hi = k;  // first gt index with sg >= v
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (sg[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        mine += (lo < k && sg[lo] == v) ? 1u : 0u;
    }
    atomicAdd(&s_cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(matches, static_cast<unsigned long long>(s_cnt));
}

// out[i * k + j] = dist(a_i, b[ids[i * stride + j]]); a quad of lanes per pair.
// An id at or beyond nb (the invalid vertex of a padded row) gives +inf.
__global__ __launch_bounds__(256) void distance_rows_kernel(int metric, const float* __restrict__ a, uint32_t na,
                                                            const float* __restrict__ b, uint32_t nb,
                                                            const uint32_t* __restrict__ ids,
                                                            uint32_t stride, uint32_t k, uint32_t D,
                                                            float* __restrict__ out) {
    const uint64_t pair = uint64_t(blockIdx.x) * 64 + threadIdx.x / 4;
    const uint32_t sub = threadIdx.x & 3u;
    const bool ok = pair < uint64_t(na) * k;
    const uint32_t i = ok ? uint32_t(pair / k) : 0, j = ok ? uint32_t(pair % k) : 0;
    const float* x = a + size_t(i) * D;
    const uint32_t id = ok ? ids[size_t(i) * stride + j] : 0u;
    const bool in = id < nb;
    const float* y = b + size_t(in ? id : 0u) * D;
    const float qn2 = metric == kCosine ? quad_norm2(x, int(D / 16), sub) : 0.f;
    const float d = quad_distance<0>(x, y, int(D / 16), metric, sub, qn2);
    if (ok && sub == 0) out[size_t(i) * k + j] = in ? d : INFINITY;
}

constexpr uint32_t kProbeG = 8;       // samples per block of the k-NN probe
constexpr uint32_t kProbeBuf = 1024;  // candidate buffer per sample (best kProbeK first)

// The kProbeK smallest dist(sample, j) over every j != sample, ascending, as
// a multiset of values (the reference keeps them in a max-heap; ties do not
// change the multiset), for kProbeG samples per block.  Each lane takes one
// base row per pass (256 rows) and runs dist_func_t's 16 accumulator chains
// for all kProbeG samples in its own registers, so every row is read once
// for the kProbeG samples; the sample rows are wave-uniform.  A value enters
// its sample's candidate buffer when it is below that sample's current
// kProbeK-th smallest (+inf until kProbeK are held); a buffer that could
// overflow in the next pass is sorted and cut back to its best kProbeK.
//
// The rows are cut into gridDim.y slices of `slice` rows (a block per sample
// group and slice, so a few hundred samples still fill the chip); each block
// writes each sample's slice best (padded with +inf) to
// out[(s * gridDim.y + y) * kProbeK ..], which probe_merge_kernel reduces.
template <int METRIC>
__global__ __launch_bounds__(256) void probe_knn_kernel(const float* __restrict__ base, uint32_t n, uint32_t D,
                                                        const uint32_t* __restrict__ samples, uint32_t ns,
                                                        uint32_t slice, float* __restrict__ out) {
    __shared__ float buf[kProbeG][kProbeBuf];
    __shared__ uint32_t s_fill[kProbeG];
    __shared__ float s_thr[kProbeG];
    const uint32_t C = D / 16;
    const uint32_t r_begin = blockIdx.y * slice, r_end = min(n, r_begin + slice);
    uint32_t vid[kProbeG];
    const float* q[kProbeG];
    float qn2[kProbeG];
#pragma unroll
    for (uint32_t g = 0; g < kProbeG; ++g) {
        vid[g] = samples[min(blockIdx.x * kProbeG + g, ns - 1)];
        q[g] = base + size_t(vid[g]) * D;
        qn2[g] = 0.f;
        if (METRIC == kCosine) {
            float a[16] = {};
            for (uint32_t c = 0; c < C; ++c)
#pragma unroll
                for (int l = 0; l < 16; ++l) a[l] = __builtin_fmaf(q[g][c * 16 + l], q[g][c * 16 + l], a[l]);
            qn2[g] = reduce16_lane(a);
        }
    }
    if (threadIdx.x < kProbeG) {
        s_fill[threadIdx.x] = 0;
        s_thr[threadIdx.x] = INFINITY;
    }
    __syncthreads();
    // Entered by every thread with the same g and fill f (read between barriers).
    auto compact = [&](uint32_t g, uint32_t f) {
        const uint32_t P = max(pow2_at_least(f), kProbeK);
        for (uint32_t i = f + threadIdx.x; i < P; i += blockDim.x) buf[g][i] = INFINITY;
        __syncthreads();
        block_bitonic(buf[g], P);
        if (threadIdx.x == 0) {
            s_fill[g] = min(f, kProbeK);
            if (f >= kProbeK) s_thr[g] = buf[g][kProbeK - 1];
        }
        __syncthreads();
    };
    for (uint32_t row0 = r_begin; row0 < r_end; row0 += 256) {
        const uint32_t j = row0 + threadIdx.x;
        const bool ok = j < r_end;
        const float* x = base + size_t(ok ? j : r_begin) * D;
        float acc[kProbeG][16], xn[16];
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            xn[l] = 0.f;
#pragma unroll
            for (uint32_t g = 0; g < kProbeG; ++g) acc[g][l] = 0.f;
        }
        for (uint32_t c = 0; c < C; ++c) {
            float xv[16];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float4 t = *reinterpret_cast<const float4*>(x + c * 16 + v * 4);
                xv[4 * v] = t.x;
                xv[4 * v + 1] = t.y;
                xv[4 * v + 2] = t.z;
                xv[4 * v + 3] = t.w;
            }
#pragma unroll
            for (uint32_t g = 0; g < kProbeG; ++g) {
#pragma unroll
                for (int l = 0; l < 16; ++l) {
                    const float qv = q[g][c * 16 + l];
                    if (METRIC == kL2Sqr || METRIC == kL2) {
                        const float d = qv - xv[l];
                        acc[g][l] = __builtin_fmaf(d, d, acc[g][l]);
                    } else {
                        acc[g][l] = __builtin_fmaf(qv, xv[l], acc[g][l]);
                    }
                }
            }
            if (METRIC == kCosine) {
#pragma unroll
                for (int l = 0; l < 16; ++l) xn[l] = __builtin_fmaf(xv[l], xv[l], xn[l]);
            }
        }
        const float sq2 = METRIC == kCosine ? reduce16_lane(xn) : 0.f;
#pragma unroll
        for (uint32_t g = 0; g < kProbeG; ++g) {
            float d = reduce16_lane(acc[g]);
            if (METRIC == kL2) d = __builtin_sqrtf(d);
            if (METRIC == kDot) d = -d;
            if (METRIC == kCosine) {
                const float denom = __builtin_sqrtf(qn2[g] * sq2);
                d = denom > 0.0f ? 1.0f - d / denom : 1.0f;
            }
            if (ok && j != vid[g] && d < s_thr[g]) {
                const uint32_t at = atomicAdd(&s_fill[g], 1u);
                buf[g][at] = d;  // at < kProbeBuf: compacted below whenever 256 more could overflow
            }
        }
        __syncthreads();
        // Every thread reads the fills before any thread appends again.
        uint32_t f[kProbeG];
#pragma unroll
        for (uint32_t g = 0; g < kProbeG; ++g) f[g] = s_fill[g];
        __syncthreads();
        for (uint32_t g = 0; g < kProbeG; ++g)
            if (f[g] + 256 > kProbeBuf) compact(g, f[g]);
    }
    for (uint32_t g = 0; g < kProbeG; ++g) compact(g, s_fill[g]);
    for (uint32_t g = 0; g < kProbeG; ++g) {
        const uint32_t s = blockIdx.x * kProbeG + g;
        if (s >= ns) break;
        float* o = out + (size_t(s) * gridDim.y + blockIdx.y) * kProbeK;
        for (uint32_t i = threadIdx.x; i < kProbeK; i += blockDim.x) o[i] = buf[g][i];
    }
}

constexpr uint32_t kMergeBuf = 8192;  // slices * kProbeK values merged per sample (slices <= 64)

// Per sample: the best kProbeK of its `slices` partial rows (slices *
// kProbeK <= kMergeBuf values), ascending.
__global__ __launch_bounds__(256) void probe_merge_kernel(const float* __restrict__ part, uint32_t slices,
                                                          float* __restrict__ out) {
    __shared__ float buf[kMergeBuf];
    const uint32_t s = blockIdx.x, f = slices * kProbeK, P = pow2_at_least(f);
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) buf[i] = i < f ? part[size_t(s) * f + i] : INFINITY;
    __syncthreads();
    block_bitonic(buf, P);
    for (uint32_t i = threadIdx.x; i < kProbeK; i += blockDim.x) out[size_t(s) * kProbeK + i] = buf[i];
}

// Per sample (one block): max over j != sample of dist(sample, j).
__global__ __launch_bounds__(256) void probe_farthest_kernel(int metric, const float* __restrict__ base, uint32_t n,
                                                             uint32_t D, const uint32_t* __restrict__ samples,
                                                             float* __restrict__ out) {
    __shared__ float s_max[4];
    const uint32_t s = blockIdx.x, vid = samples[s];
    const uint32_t sub = threadIdx.x & 3u, C = D / 16;
    const float* q = base + size_t(vid) * D;
    const float qn2 = metric == kCosine ? quad_norm2(q, int(C), sub) : 0.f;
    float m = -FLT_MAX;  // numeric_limits<distance_t>::lowest()
    for (uint32_t row0 = 0; row0 < n; row0 += 64) {
        const uint32_t j = row0 + threadIdx.x / 4;
        if (j < n && j != vid) m = fmaxf(m, quad_distance<0>(q, base + size_t(j) * D, int(C), metric, sub, qn2));
    }
    m = -wave_min(-m);
    if (lane_id() == 0) s_max[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[s] = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
}

// ADR@1 terms: ratio[q] = double(dist(q, retrieved)) / double(dist(q, nn))
// and used[q] = 1, or used[q] = 0 when the ground-truth distance is <= 0
// (the reference skips such queries) or either id is not a vertex.
__global__ __launch_bounds__(256) void adr_kernel(int metric, const float* __restrict__ base, uint32_t n,
                                                  const float* __restrict__ queries, uint32_t nq, uint32_t D,
                                                  const uint32_t* __restrict__ pred, uint32_t pstride,
                                                  const uint32_t* __restrict__ gt, uint32_t gstride,
                                                  double* __restrict__ ratio, uint8_t* __restrict__ used) {
    const uint32_t q = blockIdx.x * 64 + threadIdx.x / 4, sub = threadIdx.x & 3u;
    const bool ok = q < nq;
    const uint32_t r = ok ? pred[size_t(q) * pstride] : 0u, t = ok ? gt[size_t(q) * gstride] : 0u;
    const bool valid = ok && r < n && t < n;
    const float* x = queries + size_t(ok ? q : 0u) * D;
    const float qn2 = metric == kCosine ? quad_norm2(x, int(D / 16), sub) : 0.f;
    const float dr = quad_distance<0>(x, base + size_t(valid ? r : 0u) * D, int(D / 16), metric, sub, qn2);
    const float dt = quad_distance<0>(x, base + size_t(valid ? t : 0u) * D, int(D / 16), metric, sub, qn2);
    // The reference skips a query whose ground-truth distance is <= 0; a NaN
    // distance passes that test and its ratio enters the sum.
    if (ok && sub == 0) {
        const bool use = valid && !(dt <= 0.0f);
        ratio[q] = use ? double(dr) / double(dt) : 0.0;
        used[q] = use ? 1 : 0;
    }
}

int check_base(int metric, const void* base, uint32_t dim_pad, const char* who) {
    ARTEA_CHECK_ARG(base != nullptr, std::string(who) + ": null base");
    ARTEA_CHECK_ARG(metric >= 0 && metric <= 3, std::string(who) + ": metric must be 0..3");
    ARTEA_CHECK_ARG(dim_pad > 0 && dim_pad % 16 == 0, std::string(who) + ": dim_pad must be a positive multiple of 16");
    return ARTEA_OK;
}

// Device copy of host sample ids, checked against n.
int upload_samples(const uint32_t* samples, uint32_t ns, uint32_t n, hipStream_t s, uint32_t** d_out,
                   const char* who) {
    ARTEA_CHECK_ARG(samples != nullptr && ns >= 1, std::string(who) + ": num_samples must be at least 1");
    for (uint32_t i = 0; i < ns; ++i)
        if (samples[i] >= n) {
            set_error(std::string(who) + ": sample id " + std::to_string(samples[i]) + " is not below " +
                      std::to_string(n));
            return ARTEA_EINVAL;
        }
    ARTEA_HIP_TRY(hipMalloc(d_out, size_t(ns) * 4));
    const hipError_t e = hipMemcpyAsync(*d_out, samples, size_t(ns) * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        (void)hipFree(*d_out);
        *d_out = nullptr;
        set_error(std::string(who) + ": sample upload failed: " + hipGetErrorString(e));
        return ARTEA_EHIP;
    }
    return ARTEA_OK;
}

}  // namespace

extern "C" {

int artea_recall_at_k_device(const uint32_t* d_pred, uint32_t pred_stride, const uint32_t* d_gt, uint32_t gt_stride,
                             uint32_t nq, uint32_t k, uint64_t* matches, void* stream) {
    clear_error();
    ARTEA_CHECK_ARG(matches, "artea_recall_at_k_device: null output");
    ARTEA_CHECK_ARG(k >= 1 && k <= kRecallMaxK, "artea_recall_at_k_device: k must be in [1, 1024]");
    ARTEA_CHECK_ARG(pred_stride >= k, "artea_recall_at_k_device: prediction rows hold fewer than k ids");
    if (gt_stride < k) {
        set_error("Ground truth dimension " + std::to_string(gt_stride) + " is less than k=" + std::to_string(k));
        return ARTEA_EINVAL;
    }
    *matches = 0;
    if (nq == 0) return ARTEA_OK;
    ARTEA_CHECK_ARG(d_pred && d_gt, "artea_recall_at_k_device: null buffer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* d_m = nullptr;
    ARTEA_HIP_TRY(hipMalloc(&d_m, sizeof(unsigned long long)));
    unsigned long long h = 0;
    hipError_t e = hipMemsetAsync(d_m, 0, sizeof h, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(recall_kernel, dim3(nq), dim3(256), 0, s, d_pred, pred_stride, d_gt, gt_stride, k, d_m);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d_m, sizeof h, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_m);
    ARTEA_HIP_TRY(e);
    *matches = h;
    return ARTEA_OK;
}

int artea_distance_rows_device(int metric, const float* d_a, uint32_t na, const float* d_b, uint32_t nb,
                               const uint32_t* d_ids, uint32_t ids_stride, uint32_t k, uint32_t dim_pad, float* d_out,
                               void* stream) {
    clear_error();
    int rc = check_base(metric, d_b, dim_pad, "artea_distance_rows_device");
    if (rc) return rc;
    ARTEA_CHECK_ARG(ids_stride >= k, "artea_distance_rows_device: id rows hold fewer than k ids");
    if (na == 0 || k == 0) return ARTEA_OK;
    ARTEA_CHECK_ARG(d_a && d_ids && d_out, "artea_distance_rows_device: null buffer");
    ARTEA_CHECK_ARG(nb > 0, "artea_distance_rows_device: empty base");
    const uint64_t pairs = uint64_t(na) * k;
    ARTEA_CHECK_ARG(pairs <= uint64_t(0xFFFFFFFFu) * 64, "artea_distance_rows_device: too many pairs");
    hipLaunchKernelGGL(distance_rows_kernel, dim3(uint32_t((pairs + 63) / 64)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), metric, d_a, na, d_b, nb, d_ids, ids_stride, k, dim_pad, d_out);
    ARTEA_HIP_TRY(hipGetLastError());
    return ARTEA_OK;
}

int artea_probe_knn_device(int metric, const float* d_base, uint32_t n, uint32_t dim_pad, const uint32_t* samples,
                           uint32_t num_samples, float* knn, float* min_nearest, void* stream) {
    clear_error();
    int rc = check_base(metric, d_base, dim_pad, "artea_probe_knn_device");
    if (rc) return rc;
    if (n <= kProbeK) {
        set_error("Dataset must contain more than 128 vectors, got " + std::to_string(n));
        return ARTEA_EINVAL;
    }
    ARTEA_CHECK_ARG(knn != nullptr, "artea_probe_knn_device: null output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t* d_s = nullptr;
    rc = upload_samples(samples, num_samples, n, s, &d_s, "artea_probe_knn_device");
    if (rc) return rc;
    // Slices: enough blocks for several rounds over the chip (>= 4096, so the
    // last round's tail is short), slices of at least 4096 rows, at most
    // kMergeBuf / kProbeK of them per sample.
    const uint32_t groups = (num_samples + kProbeG - 1) / kProbeG;
    uint32_t slices = 1;
    while (slices < kMergeBuf / kProbeK && uint64_t(groups) * slices < 4096 && n / (2 * slices) >= 4096) slices *= 2;
    const uint32_t slice = (n + slices - 1) / slices;
    float* d_out = nullptr;
    const size_t out_bytes = size_t(num_samples) * kProbeK * 4;
    hipError_t e = hipMalloc(&d_out, out_bytes * (slices > 1 ? 1 + slices : 1));
    if (e == hipSuccess) {
        float* part = slices > 1 ? d_out + size_t(num_samples) * kProbeK : d_out;
        auto* kern = metric == kL2Sqr ? probe_knn_kernel<kL2Sqr>
                     : metric == kDot  ? probe_knn_kernel<kDot>
                     : metric == kCosine ? probe_knn_kernel<kCosine>
                                         : probe_knn_kernel<kL2>;
        hipLaunchKernelGGL(kern, dim3(groups, slices), dim3(256), 0, s, d_base, n, dim_pad, d_s, num_samples, slice,
                           part);
        e = hipGetLastError();
        if (e == hipSuccess && slices > 1) {
            hipLaunchKernelGGL(probe_merge_kernel, dim3(num_samples), dim3(256), 0, s, part, slices, d_out);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(knn, d_out, size_t(num_samples) * kProbeK * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_s);
    if (d_out) (void)hipFree(d_out);
    ARTEA_HIP_TRY(e);
    if (min_nearest) {
        float m = FLT_MAX;
        for (uint32_t i = 0; i < num_samples; ++i) m = std::min(m, knn[size_t(i) * kProbeK]);
        *min_nearest = m;
    }
    return ARTEA_OK;
}

int artea_probe_farthest_device(int metric, const float* d_base, uint32_t n, uint32_t dim_pad, const uint32_t* samples,
                                uint32_t num_samples, float* farthest, void* stream) {
    clear_error();
    int rc = check_base(metric, d_base, dim_pad, "artea_probe_farthest_device");
    if (rc) return rc;
    if (n < 2) {
        set_error("Dataset must contain at least 2 vectors, got " + std::to_string(n));
        return ARTEA_EINVAL;
    }
    ARTEA_CHECK_ARG(farthest != nullptr, "artea_probe_farthest_device: null output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t* d_s = nullptr;
    rc = upload_samples(samples, num_samples, n, s, &d_s, "artea_probe_farthest_device");
    if (rc) return rc;
    float* d_out = nullptr;
    hipError_t e = hipMalloc(&d_out, size_t(num_samples) * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(probe_farthest_kernel, dim3(num_samples), dim3(256), 0, s, metric, d_base, n, dim_pad, d_s,
                           d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(farthest, d_out, size_t(num_samples) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_s);
    if (d_out) (void)hipFree(d_out);
    ARTEA_HIP_TRY(e);
    return ARTEA_OK;
}

int artea_adr_at_1_device(int metric, const float* d_base, uint32_t n, const float* d_queries, uint32_t nq,
                          uint32_t dim_pad, const uint32_t* d_pred, uint32_t pred_stride, const uint32_t* d_gt,
                          uint32_t gt_stride, double* adr, uint64_t* counted, void* stream) {
    clear_error();
    int rc = check_base(metric, d_base, dim_pad, "artea_adr_at_1_device");
    if (rc) return rc;
    ARTEA_CHECK_ARG(adr != nullptr, "artea_adr_at_1_device: null output");
    ARTEA_CHECK_ARG(pred_stride >= 1 && gt_stride >= 1, "Predictions/GT dimensions must be >= 1 for ADR@1");
    *adr = 0.0;
    if (counted) *counted = 0;
    if (nq == 0) return ARTEA_OK;
    ARTEA_CHECK_ARG(d_queries && d_pred && d_gt, "artea_adr_at_1_device: null buffer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* d_r = nullptr;
    ARTEA_HIP_TRY(hipMalloc(&d_r, size_t(nq) * 9));
    uint8_t* d_used = reinterpret_cast<uint8_t*>(d_r + nq);
    std::vector<double> r(nq);
    std::vector<uint8_t> used(nq);
    hipLaunchKernelGGL(adr_kernel, dim3((nq + 63) / 64), dim3(256), 0, s, metric, d_base, n, d_queries, nq, dim_pad,
                       d_pred, pred_stride, d_gt, gt_stride, d_r, d_used);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(r.data(), d_r, size_t(nq) * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(used.data(), d_used, nq, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_r);
    ARTEA_HIP_TRY(e);
    // Sum in query order (the reference's reduction order is TBB's).
    double sum = 0.0;
    uint64_t c = 0;
    for (uint32_t i = 0; i < nq; ++i)
        if (used[i]) {
            sum += r[i];
            ++c;
        }
    if (counted) *counted = c;
    *adr = c ? sum / double(c) : 0.0;
    return ARTEA_OK;
}

}  // extern "C"

// Exact brute-force k-NN on the device (the recall ground truth): an FP64
// kernel, and a BF16 MFMA filter with FP64 rechecks that returns the FP64
// kernel's result.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"

using namespace artea;
using namespace artea::dev;

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


// ------------------------------------------------------------ BF16 MFMA path
// Exact top-k through a BF16 filter:
//  1. thresholds: the FP64 kernel on a prefix sample of the base gives, per
//     query, the sample's k-th distance T_q, an upper bound of the full
//     base's k-th distance;
//  2. filter: every (query, base) pair gets an approximate key from a BF16
//     MFMA GEMM with FP32 accumulation (two-term BF16 split of both operands,
//     L2 rows centred on the base mean first) and a proven error radius e;
//     every pair with key - e <= T_q is appended to the query's candidate list
//     (that set holds every pair with exact distance <= T_q, hence the true
//     top-k);
//  3. recheck: each candidate's distance is recomputed by the FP64 kernel's
//     own expression, and the candidates are sorted by (distance, vid).
// A query whose list overflows is recomputed by the FP64 kernel alone, so
// the result is always the FP64 kernel's result.

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t kCandCap = 4096;   // candidates kept per query
constexpr uint32_t kFilterQ = 128;    // queries per filter tile (4 waves x 32)

// Column sums in FP64 (per-block partials, then the mean in FP32).
__global__ void colsum_kernel(const float* vecs, uint32_t n, uint32_t D, uint32_t per_block, double* part) {
    const uint32_t b = blockIdx.x;
    const uint32_t lo = b * per_block, hi = min(n, lo + per_block);
    for (uint32_t d = threadIdx.x; d < D; d += blockDim.x) {
        double s = 0;
        for (uint32_t v = lo; v < hi; ++v) s += double(vecs[size_t(v) * D + d]);
        part[size_t(b) * D + d] = s;
    }
}
__global__ void colmean_kernel(const double* part, uint32_t blocks, uint32_t n, uint32_t D, float* out) {
    for (uint32_t d = threadIdx.x; d < D; d += blockDim.x) {
        double s = 0;
        for (uint32_t b = 0; b < blocks; ++b) s += part[size_t(b) * D + d];
        out[d] = float(s / double(n));
    }
}

// One wavefront per row: y = x - mu (FP32), split into BF16 hi / lo, stored
// per 16-coordinate chunk as [hi 0-7][lo 0-7][hi 8-15][lo 8-15] (4 dwords
// each, the MFMA operand order), plus |y|^2 and |y| in FP32.
__global__ __launch_bounds__(256) void split_kernel(const float* __restrict__ src, uint32_t rows, uint32_t D,
                                                    const float* __restrict__ mu, uint32_t* __restrict__ dst,
                                                    float* __restrict__ nrm2, float* __restrict__ nrm) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t groups = D / 8;
    for (uint32_t r = blockIdx.x * 4 + threadIdx.x / 64; r < rows; r += gridDim.x * 4) {
        float s2 = 0.f;
        if (lane < groups) {
            const float* x = src + size_t(r) * D + lane * 8;
            float v[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) v[t] = x[t] - (mu ? mu[lane * 8 + t] : 0.f);
            uint32_t hw[4], lw[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const f32x2 a = {v[2 * p], v[2 * p + 1]};
                const bf16x2 h = __builtin_convertvector(a, bf16x2);
                const f32x2 hf = __builtin_convertvector(h, f32x2);
                const f32x2 res = {a[0] - hf[0], a[1] - hf[1]};
                const bf16x2 l = __builtin_convertvector(res, bf16x2);
                hw[p] = __builtin_bit_cast(uint32_t, h);
                lw[p] = __builtin_bit_cast(uint32_t, l);
                s2 = __builtin_fmaf(v[2 * p], v[2 * p], s2);
                s2 = __builtin_fmaf(v[2 * p + 1], v[2 * p + 1], s2);
            }
            uint32_t* o = dst + size_t(r) * D + (lane / 2) * 16 + (lane & 1u) * 8;
            *reinterpret_cast<uint4*>(o) = make_uint4(hw[0], hw[1], hw[2], hw[3]);
            *reinterpret_cast<uint4*>(o + 4) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
        }
        for (int off = 32; off >= 1; off >>= 1) s2 += __shfl_xor(s2, off);
        if (lane == 0) {
            nrm2[r] = s2;
            nrm[r] = sqrtf(s2);
        }
    }
}

__device__ __forceinline__ float round_up_f32(double x) {
    float f = float(x);
    if (double(f) < x) f = nextafterf(f, INFINITY);
    return f;
}

// Per-query filter coefficients from the sample threshold T (FP64 distance
// of the sample's k-th neighbour).  A pair passes when
//   c0 * s + Pb[j] + Pq[i] + Qq[i] * |y_j| <= 0,
// which is  key - e <= T  written as one affine test (s = the MFMA dot
// product of the centred split rows, e = the pair's error radius):
//   L2:     key = |y_i|^2 + |y_j|^2 - 2s,  e = eps (2|y_i||y_j| + |y_i|^2 + |y_j|^2)
//   DOT:    key = -s,                      e = 2 eps |y_i||y_j|
//   COSINE: key = 1 - s / (|y_i||y_j|),    e = 4 eps
// T is rounded up to FP32 (squared first for L2); the factor 2 built into
// eps covers the FP32 roundings of the rearranged test.
__global__ void query_coef_kernel(int metric, const double* sample_d, uint32_t nq, uint32_t k, const float* qn2,
                                  const float* qn, float eps, float* qP, float* qQ) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double t64 = sample_d[size_t(q) * k + (k - 1)];
    const float t = round_up_f32(metric == kL2 ? t64 * t64 : t64);
    if (metric == kL2Sqr || metric == kL2) {
        qP[q] = (1.0f - eps) * qn2[q] - t;
        qQ[q] = -2.0f * eps * qn[q];
    } else if (metric == kDot) {
        qP[q] = -t;
        qQ[q] = -2.0f * eps * qn[q];
    } else {
        qP[q] = 0.0f;
        qQ[q] = (1.0f - 4.0f * eps - t) * qn[q];
    }
}

__global__ void base_coef_kernel(int metric, const float* xn2, uint32_t n, float eps, float* xP) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) xP[j] = (metric == kL2Sqr || metric == kL2) ? (1.0f - eps) * xn2[j] : 0.0f;
}

// Filter GEMM: base-stationary.  A workgroup holds TB split base rows in
// LDS and streams every 128-query tile past them; wave w computes queries
// [32w, 32w + 32) of the tile against the TB rows as TB/32 MFMA 32x32x16
// tiles (hi*hi + hi*lo + lo*hi into one FP32 accumulator), then appends the
// pairs that pass the affine test above.
template <int TB>
__global__ __launch_bounds__(256) void filter_kernel(const uint32_t* __restrict__ xs, const float* __restrict__ xP,
                                                     const float* __restrict__ xn, uint32_t n,
                                                     const uint32_t* __restrict__ qs, const float* __restrict__ qP,
                                                     const float* __restrict__ qQ, uint32_t nq, uint32_t D, float c0,
                                                     uint32_t* __restrict__ cand, uint32_t* __restrict__ cnt) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t stride = D + 4;  // dwords per LDS row (16-B aligned, rows 4 banks apart)
    uint32_t* sB = reinterpret_cast<uint32_t*>(smem);
    uint32_t* sQ = sB + TB * stride;
    float* sBP = reinterpret_cast<float*>(sQ + kFilterQ * stride);
    float* sBn = sBP + TB;
    float* sQP = sBn + TB;
    float* sQQ = sQP + kFilterQ;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid / 64;
    const uint32_t b0 = blockIdx.x * TB;
    const uint32_t chunks4 = D / 4;  // uint4 per row
    for (uint32_t i = tid; i < TB * chunks4; i += 256) {
        const uint32_t r = i / chunks4, c = i % chunks4;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (b0 + r < n) v = *reinterpret_cast<const uint4*>(xs + size_t(b0 + r) * D + c * 4);
        *reinterpret_cast<uint4*>(sB + r * stride + c * 4) = v;
    }
    for (uint32_t i = tid; i < TB; i += 256) {
        sBP[i] = b0 + i < n ? xP[b0 + i] : 0.f;
        sBn[i] = b0 + i < n ? xn[b0 + i] : 0.f;
    }
    for (uint32_t q0 = 0; q0 < nq; q0 += kFilterQ) {
        __syncthreads();  // previous query tile consumed; base tile staged
        for (uint32_t i = tid; i < kFilterQ * chunks4; i += 256) {
            const uint32_t r = i / chunks4, c = i % chunks4;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (q0 + r < nq) v = *reinterpret_cast<const uint4*>(qs + size_t(q0 + r) * D + c * 4);
            *reinterpret_cast<uint4*>(sQ + r * stride + c * 4) = v;
        }
        for (uint32_t i = tid; i < kFilterQ; i += 256) {
            sQP[i] = q0 + i < nq ? qP[q0 + i] : 0.f;
            sQQ[i] = q0 + i < nq ? qQ[q0 + i] : 0.f;
        }
        __syncthreads();
        f32x16 acc[TB / 32];
#pragma unroll
        for (int t = 0; t < TB / 32; ++t) acc[t] = f32x16{};
        const uint32_t* sa = sQ + (32 * w + (lane & 31u)) * stride + 8 * (lane >> 5);
        for (uint32_t k0 = 0; k0 < D; k0 += 16) {
            const bf16x8 ahi = *reinterpret_cast<const bf16x8*>(sa + k0);
            const bf16x8 alo = *reinterpret_cast<const bf16x8*>(sa + k0 + 4);
#pragma unroll
            for (int t = 0; t < TB / 32; ++t) {
                const uint32_t* sb = sB + (32 * t + (lane & 31u)) * stride + 8 * (lane >> 5) + k0;
                const bf16x8 bhi = *reinterpret_cast<const bf16x8*>(sb);
                const bf16x8 blo = *reinterpret_cast<const bf16x8*>(sb + 4);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi, acc[t], 0, 0, 0);
            }
        }
        // Element q of tile t: query 32w + 8(q/4) + 4(lane/32) + q%4 of the
        // tile, base row 32t + lane%32.
        float Pj[TB / 32], nj[TB / 32];
#pragma unroll
        for (int t = 0; t < TB / 32; ++t) {
            Pj[t] = sBP[32 * t + (lane & 31u)];
            nj[t] = sBn[32 * t + (lane & 31u)];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t il = 32 * w + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            const uint32_t qi = q0 + il;
            const float Pi = sQP[il], Qi = sQQ[il];
#pragma unroll
            for (int t = 0; t < TB / 32; ++t) {
                const uint32_t bj = b0 + 32 * t + (lane & 31u);
                const float g = __builtin_fmaf(Qi, nj[t], __builtin_fmaf(c0, acc[t][q], Pj[t]) + Pi);
                if (g <= 0.0f && qi < nq && bj < n) {
                    const uint32_t pos = atomicAdd(cnt + qi, 1u);
                    if (pos < kCandCap) cand[size_t(qi) * kCandCap + pos] = bj;
                }
            }
        }
    }
}

// FP64 distance of one pair by the FP64 kernel's own expression (chains
// over the padded dimension in order; L2 squares differences, dot/cosine
// use the product and norm chains).
__device__ __forceinline__ double fp64_distance(int metric, const float* y, const float* x, uint32_t D, double qn) {
    double s = 0, dt = 0, xn = 0;
    if (metric == kL2Sqr || metric == kL2) {
        for (uint32_t d = 0; d < D; ++d) {
            const double df = double(y[d]) - double(x[d]);
            s = fma(df, df, s);
        }
        return metric == kL2 ? sqrt(s) : s;
    }
    for (uint32_t d = 0; d < D; ++d) {
        dt = fma(double(y[d]), double(x[d]), dt);
        xn = fma(double(x[d]), double(x[d]), xn);
    }
    if (metric == kDot) return -dt;
    const double den = sqrt(qn * xn);
    return den > 0 ? 1.0 - dt / den : 1.0;
}

// Recheck: one workgroup per query.  Candidates' FP64 distances, a bitonic
// sort of (distance, vid) in LDS, the first k written out.  Overflowed
// queries are left to the host.
__global__ __launch_bounds__(256) void recheck_kernel(int metric, const float* __restrict__ base,
                                                      const float* __restrict__ queries, uint32_t D, uint32_t k,
                                                      const uint32_t* __restrict__ cand,
                                                      const uint32_t* __restrict__ cnt, uint32_t* __restrict__ out_v,
                                                      double* __restrict__ out_d) {
    __shared__ double sd[kCandCap];
    __shared__ uint32_t sv[kCandCap];
    __shared__ float sq[1024];
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint32_t m = cnt[q];
    if (m > kCandCap) return;
    for (uint32_t d = tid; d < D; d += 256) sq[d] = queries[size_t(q) * D + d];
    __syncthreads();
    double qn = 0;
    for (uint32_t d = 0; d < D; ++d) qn = fma(double(sq[d]), double(sq[d]), qn);
    uint32_t P = 1;
    while (P < m) P <<= 1;
    for (uint32_t i = tid; i < P; i += 256) {
        if (i < m) {
            const uint32_t v = cand[size_t(q) * kCandCap + i];
            sv[i] = v;
            sd[i] = fp64_distance(metric, sq, base + size_t(v) * D, D, qn);
        } else {
            sv[i] = 0xFFFFFFFFu;
            sd[i] = INFINITY;
        }
    }
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = tid; i < P; i += 256) {
                const uint32_t j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const double di = sd[i], dj = sd[j];
                    const uint32_t vi = sv[i], vj = sv[j];
                    const bool gt = di > dj || (di == dj && vi > vj);
                    if (gt == up) {
                        sd[i] = dj;
                        sd[j] = di;
                        sv[i] = vj;
                        sv[j] = vi;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = tid; i < k; i += 256) {
        out_v[size_t(q) * k + i] = i < m ? sv[i] : kInvalidVid;
        out_d[size_t(q) * k + i] = i < m ? sd[i] : INFINITY;
    }
}

// Device scratch of one MFMA brute-force call, freed on every exit path.
struct BfScratch {
    std::vector<void*> ptrs;
    ~BfScratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int get(T** out, size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
            set_error("artea_bruteforce_knn_mfma_device: device allocation failed");
            return ARTEA_ENOMEM;
        }
        ptrs.push_back(p);
        *out = static_cast<T*>(p);
        return ARTEA_OK;
    }
};

}  // namespace

extern "C" {

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

}  // extern "C"

extern "C" int artea_bruteforce_knn_mfma_device(int metric, const float* d_base, uint32_t n, const float* d_queries,
                                                uint32_t nq, uint32_t D, uint32_t k, uint32_t* d_out_vids,
                                                double* d_out_dists, void* stream, uint64_t* stats) {
    clear_error();
    ARTEA_CHECK_ARG(metric >= 0 && metric <= 3, "artea_bruteforce_knn_mfma_device: metric must be 0..3");
    ARTEA_CHECK_ARG(D > 0 && D % 16 == 0 && D <= 1024,
                    "artea_bruteforce_knn_mfma_device: dim_pad must be a multiple of 16, <= 1024");
    ARTEA_CHECK_ARG(k >= 1 && k <= 256, "artea_bruteforce_knn_mfma_device: k must be in [1, 256]");
    if (stats) stats[0] = stats[1] = stats[2] = 0;
    if (nq == 0) return ARTEA_OK;
    ARTEA_CHECK_ARG(d_base && d_queries && d_out_vids && d_out_dists, "artea_bruteforce_knn_mfma_device: null buffer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // Sample size: the k-th distance of a prefix of ns rows bounds the full
    // k-th distance; ns is chosen so that about k * n / ns <= cap / 2 rows
    // pass per query.  Small bases and wide rows take the FP64 kernel alone.
    const uint64_t ns64 = std::max<uint64_t>(n / 16, (2ull * k * n + kCandCap - 1) / kCandCap);
    if (D > 128 || n < 8192 || ns64 >= n)
        return artea_bruteforce_knn_device(metric, d_base, n, d_queries, nq, D, k, d_out_vids, d_out_dists, stream);
    const uint32_t ns = uint32_t(ns64);
    const bool l2 = metric == kL2Sqr || metric == kL2;
    // Error radius constant: split and product terms 3.2 * 2^-16, FP32
    // accumulation over 3D products and the norm chains (4D + 8) * 2^-24,
    // doubled for the roundings of the affine test.
    const float eps = float(2.0 * (3.2 * std::ldexp(1.0, -16) + (4.0 * D + 8.0) * std::ldexp(1.0, -24)));
    BfScratch sc;
    uint32_t *xs, *qs, *cand, *cnt, *sv;
    float *xn2, *xn, *xP, *qn2, *qn, *qP, *qQ, *mu = nullptr;
    double *sd, *part = nullptr;
    int rc;
    if ((rc = sc.get(&xs, size_t(n) * D)) || (rc = sc.get(&qs, size_t(nq) * D)) ||
        (rc = sc.get(&cand, size_t(nq) * kCandCap)) || (rc = sc.get(&cnt, nq)) || (rc = sc.get(&sv, size_t(nq) * k)) ||
        (rc = sc.get(&xn2, n)) || (rc = sc.get(&xn, n)) || (rc = sc.get(&xP, n)) || (rc = sc.get(&qn2, nq)) ||
        (rc = sc.get(&qn, nq)) || (rc = sc.get(&qP, nq)) || (rc = sc.get(&qQ, nq)) || (rc = sc.get(&sd, size_t(nq) * k)))
        return rc;
    if (l2) {
        const uint32_t per = 4096, blocks = (n + per - 1) / per;
        if ((rc = sc.get(&mu, D)) || (rc = sc.get(&part, size_t(blocks) * D))) return rc;
        hipLaunchKernelGGL(colsum_kernel, dim3(blocks), dim3(256), 0, s, d_base, n, D, per, part);
        hipLaunchKernelGGL(colmean_kernel, dim3(1), dim3(256), 0, s, part, blocks, n, D, mu);
    }
    hipLaunchKernelGGL(split_kernel, dim3(std::min<uint32_t>((n + 3) / 4, 65536)), dim3(256), 0, s, d_base, n, D,
                       mu, xs, xn2, xn);
    hipLaunchKernelGGL(split_kernel, dim3(std::min<uint32_t>((nq + 3) / 4, 65536)), dim3(256), 0, s, d_queries, nq,
                       D, mu, qs, qn2, qn);
    ARTEA_HIP_TRY(hipGetLastError());
    // 1. thresholds from the sample
    rc = artea_bruteforce_knn_device(metric, d_base, ns, d_queries, nq, D, k, sv, sd, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(query_coef_kernel, dim3((nq + 255) / 256), dim3(256), 0, s, metric, sd, nq, k, qn2, qn, eps,
                       qP, qQ);
    hipLaunchKernelGGL(base_coef_kernel, dim3((n + 255) / 256), dim3(256), 0, s, metric, xn2, n, eps, xP);
    ARTEA_HIP_TRY(hipMemsetAsync(cnt, 0, size_t(nq) * 4, s));
    // 2. filter
    constexpr int TB = 128;
    const size_t lds = (size_t(TB) + kFilterQ) * (D + 4) * 4 + (2 * size_t(TB) + 2 * kFilterQ) * 4;
    ARTEA_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(filter_kernel<TB>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    hipLaunchKernelGGL(filter_kernel<TB>, dim3((n + TB - 1) / TB), dim3(256), lds, s, xs, xP, xn, n, qs, qP, qQ, nq,
                       D, l2 ? -2.0f : -1.0f, cand, cnt);
    ARTEA_HIP_TRY(hipGetLastError());
    // 3. recheck
    hipLaunchKernelGGL(recheck_kernel, dim3(nq), dim3(256), 0, s, metric, d_base, d_queries, D, k, cand, cnt,
                       d_out_vids, d_out_dists);
    ARTEA_HIP_TRY(hipGetLastError());
    std::vector<uint32_t> hc(nq);
    ARTEA_HIP_TRY(hipMemcpyAsync(hc.data(), cnt, size_t(nq) * 4, hipMemcpyDeviceToHost, s));
    ARTEA_HIP_TRY(hipStreamSynchronize(s));
    uint64_t total = 0;
    std::vector<uint32_t> over;
    for (uint32_t q = 0; q < nq; ++q) {
        total += std::min(hc[q], kCandCap);
        if (hc[q] > kCandCap) over.push_back(q);
    }
    // Overflowed queries: the FP64 kernel over the whole base, gathered into
    // a contiguous batch and scattered back.
    if (!over.empty()) {
        const uint32_t m = uint32_t(over.size());
        float* oq;
        uint32_t* ov;
        double* od;
        if ((rc = sc.get(&oq, size_t(m) * D)) || (rc = sc.get(&ov, size_t(m) * k)) || (rc = sc.get(&od, size_t(m) * k)))
            return rc;
        for (uint32_t i = 0; i < m; ++i)
            ARTEA_HIP_TRY(hipMemcpyAsync(oq + size_t(i) * D, d_queries + size_t(over[i]) * D, size_t(D) * 4,
                                         hipMemcpyDeviceToDevice, s));
        rc = artea_bruteforce_knn_device(metric, d_base, n, oq, m, D, k, ov, od, stream);
        if (rc) return rc;
        for (uint32_t i = 0; i < m; ++i) {
            ARTEA_HIP_TRY(hipMemcpyAsync(d_out_vids + size_t(over[i]) * k, ov + size_t(i) * k, size_t(k) * 4,
                                         hipMemcpyDeviceToDevice, s));
            ARTEA_HIP_TRY(hipMemcpyAsync(d_out_dists + size_t(over[i]) * k, od + size_t(i) * k, size_t(k) * 8,
                                         hipMemcpyDeviceToDevice, s));
        }
    }
    ARTEA_HIP_TRY(hipStreamSynchronize(s));
    if (stats) {
        stats[0] = total;
        stats[1] = over.size();
        stats[2] = ns;
    }
    return ARTEA_OK;
}

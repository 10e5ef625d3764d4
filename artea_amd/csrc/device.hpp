// Device-side building blocks shared by the router, refine and utility
// kernels: wave-level primitives for 64-wide CDNA4 wavefronts and the FP32
// distance kernel that reproduces dist_func_t bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace artea {
namespace dev {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ __forceinline__ float readlane_f(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ uint32_t readlane_u(uint32_t x, int l) {
    return uint32_t(__builtin_amdgcn_readlane(int(x), l));
}
__device__ __forceinline__ uint32_t readfirst_u(uint32_t x) {
    return uint32_t(__builtin_amdgcn_readfirstlane(int(x)));
}

// Orders LDS traffic between lanes of one wavefront (no s_barrier: the
// waves of a block run independent queries).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP quad permutes: xor-2 = [2,3,0,1], xor-1 = [1,0,3,2].
__device__ __forceinline__ float quad_xor2(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, false));
}
__device__ __forceinline__ float quad_xor1(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}

// One DPP min step; lanes without a source keep their own value.
template <int CTRL>
__device__ __forceinline__ float dpp_min_step(float x) {
    const int y = __builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), CTRL, 0xF, 0xF, false);
    return fminf(x, __int_as_float(y));
}

// Minimum over the 64 lanes (quad permutes, row shifts, row broadcasts; the
// full minimum lands in lane 63), returned to every lane.
__device__ __forceinline__ float wave_min(float x) {
    x = dpp_min_step<0xB1>(x);   // quad_perm [1,0,3,2]
    x = dpp_min_step<0x4E>(x);   // quad_perm [2,3,0,1]
    x = dpp_min_step<0x114>(x);  // row_shr:4
    x = dpp_min_step<0x118>(x);  // row_shr:8
    x = dpp_min_step<0x142>(x);  // row_bcast:15
    x = dpp_min_step<0x143>(x);  // row_bcast:31
    return readlane_f(x, 63);
}

// Sum of the 16 lane accumulators of one AVX-512 register in the order of
// GCC's _mm512_reduce_add_ps: t8[i] = a[i+8] + a[i], t4[i] = t8[i+4] + t8[i],
// (t4[0]+t4[2]) + (t4[1]+t4[3]).  A quad of lanes holds the register:
// quad lane j owns accumulators 4j..4j+3 in a0..a3.  Every lane of the quad
// returns the same value.
__device__ __forceinline__ float reduce16_quad(float a0, float a1, float a2, float a3) {
    // t8: quad lanes 0,1 add partner 2,3 (accumulator i and i+8).
    const float b0 = a0 + quad_xor2(a0), b1 = a1 + quad_xor2(a1);
    const float b2 = a2 + quad_xor2(a2), b3 = a3 + quad_xor2(a3);
    // t4: quad lane 0 holds t8[0..3], lane 1 holds t8[4..7].
    const float c0 = b0 + quad_xor1(b0), c1 = b1 + quad_xor1(b1);
    const float c2 = b2 + quad_xor1(b2), c3 = b3 + quad_xor1(b3);
    const float x0 = c0 + c2;
    const float x1 = c1 + c3;
    return x0 + x1;
}

enum Metric : int { kL2Sqr = 0, kDot = 1, kCosine = 2, kL2 = 3 };

// dist_func_t(q, x) (utils/simd_distance.hpp:64-79, unroll 1) computed by a
// quad of lanes: quad lane `sub` holds AVX lanes 4sub..4sub+3 of each
// 16-float chunk, so each accumulator sees chunks 0..C-1 in order, exactly
// as _mm512_fmadd_ps does.  q may live in LDS or global memory; qn2 is the
// query's own squared norm chain (cosine only).
template <int CT>
__device__ __forceinline__ float quad_distance(const float* __restrict__ q, const float* __restrict__ x, int C,
                                               int metric, uint32_t sub, float qn2) {
    const int Cn = CT > 0 ? CT : C;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    if (metric == kL2Sqr || metric == kL2) {
#pragma unroll(CT > 0 ? CT : 4)
        for (int c = 0; c < Cn; ++c) {
            const float4 xv = *reinterpret_cast<const float4*>(x + c * 16 + sub * 4);
            const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
            const float d0 = qv.x - xv.x, d1 = qv.y - xv.y, d2 = qv.z - xv.z, d3 = qv.w - xv.w;
            a0 = __builtin_fmaf(d0, d0, a0);
            a1 = __builtin_fmaf(d1, d1, a1);
            a2 = __builtin_fmaf(d2, d2, a2);
            a3 = __builtin_fmaf(d3, d3, a3);
        }
        const float s = reduce16_quad(a0, a1, a2, a3);
        return metric == kL2 ? __builtin_sqrtf(s) : s;
    }
    if (metric == kDot) {
#pragma unroll(CT > 0 ? CT : 4)
        for (int c = 0; c < Cn; ++c) {
            const float4 xv = *reinterpret_cast<const float4*>(x + c * 16 + sub * 4);
            const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
            a0 = __builtin_fmaf(qv.x, xv.x, a0);
            a1 = __builtin_fmaf(qv.y, xv.y, a1);
            a2 = __builtin_fmaf(qv.z, xv.z, a2);
            a3 = __builtin_fmaf(qv.w, xv.w, a3);
        }
        return -reduce16_quad(a0, a1, a2, a3);
    }
    // Cosine (_impl_cosine :241-275): cross and |x|^2 chains here, |q|^2 given.
    float n0 = 0.f, n1 = 0.f, n2 = 0.f, n3 = 0.f;
#pragma unroll(CT > 0 ? CT : 4)
    for (int c = 0; c < Cn; ++c) {
        const float4 xv = *reinterpret_cast<const float4*>(x + c * 16 + sub * 4);
        const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
        a0 = __builtin_fmaf(qv.x, xv.x, a0);
        a1 = __builtin_fmaf(qv.y, xv.y, a1);
        a2 = __builtin_fmaf(qv.z, xv.z, a2);
        a3 = __builtin_fmaf(qv.w, xv.w, a3);
        n0 = __builtin_fmaf(xv.x, xv.x, n0);
        n1 = __builtin_fmaf(xv.y, xv.y, n1);
        n2 = __builtin_fmaf(xv.z, xv.z, n2);
        n3 = __builtin_fmaf(xv.w, xv.w, n3);
    }
    const float dot = reduce16_quad(a0, a1, a2, a3);
    const float sq2 = reduce16_quad(n0, n1, n2, n3);
    const float denom = __builtin_sqrtf(qn2 * sq2);
    return denom > 0.0f ? 1.0f - dot / denom : 1.0f;
}

// The metric value of one row whose quad slice (CT float4 per lane, chunk
// order) is already in registers; same chains and rounding as quad_distance.
template <int CT>
__device__ __forceinline__ float quad_finish(const float* __restrict__ q, const float4 (&xv)[CT], int metric,
                                             uint32_t sub, float qn2) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    if (metric == kL2Sqr || metric == kL2) {
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
            const float d0 = qv.x - xv[c].x, d1 = qv.y - xv[c].y, d2 = qv.z - xv[c].z, d3 = qv.w - xv[c].w;
            a0 = __builtin_fmaf(d0, d0, a0);
            a1 = __builtin_fmaf(d1, d1, a1);
            a2 = __builtin_fmaf(d2, d2, a2);
            a3 = __builtin_fmaf(d3, d3, a3);
        }
        const float s = reduce16_quad(a0, a1, a2, a3);
        return metric == kL2 ? __builtin_sqrtf(s) : s;
    }
    if (metric == kDot) {
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
            a0 = __builtin_fmaf(qv.x, xv[c].x, a0);
            a1 = __builtin_fmaf(qv.y, xv[c].y, a1);
            a2 = __builtin_fmaf(qv.z, xv[c].z, a2);
            a3 = __builtin_fmaf(qv.w, xv[c].w, a3);
        }
        return -reduce16_quad(a0, a1, a2, a3);
    }
    float n0 = 0.f, n1 = 0.f, n2 = 0.f, n3 = 0.f;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
        a0 = __builtin_fmaf(qv.x, xv[c].x, a0);
        a1 = __builtin_fmaf(qv.y, xv[c].y, a1);
        a2 = __builtin_fmaf(qv.z, xv[c].z, a2);
        a3 = __builtin_fmaf(qv.w, xv[c].w, a3);
        n0 = __builtin_fmaf(xv[c].x, xv[c].x, n0);
        n1 = __builtin_fmaf(xv[c].y, xv[c].y, n1);
        n2 = __builtin_fmaf(xv[c].z, xv[c].z, n2);
        n3 = __builtin_fmaf(xv[c].w, xv[c].w, n3);
    }
    const float dot = reduce16_quad(a0, a1, a2, a3);
    const float sq2 = reduce16_quad(n0, n1, n2, n3);
    const float denom = __builtin_sqrtf(qn2 * sq2);
    return denom > 0.0f ? 1.0f - dot / denom : 1.0f;
}

// Two rows per quad: both rows' slices are loaded before either is reduced,
// so one memory round trip covers 32 rows per wavefront.  The chunk loop
// feeds both rows from one read of the query chunk; every accumulator sees
// the same operations in the same order as in quad_distance, so the results
// equal two quad_distance calls bit for bit.
template <int CT>
__device__ __forceinline__ void quad_distance2(const float* __restrict__ q, const float* __restrict__ xa,
                                               const float* __restrict__ xb, int C, int metric, uint32_t sub,
                                               float qn2, float* da, float* db) {
    if constexpr (CT == 0) {
        *da = quad_distance<0>(q, xa, C, metric, sub, qn2);
        *db = quad_distance<0>(q, xb, C, metric, sub, qn2);
    } else {
        float4 va[CT], vb[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) va[c] = *reinterpret_cast<const float4*>(xa + c * 16 + sub * 4);
#pragma unroll
        for (int c = 0; c < CT; ++c) vb[c] = *reinterpret_cast<const float4*>(xb + c * 16 + sub * 4);
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, b0 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
        if (metric == kL2Sqr || metric == kL2) {
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
                const float x0 = qv.x - va[c].x, x1 = qv.y - va[c].y, x2 = qv.z - va[c].z, x3 = qv.w - va[c].w;
                a0 = __builtin_fmaf(x0, x0, a0);
                a1 = __builtin_fmaf(x1, x1, a1);
                a2 = __builtin_fmaf(x2, x2, a2);
                a3 = __builtin_fmaf(x3, x3, a3);
                const float y0 = qv.x - vb[c].x, y1 = qv.y - vb[c].y, y2 = qv.z - vb[c].z, y3 = qv.w - vb[c].w;
                b0 = __builtin_fmaf(y0, y0, b0);
                b1 = __builtin_fmaf(y1, y1, b1);
                b2 = __builtin_fmaf(y2, y2, b2);
                b3 = __builtin_fmaf(y3, y3, b3);
            }
            const float sa = reduce16_quad(a0, a1, a2, a3), sb = reduce16_quad(b0, b1, b2, b3);
            *da = metric == kL2 ? __builtin_sqrtf(sa) : sa;
            *db = metric == kL2 ? __builtin_sqrtf(sb) : sb;
            return;
        }
        if (metric == kDot) {
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
                a0 = __builtin_fmaf(qv.x, va[c].x, a0);
                a1 = __builtin_fmaf(qv.y, va[c].y, a1);
                a2 = __builtin_fmaf(qv.z, va[c].z, a2);
                a3 = __builtin_fmaf(qv.w, va[c].w, a3);
                b0 = __builtin_fmaf(qv.x, vb[c].x, b0);
                b1 = __builtin_fmaf(qv.y, vb[c].y, b1);
                b2 = __builtin_fmaf(qv.z, vb[c].z, b2);
                b3 = __builtin_fmaf(qv.w, vb[c].w, b3);
            }
            *da = -reduce16_quad(a0, a1, a2, a3);
            *db = -reduce16_quad(b0, b1, b2, b3);
            return;
        }
        // Cosine keeps three chains per row; two rows' worth would outgrow the
        // register budget of the search kernels, so the rows go one by one.
        *da = quad_finish<CT>(q, va, kCosine, sub, qn2);
        *db = quad_finish<CT>(q, vb, kCosine, sub, qn2);
    }
}

// ---- one-byte rows -------------------------------------------------------
// When every value of a row and of the query is a whole number in [0, 255]
// and a row holds at most 256 values, every operation of the FP32 chains
// above is exact (differences, squares and products are integers below 2^16,
// partial sums stay below 2^24), so the chains' results are the exact
// integer sums whatever their order.  Either zero qualifies: the chains'
// accumulators start at +0, a zero product or square added to +0 gives +0,
// and a difference of zeros is squared, so no -0 reaches a result.  The
// one-byte path computes those sums in integers: a quad of lanes per row,
// lane `sub` holding CT words (4*CT bytes) of it, v_dot4_u32_u8 for q.x and
// x.x, quad sums through DPP.
__device__ __forceinline__ bool byte_valued(float f) {
    return f >= 0.f && f <= 255.f && f == __builtin_truncf(f);
}

__device__ __forceinline__ uint32_t quad_sum_u(uint32_t x) {
    x += uint32_t(__builtin_amdgcn_mov_dpp(int(x), 0x4E, 0xF, 0xF, false));
    x += uint32_t(__builtin_amdgcn_mov_dpp(int(x), 0xB1, 0xF, 0xF, false));
    return x;
}

// CT words of one row slice (16-, 8- or 4-byte loads, as the slice's
// alignment allows: slices start at 4*CT*sub bytes of a 16*CT-byte row).
template <int CT>
__device__ __forceinline__ void load_u8_slice(const uint32_t* __restrict__ p, uint32_t (&x)[CT]) {
    if constexpr (CT % 4 == 0) {
#pragma unroll
        for (int i = 0; i < CT / 4; ++i) {
            const uint4 v = reinterpret_cast<const uint4*>(p)[i];
            x[4 * i] = v.x, x[4 * i + 1] = v.y, x[4 * i + 2] = v.z, x[4 * i + 3] = v.w;
        }
    } else if constexpr (CT % 2 == 0) {
#pragma unroll
        for (int i = 0; i < CT / 2; ++i) {
            const uint2 v = reinterpret_cast<const uint2*>(p)[i];
            x[2 * i] = v.x, x[2 * i + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < CT; ++i) x[i] = p[i];
    }
}

// q.x and x.x of one row slice, summed over the quad.
template <int CT>
__device__ __forceinline__ void u8_dots(const uint32_t (&q)[CT], const uint32_t (&x)[CT], uint32_t* qx,
                                        uint32_t* xx) {
    uint32_t a = 0, b = 0;
#pragma unroll
    for (int k = 0; k < CT; ++k) {
        a = __builtin_amdgcn_udot4(q[k], x[k], a, false);
        b = __builtin_amdgcn_udot4(x[k], x[k], b, false);
    }
    *qx = quad_sum_u(a);
    *xx = quad_sum_u(b);
}

// The metric value from the exact sums qn = q.q, qx = q.x, xx = x.x: the same
// final operations as the FP32 chains apply to the same (exact) values.
__device__ __forceinline__ float u8_finish(int metric, uint32_t qn, uint32_t qx, uint32_t xx) {
    if (metric == kL2Sqr || metric == kL2) {
        const float s = float(qn + xx - 2u * qx);
        return metric == kL2 ? __builtin_sqrtf(s) : s;
    }
    if (metric == kDot) return -float(qx);
    const float dot = float(qx), sq2 = float(xx), qn2 = float(qn);
    const float denom = __builtin_sqrtf(qn2 * sq2);
    return denom > 0.0f ? 1.0f - dot / denom : 1.0f;
}

// |q|^2 chain of the cosine kernel (same lane order), by a quad.
__device__ __forceinline__ float quad_norm2(const float* __restrict__ q, int C, uint32_t sub) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    for (int c = 0; c < C; ++c) {
        const float4 qv = *reinterpret_cast<const float4*>(q + c * 16 + sub * 4);
        a0 = __builtin_fmaf(qv.x, qv.x, a0);
        a1 = __builtin_fmaf(qv.y, qv.y, a1);
        a2 = __builtin_fmaf(qv.z, qv.z, a2);
        a3 = __builtin_fmaf(qv.w, qv.w, a3);
    }
    return reduce16_quad(a0, a1, a2, a3);
}

// The 16 accumulators of one AVX-512 register summed in the order of GCC's
// _mm512_reduce_add_ps (the tree reduce16_quad spreads over a quad), in one
// lane: t8[i] = a[i+8] + a[i], t4[i] = t8[i+4] + t8[i], then
// (t4[0] + t4[2]) + (t4[1] + t4[3]).
__device__ __forceinline__ float reduce16_lane(const float (&a)[16]) {
    float t8[8], t4[4];
#pragma unroll
    for (int i = 0; i < 8; ++i) t8[i] = a[i + 8] + a[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) t4[i] = t8[i + 4] + t8[i];
    return (t4[0] + t4[2]) + (t4[1] + t4[3]);
}

// dist_func_t(a, b) (utils/simd_distance.hpp:64-79, 241-275) by one lane:
// the 16 accumulator chains over chunks 0..C-1 in order, then the reduce
// tree; the same bits as quad_distance.  CT > 0: C == CT, the chunk loop
// unrolled so every load of the row is in flight at once.
template <int CT = 0>
__device__ __forceinline__ float lane_distance(const float* __restrict__ a, const float* __restrict__ b, uint32_t C,
                                               int metric) {
    float acc[16], na[16], nb[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) acc[l] = na[l] = nb[l] = 0.f;
    const uint32_t nc = CT > 0 ? uint32_t(CT) : C;
#pragma unroll(CT > 0 ? CT : 1)
    for (uint32_t c = 0; c < nc; ++c) {
        float x[16], y[16];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float4 s = *reinterpret_cast<const float4*>(a + c * 16 + v * 4);
            const float4 t = *reinterpret_cast<const float4*>(b + c * 16 + v * 4);
            x[4 * v] = s.x, x[4 * v + 1] = s.y, x[4 * v + 2] = s.z, x[4 * v + 3] = s.w;
            y[4 * v] = t.x, y[4 * v + 1] = t.y, y[4 * v + 2] = t.z, y[4 * v + 3] = t.w;
        }
        if (metric == kL2Sqr || metric == kL2) {
#pragma unroll
            for (int l = 0; l < 16; ++l) {
                const float d = x[l] - y[l];
                acc[l] = __builtin_fmaf(d, d, acc[l]);
            }
        } else {
#pragma unroll
            for (int l = 0; l < 16; ++l) acc[l] = __builtin_fmaf(x[l], y[l], acc[l]);
            if (metric == kCosine) {
#pragma unroll
                for (int l = 0; l < 16; ++l) {
                    na[l] = __builtin_fmaf(x[l], x[l], na[l]);
                    nb[l] = __builtin_fmaf(y[l], y[l], nb[l]);
                }
            }
        }
    }
    const float s = reduce16_lane(acc);
    if (metric == kL2Sqr) return s;
    if (metric == kL2) return __builtin_sqrtf(s);
    if (metric == kDot) return -s;
    const float denom = __builtin_sqrtf(reduce16_lane(na) * reduce16_lane(nb));
    return denom > 0.0f ? 1.0f - s / denom : 1.0f;
}

// Total order of CandidateEntry / Neighbor (candidate_entry.hpp:170-180,
// neighbor.hpp:247-258): distance, then vid.
__device__ __forceinline__ bool key_less(float da, uint32_t va, float db, uint32_t vb) {
    return da < db || (da == db && va < vb);
}

}  // namespace dev
}  // namespace artea

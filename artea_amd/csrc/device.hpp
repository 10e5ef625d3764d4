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

// Total order of CandidateEntry / Neighbor (candidate_entry.hpp:170-180,
// neighbor.hpp:247-258): distance, then vid.
__device__ __forceinline__ bool key_less(float da, uint32_t va, float db, uint32_t vb) {
    return da < db || (da == db && va < vb);
}

}  // namespace dev
}  // namespace artea

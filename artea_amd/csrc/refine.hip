// Propagate engine on MI355X (refiner/propagate_engine.hpp:38-325 and its
// updaters), plus the full-hierarchy build that drives it.
//
// State: one row of up to `cap` Neighbor records (vid:31|new:1, FP32
// distance) per local vertex, sorted by (distance, vid).  Every `next` is a
// BSP step exactly like PropagateEngine::next: a producer kernel updates rows
// in place and emits log records (target row, vid, distance); the logs are
// then bucketed per target row (count, exclusive scan, scatter) and merged
// by NbrLogTable::apply_logs's rule (nbr_log_table.hpp:127-182): sort by
// (distance, vid), row entries before equal logs, adjacent-unique by vid.
// The merge keeps the first `cap` entries, which is the reference's merge
// followed by the arc (truncate) updater that refine_layer always runs next
// (index/artea_graph/index_factory.hpp:385-393); rows therefore never exceed
// the layer's max_nbr_size on the device.  Log order inside a bucket does
// not matter (the merge sorts), so the result is deterministic.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "refine_ctx.hpp"
#include "rnet.hpp"

using namespace artea;
using namespace artea::dev;
using namespace artea::rf;

namespace {

constexpr uint32_t kMergeN = 256;      // merge: entries sorted per round by one wavefront

// The same counter-based draw as the oracle: SplitMix64 finaliser chained
// over (seed, stream, row, j), mapped to [lo, hi) by multiply-high.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint32_t rand_draw(uint64_t seed, uint64_t stream, uint32_t row, uint32_t j, uint32_t lo,
                                              uint32_t hi) {
    const uint64_t x = mix64(mix64(mix64(seed ^ mix64(stream)) ^ uint64_t(row)) + uint64_t(j));
    return lo + uint32_t((((x >> 32) * uint64_t(hi - lo)) >> 32));
}

__device__ __forceinline__ void emit(const Ctx& c, uint32_t src, uint32_t slot, uint32_t target, uint32_t vid,
                                     float d) {
    c.e_t[size_t(src) * c.cap + slot] = target;
    c.e_vd[size_t(src) * c.cap + slot] = make_uint2(vid, __float_as_uint(d));
    atomicAdd(c.t_cnt + target, 1u);
}

// ---------------------------------------------------------------- prefill
// RandomUpdater::update_impl (refiner/updaters/random_updater.hpp:97-142),
// one wavefront per row of [row_begin, row_end).
__global__ __launch_bounds__(256) void prefill_kernel(Ctx c, uint32_t gen, uint32_t lo, uint32_t hi, uint32_t rb,
                                                      uint32_t re, uint64_t seed, uint64_t stream) {
    __shared__ uint32_t s_cand[4][64];
    __shared__ float s_d[4][64];
    const uint32_t wid = threadIdx.x / 64, lane = lane_id();
    for (uint32_t row = rb + blockIdx.x * 4 + wid; row < re; row += gridDim.x * 4) {
        const uint32_t n0 = c.cnt[row];
        if (lane == 0) c.e_cnt[row] = 0;
        if (n0 >= gen) continue;
        const uint32_t need = gen - n0;
        const uint32_t sv = storage_of(c, row);
        const float* pv = c.vecs + size_t(sv) * c.D;
        const float qn2 = c.metric == kCosine ? quad_norm2(pv, int(c.C), lane & 3u) : 0.f;
        uint32_t emitted = 0;
        for (uint32_t j0 = 0; j0 < need; j0 += 64) {
            const uint32_t j = j0 + lane;
            uint32_t rg = 0xFFFFFFFFu;
            if (j < need) rg = storage_of(c, rand_draw(seed, stream, row, j, lo, hi));
            const bool ok = j < need && rg != sv;
            const uint64_t m = ballot(ok);
            if (ok) s_cand[wid][lanes_below(m)] = rg;
            wave_sync();
            const uint32_t k = uint32_t(__popcll(m));
            for (uint32_t g = 0; g < k; g += 16) {
                const uint32_t r = g + (lane >> 2);
                float d = 0.f;
                if (r < k) d = quad_distance<0>(pv, c.vecs + size_t(s_cand[wid][r]) * c.D, int(c.C), c.metric,
                                                lane & 3u, qn2);
                if (r < k && (lane & 3u) == 0) s_d[wid][r] = d;
            }
            wave_sync();
            // Row-full skip of the reference (never true below the gate, kept for fidelity).
            bool keep = false;
            float d = 0.f;
            if (lane < k) {
                d = s_d[wid][lane];
                keep = !(n0 >= c.cap && __uint_as_float(c.rows[size_t(row) * c.cap + c.cap - 1].y) <= d);
            }
            const uint64_t km = ballot(keep);
            if (keep) emit(c, row, emitted + lanes_below(km), row, s_cand[wid][lane], d);
            emitted += uint32_t(__popcll(km));
            wave_sync();
        }
        if (lane == 0) c.e_cnt[row] = emitted;
    }
}

// ---------------------------------------------------------------- reverse
// ReverseUpdater::update_impl (refiner/updaters/reverse_updater.hpp:72-100).
__global__ void reverse_kernel(Ctx c) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(c.n) * c.cap) return;
    const uint32_t row = uint32_t(i / c.cap), k = uint32_t(i % c.cap);
    const uint32_t nr = (row >= c.rb && row < c.re) ? c.cnt[row] : 0u;  // outside the window: no logs
    if (k == 0) c.e_cnt[row] = nr;
    if (k >= nr) return;
    const uint2 e = c.rows[i];
    uint32_t t = local_of(c, e.x & kMask);
    if (t == 0xFFFFFFFFu) t = c.n;  // outside the layer: parked in a dead bucket, never merged
    c.e_t[i] = t;
    c.e_vd[i] = make_uint2(storage_of(c, row), e.y);
    atomicAdd(c.t_cnt + t, 1u);
}

// --------------------------------------------------------------- triangle
// TriangleUpdater::update_impl / _internal_check
// (refiner/updaters/triangle_updater.hpp:114-157, 200-237), one 256-thread
// block per row, rows sorted by (distance, vid).  The reference walks the
// entries in order: entry i is compared with the retained entries from the
// farthest down (pairs where both are old are skipped); the first retained r
// with dist(i, r) < d_i / scale - shift_term prunes i, and the closest
// retained r at or beyond that point with dist(i, r) < d_i receives i as a
// recommendation (equal distances: the later position).  Those decisions
// only depend on two per-pair predicates, so the block evaluates them for all
// pairs at once as bitmasks over positions:
//   H_i = { r < i : pair needed, dist(i, r) < prune_i }   (prune hits)
//   C_i = { r < i : pair needed, dist(i, r) < d_i }       (recommendation candidates)
// The retained set R is then the fixed point of "i in R iff H_i and R are
// disjoint" (evaluated in rounds of a wave, each round deciding at least the
// lowest open entry), and entry i recommends to the closest member of
// C_i & R & [jstar_i, i) with jstar_i = the highest member of H_i & R.
//
// Pair distances come from one of two paths:
//  * FP32: every needed pair by the FP32 chain of dist_func_t;
//  * BF16: the row's Gram matrix in 32x32 tiles on MFMA (two-term BF16 split
//    of each coordinate, FP32 accumulation; L2 vectors are first centred on
//    the row's own vertex) gives every pair a "key" (the squared distance for
//    L2, the metric value otherwise) inside a proven interval.  A pair whose
//    interval straddles prune_i or d_i, and every candidate that may be the
//    recommendation target, is recomputed by the FP32 chain, so every
//    decision and every emitted distance equals the FP32 path's.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kStageBatch = 8;      // row-vector gathers in flight per thread
constexpr uint32_t kListCap = 768;  // FP32 recheck queue per row; overflow goes to a bitmap
constexpr uint32_t kMfmaMinCap = 64; // BF16 tiles only for rows of more entries

// Half-widths of the BF16 key intervals (derived on the host from D).
struct JoinBounds {
    float l2;   // x (N_i + N_j) on the squared distance, N = centred squared norms
    float dot;  // x sqrt(N_i) sqrt(N_j)
    float cos;  // absolute
};

struct TriLayout {
    uint32_t vid, dist, norm, rn, pm, thr, best, H, Cw, X, pend, list, P, vec, vstride, total;
};

__host__ __device__ inline uint32_t align16u(uint32_t x) { return (x + 15u) & ~15u; }

// LDS of one row: entries up to cap (rounded to the 32-row tile), the pair
// triangle, the bitmasks, the BF16 path's per-entry data and recheck queue,
// and the row's vectors (FP32, or BF16 hi/lo at one dword per coordinate).
// Row stride: the MFMA operand reads (32 rows, 16 B each) want rows 4 banks
// apart, the FP32 chain's quad reads (16 consecutive dwords per vector) want
// rows 16 banks apart.
__host__ __device__ inline TriLayout tri_layout(uint32_t cap, uint32_t D, bool staged, bool mfma) {
    TriLayout L;
    const uint32_t capr = (cap + 31) / 32 * 32;
    const uint32_t np = cap * (cap - 1) / 2;
    const uint32_t xw = align16u((np + 31) / 32 * 4);
    L.vstride = mfma ? D + 4 : ((D % 64 == 16 || D % 64 == 48) ? D : D + 16);
    uint32_t o = 0;
    auto take = [&](uint32_t& field, uint32_t bytes) {
        field = o;
        o += align16u(bytes);
    };
    take(L.vid, capr * 4);
    take(L.dist, capr * 4);
    take(L.H, capr * 16);
    take(L.Cw, capr * 16);
    take(L.P, np * 4);
    const uint32_t m = mfma ? 1u : 0u;
    take(L.norm, m * capr * 4);
    take(L.rn, m * capr * 4);
    take(L.pm, m * capr * 4);
    take(L.thr, m * capr * 16);
    take(L.best, m * capr * 8);
    take(L.X, m * xw);
    take(L.pend, m * xw);
    take(L.list, m * kListCap * 2);
    take(L.vec, staged ? capr * L.vstride * 4 : 0u);
    L.total = o;
    return L;
}

__device__ __forceinline__ uint32_t tri_index(uint32_t i, uint32_t j) { return i * (i - 1) / 2 + j; }  // i > j

// Work item t of the pair triangle's words: entry i (1 <= i < n) owns
// ceil(i / 32) words of 32 columns j < i; entries 1..32 have one word, 33..64
// two, 65..96 three, 97..127 four, listed entry by entry.  False once i >= n.
__device__ __forceinline__ bool pair_word(uint32_t t, uint32_t n, uint32_t& i, uint32_t& w) {
    uint32_t first = 0;
#pragma unroll
    for (uint32_t words = 1; words <= 4; ++words) {
        if (t < first + 32u * words) {
            i = 32u * (words - 1) + 1 + (t - first) / words;
            w = (t - first) % words;
            return i < n;
        }
        first += 32u * words;
    }
    return false;
}

__device__ __forceinline__ uint32_t word_below(uint32_t i, uint32_t w) {  // bits [0, i) of word w
    const uint32_t b = 32u * w;
    return i >= b + 32u ? 0xFFFFFFFFu : (i > b ? (1u << (i - b)) - 1u : 0u);
}

// Two-term BF16 split of 8 coordinates: x = hi + lo + O(2^-16 |x|).
__device__ __forceinline__ void split8(float4 x0, float4 x1, bf16x8& hi, bf16x8& lo) {
    const float v[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const f32x2 a = {v[2 * p], v[2 * p + 1]};
        const bf16x2 h = __builtin_convertvector(a, bf16x2);
        const f32x2 hf = __builtin_convertvector(h, f32x2);
        const f32x2 r = {a[0] - hf[0], a[1] - hf[1]};
        const bf16x2 l = __builtin_convertvector(r, bf16x2);
        hi[2 * p] = h[0];
        hi[2 * p + 1] = h[1];
        lo[2 * p] = l[0];
        lo[2 * p + 1] = l[1];
    }
}

// Threshold x in the key domain as (certainly below, certainly at or above)
// bounds: key_hi < first implies dist < x, key_lo >= second implies
// dist >= x.  For L2 the FP32 chain's sqrt is correctly rounded, so squared
// keys below x^2 (1 - 2^-20) give distances below x and keys at or above
// x^2 (1 + 2^-20) give distances at or above x; distances are never negative.
__device__ __forceinline__ float2 key_threshold(int metric, float x) {
    if (metric != kL2) return make_float2(x, x);
    if (!(x > 0.0f)) return make_float2(-INFINITY, -INFINITY);
    const float x2 = x * x;
    return make_float2(x2 * (1.0f - 0x1p-20f), x2 * (1.0f + 0x1p-20f));
}

// Interval of the key of an FP32-chain distance v.
__device__ __forceinline__ void exact_key(int metric, float v, float& lo, float& hi) {
    if (metric == kL2) {
        const float v2 = v * v;
        lo = v2 * (1.0f - 0x1p-20f);
        hi = v2 * (1.0f + 0x1p-20f);
    } else {
        lo = hi = v;
    }
}

// Half-width of an approximate key; rn = sqrt(N) (dot) or 1/sqrt(N) (cosine).
__device__ __forceinline__ float key_radius(int metric, const JoinBounds& kb, float ni, float nj, float rni,
                                            float rnj) {
    if (metric == kL2Sqr || metric == kL2) return kb.l2 * (ni + nj);
    if (metric == kDot) return kb.dot * (rni * rnj);
    return kb.cos;
}

// Position of the k-th (0-based) set bit of a 128-bit mask held in 4 words.
__device__ __forceinline__ uint32_t select_in_words(const uint32_t w[4], uint32_t k) {
    uint32_t base = 0, x = w[3];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint32_t pc = __popc(w[t]);
        if (k < pc) {
            x = w[t];
            base = 32u * t;
            break;
        }
        k -= pc;
    }
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t sh = 16; sh >= 1; sh >>= 1) {
        const uint32_t c = __popc(x & ((1u << sh) - 1u));
        if (k >= c) {
            k -= c;
            x >>= sh;
            pos += sh;
        }
    }
    return base + pos;
}

// Recommendation target order: smaller distance first, then the later
// position (-0 and +0 are equal, as the FP32 comparisons see them).
__device__ __forceinline__ unsigned long long target_key(float d, uint32_t r) {
    uint32_t u = __float_as_uint(d == 0.0f ? 0.0f : d);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return (static_cast<unsigned long long>(u) << 32) | (0xFFu - r);
}

template <bool MFMA, int M>
__global__ __launch_bounds__(256, 2) void triangle_kernel(Ctx c, float inv_scale, float shift_term, int staged,
                                                          JoinBounds kb) {
    extern __shared__ __align__(16) unsigned char smem[];
    const TriLayout L = tri_layout(c.cap, c.D, staged != 0, MFMA);
    const uint32_t vstride = L.vstride;
    uint32_t* s_vid = reinterpret_cast<uint32_t*>(smem + L.vid);
    float* s_dist = reinterpret_cast<float*>(smem + L.dist);
    float* s_norm = reinterpret_cast<float*>(smem + L.norm);
    float* s_rn = reinterpret_cast<float*>(smem + L.rn);
    float* s_pm = reinterpret_cast<float*>(smem + L.pm);  // prefix bound of the radius term over r < i
    float4* s_thr = reinterpret_cast<float4*>(smem + L.thr);
    unsigned long long* s_best = reinterpret_cast<unsigned long long*>(smem + L.best);
    uint32_t* s_H = reinterpret_cast<uint32_t*>(smem + L.H);
    uint32_t* s_C = reinterpret_cast<uint32_t*>(smem + L.Cw);
    uint32_t* s_X = reinterpret_cast<uint32_t*>(smem + L.X);
    uint32_t* s_pend = reinterpret_cast<uint32_t*>(smem + L.pend);
    uint16_t* s_list = reinterpret_cast<uint16_t*>(smem + L.list);
    float* s_P = reinterpret_cast<float*>(smem + L.P);
    float* s_vec = reinterpret_cast<float*>(smem + L.vec);       // FP32 path: staged vectors
    uint32_t* s_split = reinterpret_cast<uint32_t*>(smem + L.vec);  // BF16 path: split coordinates
    __shared__ uint32_t s_newm[4], s_R[4], s_nl, s_ne;

    const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / 64;
    constexpr int metric = M;  // the build metric, fixed per instantiation
    const bool centred = MFMA && (metric == kL2Sqr || metric == kL2);
    const uint32_t capr = (c.cap + 31) / 32 * 32;
    const uint32_t xwords = (c.cap * (c.cap - 1) / 2 + 31) / 32;
    unsigned long long exact_pairs = 0, approx_pairs = 0;

    // FP32 coordinates of entry k: LDS when the FP32 path staged them, else
    // global memory (the BF16 path's rechecks: a small share of the pairs).
    auto vec_of = [&](uint32_t k) -> const float* {
        return (staged && !MFMA) ? s_vec + k * vstride : c.vecs + size_t(s_vid[k]) * c.D;
    };
    uint32_t newm[4] = {0, 0, 0, 0};
    auto is_new = [&](uint32_t k) -> bool { return (newm[k >> 5] >> (k & 31)) & 1u; };
    auto prune_of = [&](float d) -> float { return d * inv_scale - shift_term; };
    // dist_func_t's FP32 chain for one pair, by a quad of lanes (all four
    // lanes active, same pair).
    auto exact_pair_quad = [&](uint32_t i, uint32_t j) -> float {
        const uint32_t sub = lane & 3u;
        const float* a = vec_of(i);
        const float* b = vec_of(j);
        const float qn2 = metric == kCosine ? quad_norm2(a, int(c.C), sub) : 0.f;
        // Fully unrolled chunk loops for the common widths (all loads in flight).
        if (c.C == 8) return quad_distance<8>(a, b, 8, metric, sub, qn2);
        if (c.C == 6) return quad_distance<6>(a, b, 6, metric, sub, qn2);
        return quad_distance<0>(a, b, int(c.C), metric, sub, qn2);
    };
    auto mark_exact = [&](uint32_t p) { atomicOr(s_X + (p >> 5), 1u << (p & 31)); };
    auto is_exact = [&](uint32_t p) -> bool { return (s_X[p >> 5] >> (p & 31)) & 1u; };
    // Queue of pairs for the FP32 chain (slot reserved from s_nl); beyond
    // list_cap entries, a bitmap.
    auto put_recheck = [&](uint32_t slot, uint32_t i, uint32_t j) {
        if (slot < c.list_cap) {
            s_list[slot] = uint16_t((i << 7) | j);
        } else {
            const uint32_t p = tri_index(i, j);
            atomicOr(s_pend + (p >> 5), 1u << (p & 31));
        }
    };
    // Runs the queued pairs through the FP32 chain, a quad of lanes per pair,
    // and hands each result to settle(i, j, value) on the quad's first lane.
    // The overflow bitmap is cleared as it is drained.
    auto drain = [&](auto&& settle) {
        const uint32_t total = s_nl, nl = min(total, c.list_cap);
        for (uint32_t base = 0; base < nl; base += 64) {
            const uint32_t k = base + (tid >> 2);
            if (k < nl) {
                const uint32_t i = s_list[k] >> 7, j = s_list[k] & 127u;
                const float te = exact_pair_quad(i, j);
                if ((tid & 3u) == 0) settle(i, j, te);
            }
        }
        if (total > c.list_cap)
            for (uint32_t w = tid >> 2; w < xwords; w += 64) {
                const uint32_t bits = s_pend[w];
                for (uint32_t m = bits; m; m &= m - 1) {
                    const uint32_t p = 32 * w + __ffs(m) - 1;
                    uint32_t i = uint32_t((1.0f + sqrtf(1.0f + 8.0f * float(p))) * 0.5f);
                    while (i * (i - 1) / 2 > p) --i;
                    while ((i + 1) * i / 2 <= p) ++i;
                    const uint32_t j = p - i * (i - 1) / 2;
                    const float te = exact_pair_quad(i, j);
                    if ((tid & 3u) == 0) settle(i, j, te);
                }
                if (bits && (tid & 3u) == 0) s_pend[w] = 0;
            }
    };

    // The next row's count and entries are loaded while this row is processed.
    uint32_t n_pf = 0;
    uint2 e_pf = make_uint2(0u, 0u);
    auto prefetch = [&](uint32_t r) {
        if (r < c.re) {
            n_pf = c.cnt[r];
            if (tid < c.cap) e_pf = c.rows[size_t(r) * c.cap + tid];
        }
    };
    prefetch(c.rb + blockIdx.x);
    for (uint32_t row = c.rb + blockIdx.x; row < c.re; row += gridDim.x) {
        const uint32_t n = n_pf;
        const uint2 e_row = e_pf;
        prefetch(row + gridDim.x);
        if (n == 0) {
            if (tid == 0) c.e_cnt[row] = 0;
            continue;  // uniform; LDS untouched
        }
        // ---- row metadata, masks, staging
        if (tid < 128) {
            bool nw = false;
            if (tid < n) {
                s_vid[tid] = e_row.x & kMask;
                s_dist[tid] = __uint_as_float(e_row.y);
                nw = (e_row.x & kNew) != 0;
            }
            const uint64_t b = ballot(nw);
            if (lane == 0) {
                s_newm[2 * wid] = uint32_t(b);
                s_newm[2 * wid + 1] = uint32_t(b >> 32);
            }
            if (tid < capr)
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    s_H[tid * 4 + w] = 0;
                    s_C[tid * 4 + w] = 0;
                }
        }
        if (MFMA)
            for (uint32_t k = tid; k < xwords; k += 256) {
                s_X[k] = 0;
                s_pend[k] = 0;
            }
        if (tid == 0) {
            s_nl = 0;
            s_ne = 0;
        }
        __syncthreads();
        const uint32_t T = (n + 31) / 32;
        const uint32_t npad = MFMA ? T * 32 : n;
#pragma unroll
        for (int w = 0; w < 4; ++w) newm[w] = s_newm[w];
        if (MFMA) {
            // Split staging: a group of 8 lanes per vector, one 16-float chunk
            // per lane and pass, the row's four 32-vector passes in flight
            // together.  Coordinates are centred (L2), summed into squared
            // norm partials, and stored as BF16 [hi x8 | lo x8] per 8
            // coordinates (one dword per coordinate).  Padding rows are zero.
            // Per entry then: its norm, root or inverse root, and the two
            // thresholds in the key domain.
            const uint32_t C16 = c.D / 16, g = tid >> 3, sub = tid & 7u;
            const float* cen = c.vecs + size_t(storage_of(c, row)) * c.D;
            float nrm[4] = {0.f, 0.f, 0.f, 0.f};
            for (uint32_t ch = sub; ch < C16; ch += 8) {
                float4 x[4][4];
                float4 m[4] = {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f),
                               make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
                if (centred)
#pragma unroll
                    for (int u = 0; u < 4; ++u) m[u] = *reinterpret_cast<const float4*>(cen + ch * 16 + 4 * u);
#pragma unroll
                for (int ps = 0; ps < 4; ++ps) {
                    const uint32_t k = 32u * ps + g;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        x[ps][u] = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (k < n)
                            x[ps][u] = *reinterpret_cast<const float4*>(c.vecs + size_t(s_vid[k]) * c.D + ch * 16 + 4 * u);
                    }
                }
#pragma unroll
                for (int ps = 0; ps < 4; ++ps) {
                    const uint32_t k = 32u * ps + g;
                    if (k >= npad) continue;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        float4& v = x[ps][u];
                        if (centred && k < n) v = make_float4(v.x - m[u].x, v.y - m[u].y, v.z - m[u].z, v.w - m[u].w);
                        nrm[ps] = __builtin_fmaf(v.x, v.x, nrm[ps]);
                        nrm[ps] = __builtin_fmaf(v.y, v.y, nrm[ps]);
                        nrm[ps] = __builtin_fmaf(v.z, v.z, nrm[ps]);
                        nrm[ps] = __builtin_fmaf(v.w, v.w, nrm[ps]);
                    }
                    bf16x8 h0, l0, h1, l1;
                    split8(x[ps][0], x[ps][1], h0, l0);
                    split8(x[ps][2], x[ps][3], h1, l1);
                    uint32_t* dst = s_split + k * vstride + ch * 16;
                    *reinterpret_cast<bf16x8*>(dst) = h0;
                    *reinterpret_cast<bf16x8*>(dst + 4) = l0;
                    *reinterpret_cast<bf16x8*>(dst + 8) = h1;
                    *reinterpret_cast<bf16x8*>(dst + 12) = l1;
                }
            }
#pragma unroll
            for (int ps = 0; ps < 4; ++ps) {
                float v = nrm[ps];
                v += __shfl_xor(v, 4);
                v += __shfl_xor(v, 2);
                v += __shfl_xor(v, 1);
                const uint32_t k = 32u * ps + g;
                if (k < n && sub == 0) {
                    s_norm[k] = v;
                    s_rn[k] = metric == kDot ? __builtin_sqrtf(v) : (metric == kCosine ? 1.0f / __builtin_sqrtf(v) : 0.f);
                    const float dk = s_dist[k];
                    const float2 tp = key_threshold(metric, prune_of(dk)), td = key_threshold(metric, dk);
                    s_thr[k] = make_float4(tp.x, tp.y, td.x, td.y);
                }
            }
            __syncthreads();
            // Prefix bound of the radius term (norm, or root norm for dot):
            // s_pm[i] = max over r < i.  Waves 0 and 1, read only after later
            // barriers.
            if (tid < 128) {
                auto term = [&](uint32_t k) -> float {
                    return k < n ? (metric == kDot ? s_rn[k] : s_norm[k]) : 0.f;
                };
                float t = term(tid);
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const float o = __shfl_up(t, d);
                    if (int(lane) >= d) t = fmaxf(t, o);
                }
                float ex = __shfl_up(t, 1);
                if (lane == 0) ex = 0.f;
                if (wid == 1) ex = fmaxf(ex, -wave_min(-term(lane)));
                if (tid < n) s_pm[tid] = ex;
            }
        } else {
            if (staged) {
                // Gathers issued kStageBatch at a time per thread before any LDS store.
                const uint32_t per = c.D / 4;  // float4 per vector
                const uint32_t total = npad * per;
                for (uint32_t base = 0; base < total; base += 256 * kStageBatch) {
                    float4 v[kStageBatch];
#pragma unroll
                    for (int u = 0; u < kStageBatch; ++u) {
                        const uint32_t t = base + uint32_t(u) * 256 + tid;
                        const uint32_t k = t / per;
                        v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (t < total && k < n)
                            v[u] = *reinterpret_cast<const float4*>(c.vecs + size_t(s_vid[k]) * c.D + (t - k * per) * 4);
                    }
#pragma unroll
                    for (int u = 0; u < kStageBatch; ++u) {
                        const uint32_t t = base + uint32_t(u) * 256 + tid;
                        const uint32_t k = t / per;
                        if (t < total) *reinterpret_cast<float4*>(s_vec + k * vstride + (t - k * per) * 4) = v[u];
                    }
                }
                __syncthreads();
            }
            // Every needed pair by the FP32 chain, a quad of lanes per pair.
            const uint32_t np = n * (n - 1) / 2;
            for (uint32_t p0 = 0; p0 < np; p0 += 64) {
                const uint32_t p = p0 + (tid >> 2);
                if (p >= np) break;
                uint32_t i = uint32_t((1.0f + sqrtf(1.0f + 8.0f * float(p))) * 0.5f);
                while (i * (i - 1) / 2 > p) --i;
                while ((i + 1) * i / 2 <= p) ++i;
                const uint32_t j = p - i * (i - 1) / 2;
                if (!is_new(i) && !is_new(j)) continue;
                const float d = exact_pair_quad(i, j);
                if ((tid & 3u) == 0) {
                    s_P[p] = d;
                    exact_pairs++;
                }
            }
            __syncthreads();
        }
        const uint32_t ntiles = T * (T + 1) / 2;
        for (uint32_t t = wid; t < ntiles; t += 4) {
            uint32_t I = 0;
            while ((I + 1) * (I + 2) / 2 <= t) ++I;
            const uint32_t J = t - I * (I + 1) / 2;
            const uint32_t j = 32 * J + (lane & 31);
            const bool jnew = j < n && is_new(j);
            const uint32_t newI = I == 0 ? newm[0] : (I == 1 ? newm[1] : (I == 2 ? newm[2] : newm[3]));
            // Lanes 0-31 hold entry i0 and lanes 32-63 entry i0 + 4 of tile column J.
            auto write_masks = [&](int q, bool hit, bool rec) {
                const uint64_t hb = ballot(hit), cb = ballot(rec);
                if (lane == 0) {
                    const uint32_t i0 = 32 * I + 8 * (q >> 2) + (q & 3);
                    s_H[i0 * 4 + J] = uint32_t(hb);
                    s_H[(i0 + 4) * 4 + J] = uint32_t(hb >> 32);
                    s_C[i0 * 4 + J] = uint32_t(cb);
                    s_C[(i0 + 4) * 4 + J] = uint32_t(cb >> 32);
                }
            };
            if (MFMA) {
                f32x16 acc = {}, acc_hl = {}, acc_lh = {};
                const uint32_t* sa = s_split + (32 * I + (lane & 31)) * vstride + 8 * (lane >> 5);
                const uint32_t* sb = s_split + (32 * J + (lane & 31)) * vstride + 8 * (lane >> 5);
                for (uint32_t k0 = 0; k0 < c.D; k0 += 16) {
                    const bf16x8 ahi = *reinterpret_cast<const bf16x8*>(sa + k0);
                    const bf16x8 alo = *reinterpret_cast<const bf16x8*>(sa + k0 + 4);
                    const bf16x8 bhi = *reinterpret_cast<const bf16x8*>(sb + k0);
                    const bf16x8 blo = *reinterpret_cast<const bf16x8*>(sb + k0 + 4);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi, acc, 0, 0, 0);
                    acc_hl = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo, acc_hl, 0, 0, 0);
                    acc_lh = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi, acc_lh, 0, 0, 0);
                }
                // The tile only stores its keys; the decision pass after the
                // last tile takes the decisions.
                const float nj = j < n ? s_norm[j] : 0.f, rnj = j < n ? s_rn[j] : 0.f;
                uint32_t nvalid = 0;
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const uint32_t i = 32 * I + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    const bool valid = (i < n) & (j < i) & (jnew | ((newI >> (i & 31)) & 1u));
                    const float g = acc[q] + (acc_hl[q] + acc_lh[q]);
                    float key;
                    if (metric == kL2Sqr || metric == kL2)
                        key = (s_norm[i] + nj) - 2.0f * g;
                    else if (metric == kDot)
                        key = -g;
                    else
                        key = 1.0f - g * (s_rn[i] * rnj);
                    if (valid) s_P[tri_index(i, j)] = key;
                    nvalid += valid ? 1u : 0u;
                }
                approx_pairs += nvalid;
            } else {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const uint32_t i = 32 * I + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    const bool valid = i < n && j < i && (jnew || ((newI >> (i & 31)) & 1u));
                    const float tv = valid ? s_P[tri_index(i, j)] : 0.f;
                    const float di = i < n ? s_dist[i] : 0.f;
                    write_masks(q, valid && tv < prune_of(di), valid && tv < di);
                }
            }
        }
        __syncthreads();
        if (MFMA) {
            // Decisions from the tiles' keys, one thread per (entry i, word w)
            // of the pair triangle, so every thread of the block takes part
            // whatever the tile split, each word of H and C has one writer,
            // and only the pairs that need a decision are visited.  A pair
            // whose key interval straddles a threshold goes to the FP32
            // recheck queue; the drain below sets its bits.
            for (uint32_t t = tid;; t += 256) {
                uint32_t i, w;
                if (!pair_word(t, n, i, w)) break;
                const float ni = s_norm[i], rni = s_rn[i];
                const float4 th = s_thr[i];
                const uint32_t j0 = 32 * w, span = min(i - j0, 32u), b0 = tri_index(i, 0);
                uint32_t todo = span == 32 ? 0xFFFFFFFFu : (1u << span) - 1u;
                if (!is_new(i)) todo &= w == 0 ? newm[0] : (w == 1 ? newm[1] : (w == 2 ? newm[2] : newm[3]));
                uint32_t hw = 0, cw = 0, aw = 0;
                for (uint32_t m = todo; m; m &= m - 1) {
                    const uint32_t b = uint32_t(__ffs(m) - 1), j = j0 + b;
                    const float key = s_P[b0 + j];
                    const float e = key_radius(metric, kb, ni, s_norm[j], rni, s_rn[j]);
                    const float lo = key - e, hi = key + e;
                    const bool dp = hi < th.x, np_ = lo >= th.y, dd = hi < th.z, nd = lo >= th.w;
                    const bool sure = (dp | np_) & (dd | nd);
                    hw |= uint32_t(sure & dp) << b;
                    cw |= uint32_t(sure & dd) << b;
                    aw |= uint32_t(!sure) << b;
                }
                s_H[i * 4 + w] = hw;
                s_C[i * 4 + w] = cw;
                if (aw) {
                    uint32_t off = atomicAdd(&s_nl, uint32_t(__popc(aw)));
                    for (uint32_t m = aw; m; m &= m - 1) put_recheck(off++, i, j0 + uint32_t(__ffs(m) - 1));
                }
            }
            __syncthreads();
            // Ambiguous predicates: FP32 chain (a quad of lanes per pair), bits set here.
            auto settle = [&](uint32_t i, uint32_t j, float te) {
                const uint32_t p = tri_index(i, j);
                s_P[p] = te;
                mark_exact(p);
                exact_pairs++;
                const float di = s_dist[i];
                if (te < prune_of(di)) atomicOr(s_H + i * 4 + (j >> 5), 1u << (j & 31));
                if (te < di) atomicOr(s_C + i * 4 + (j >> 5), 1u << (j & 31));
            };
            drain(settle);
            __syncthreads();
        }

        // ---- retained set: rounds of "i in R iff H_i and R are disjoint"
        if (wid == 0) {
            uint32_t ha[4], hb[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                ha[w] = lane < n ? s_H[lane * 4 + w] : 0u;
                hb[w] = lane + 64 < n ? s_H[(lane + 64) * 4 + w] : 0u;
            }
            uint32_t kr[4] = {0, 0, 0, 0}, kn[4] = {0, 0, 0, 0};
            bool da = lane >= n, db = lane + 64 >= n;
            for (uint32_t round = 0; round <= kMaxCap; ++round) {
                bool ra = false, na = false, rb = false, nb = false;
                if (!da) {
                    const uint32_t hit = (ha[0] & kr[0]) | (ha[1] & kr[1]) | (ha[2] & kr[2]) | (ha[3] & kr[3]);
                    const uint32_t open = (ha[0] & ~kn[0]) | (ha[1] & ~kn[1]) | (ha[2] & ~kn[2]) | (ha[3] & ~kn[3]);
                    na = hit != 0;
                    ra = !na && open == 0;
                }
                if (!db) {
                    const uint32_t hit = (hb[0] & kr[0]) | (hb[1] & kr[1]) | (hb[2] & kr[2]) | (hb[3] & kr[3]);
                    const uint32_t open = (hb[0] & ~kn[0]) | (hb[1] & ~kn[1]) | (hb[2] & ~kn[2]) | (hb[3] & ~kn[3]);
                    nb = hit != 0;
                    rb = !nb && open == 0;
                }
                const uint64_t bra = ballot(ra), bna = ballot(na), brb = ballot(rb), bnb = ballot(nb);
                kr[0] |= uint32_t(bra);
                kr[1] |= uint32_t(bra >> 32);
                kr[2] |= uint32_t(brb);
                kr[3] |= uint32_t(brb >> 32);
                kn[0] |= uint32_t(bna);
                kn[1] |= uint32_t(bna >> 32);
                kn[2] |= uint32_t(bnb);
                kn[3] |= uint32_t(bnb >> 32);
                da = da || ra || na;
                db = db || rb || nb;
                if (ballot(!da || !db) == 0) break;
            }
            if (lane == 0) {
#pragma unroll
                for (int w = 0; w < 4; ++w) s_R[w] = kr[w];
                s_nl = 0;
            }
        }
        __syncthreads();

        // ---- recommendations: lane per entry
        const uint32_t i = tid;
        uint32_t rec[4] = {0, 0, 0, 0};
        if (i < n && i > 0) {
            uint32_t hr[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const uint32_t rb = s_R[w] & word_below(i, w);
                hr[w] = s_H[i * 4 + w] & rb;
                rec[w] = s_C[i * 4 + w] & rb;
            }
            int jstar = -1;
#pragma unroll
            for (int w = 0; w < 4; ++w)
                if (hr[w]) jstar = 32 * w + 31 - __clz(hr[w]);
            if (jstar >= 0)
#pragma unroll
                for (int w = 0; w < 4; ++w) rec[w] &= ~word_below(uint32_t(jstar), w);
        }
        const bool has_rec = (rec[0] | rec[1] | rec[2] | rec[3]) != 0;
        if (MFMA) {
            // The target must be exact.  Candidates already settled by the
            // FP32 chain give the best so far; of the others, those whose key
            // interval starts at or below the smallest high end can be the
            // closest and are queued.  Radii use the row's prefix bound
            // (max over r < i of r's norm term), so each candidate costs one
            // key load per pass.  The exact flags of row i's pairs are bits
            // [i(i-1)/2, i(i-1)/2 + i) of the pair bitmap.
            if (has_rec) {
                const uint32_t b0 = tri_index(i, 0);
                uint32_t xm[4];
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const uint32_t bit = b0 + 32u * w, wi = bit >> 5, sh = bit & 31u;
                    const uint32_t lo = wi < xwords ? s_X[wi] : 0u;
                    const uint32_t hi = (sh != 0 && wi + 1 < xwords) ? s_X[wi + 1] : 0u;
                    xm[w] = sh ? (lo >> sh) | (hi << (32u - sh)) : lo;
                }
                const float e = key_radius(metric, kb, s_norm[i], s_pm[i], s_rn[i], s_pm[i]);
                float m_hi = INFINITY;
                unsigned long long best = ~0ull;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    uint32_t m = rec[w];
                    while (m) {
                        uint32_t rr[4];
                        float v[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            rr[u] = m ? 32u * w + __ffs(m) - 1 : 0xFFFFFFFFu;
                            m &= m - 1;
                            v[u] = rr[u] != 0xFFFFFFFFu ? s_P[tri_index(i, rr[u])] : 0.f;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            if (rr[u] == 0xFFFFFFFFu) continue;
                            float lo, hi;
                            if ((xm[w] >> (rr[u] & 31u)) & 1u) {
                                exact_key(metric, v[u], lo, hi);
                                const unsigned long long k = target_key(v[u], rr[u]);
                                best = k < best ? k : best;
                            } else {
                                hi = v[u] + e;
                            }
                            m_hi = fminf(m_hi, hi);
                        }
                    }
                }
                s_best[i] = best;
                uint32_t cm[4], cnt = 0;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    cm[w] = 0;
                    uint32_t m = rec[w] & ~xm[w];
                    while (m) {
                        uint32_t rr[4];
                        float v[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            rr[u] = m ? 32u * w + __ffs(m) - 1 : 0xFFFFFFFFu;
                            m &= m - 1;
                            v[u] = rr[u] != 0xFFFFFFFFu ? s_P[tri_index(i, rr[u])] : 0.f;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (rr[u] != 0xFFFFFFFFu && !(v[u] - e > m_hi)) cm[w] |= 1u << (rr[u] & 31u);
                    }
                    cnt += __popc(cm[w]);
                }
                if (cnt) {
                    uint32_t off = atomicAdd(&s_nl, cnt);
#pragma unroll
                    for (int w = 0; w < 4; ++w)
                        for (uint32_t m = cm[w]; m; m &= m - 1) put_recheck(off++, i, 32u * w + __ffs(m) - 1);
                }
            }
            __syncthreads();
            drain([&](uint32_t a, uint32_t b, float te) {
                const uint32_t p = tri_index(a, b);
                s_P[p] = te;
                mark_exact(p);
                exact_pairs++;
                atomicMin(s_best + a, target_key(te, b));
            });
            __syncthreads();
            if (has_rec) {
                const uint32_t br = 0xFFu - uint32_t(s_best[i] & 0xFFu);
                const uint32_t target = local_of(c, s_vid[br]);
                if (target != 0xFFFFFFFFu)
                    emit(c, row, atomicAdd(&s_ne, 1u), target, s_vid[i], s_P[tri_index(i, br)]);
            }
        } else if (has_rec) {
            // Closest candidate; equal distances go to the later position.
            float best = 0.f;
            int br = -1;
            for (int w = 0; w < 4; ++w)
                for (uint32_t m = rec[w]; m; m &= m - 1) {
                    const uint32_t r = 32 * w + __ffs(m) - 1;
                    const float v = s_P[tri_index(i, r)];
                    if (br < 0 || v <= best) {
                        best = v;
                        br = int(r);
                    }
                }
            const uint32_t target = local_of(c, s_vid[br]);
            if (target != 0xFFFFFFFFu) emit(c, row, atomicAdd(&s_ne, 1u), target, s_vid[i], best);
        }
        // ---- write back the retained entries, all marked old
        if (i < n && ((s_R[i >> 5] >> (i & 31)) & 1u)) {
            uint32_t pos = 0;
#pragma unroll
            for (int w = 0; w < 4; ++w) pos += __popc(s_R[w] & word_below(i, w));
            c.rows[size_t(row) * c.cap + pos] = make_uint2(s_vid[i], __float_as_uint(s_dist[i]));
        }
        __syncthreads();
        if (tid == 0) {
            c.cnt[row] = __popc(s_R[0]) + __popc(s_R[1]) + __popc(s_R[2]) + __popc(s_R[3]);
            c.e_cnt[row] = s_ne;
        }
        __syncthreads();
    }
    __shared__ unsigned long long s_pairs[2];
    if (tid == 0) s_pairs[0] = s_pairs[1] = 0;
    __syncthreads();
    atomicAdd(&s_pairs[0], exact_pairs);
    atomicAdd(&s_pairs[1], approx_pairs);
    __syncthreads();
    if (tid == 0) {
        atomicAdd(c.stats + 1, s_pairs[0]);
        atomicAdd(c.stats + 2, s_pairs[1]);
    }
}

// ------------------------------------------------------------ log bucketing
__global__ void scatter_kernel(Ctx c) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(c.n) * c.cap) return;
    const uint32_t row = uint32_t(i / c.cap), k = uint32_t(i % c.cap);
    if (k >= c.e_cnt[row]) return;
    const uint32_t t = c.e_t[i];
    if (t >= c.n) return;
    const uint32_t pos = c.t_off[t] + atomicAdd(c.t_fill + t, 1u);
    c.seg[pos] = c.e_vd[i];
}

// Orderable encoding of (distance, vid, origin): origin 0 = row entry.
__device__ __forceinline__ unsigned long long sort_key(float d, uint32_t vid, uint32_t origin) {
    uint32_t u = __float_as_uint(d == 0.0f ? 0.0f : d);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return (static_cast<unsigned long long>(u) << 32) | (static_cast<unsigned long long>(vid & kMask) << 1) | origin;
}

// The value of lane (lane ^ j), j a power of two below 64 (a constant once
// the sort's stages are unrolled): DPP quad permutes for 1 and 2, a row
// rotate by 8 within 16 lanes for 8, ds_swizzle xor masks within 32 lanes for
// 4 and 16, ds_bpermute for 32.
__device__ __forceinline__ uint32_t lane_xor(uint32_t x, uint32_t j) {
    switch (j) {
    case 1: return uint32_t(__builtin_amdgcn_mov_dpp(int(x), 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    case 2: return uint32_t(__builtin_amdgcn_mov_dpp(int(x), 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
    case 4: return uint32_t(__builtin_amdgcn_ds_swizzle(int(x), 0x101F));              // and 0x1F, xor 4
    case 8: return uint32_t(__builtin_amdgcn_mov_dpp(int(x), 0x128, 0xF, 0xF, false)); // row_ror:8
    case 16: return uint32_t(__builtin_amdgcn_ds_swizzle(int(x), 0x401F));             // and 0x1F, xor 16
    default: return __shfl_xor(x, int(j));
    }
}

// Bitonic network over 64 * R (key, payload) pairs held four to a lane at
// most: element r * 64 + lane in k[r], p[r].  Stages with j >= 64 pair
// registers of one lane, the others pair lanes (lane ^ j).  Equal keys carry
// equal payloads here (a key holds the vid and its origin), so the network's
// instability does not show in the result.
template <int R>
__device__ __forceinline__ void wave_bitonic(unsigned long long (&k)[4], uint32_t (&p)[4]) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (uint32_t k2 = 2; k2 <= 64u * R; k2 <<= 1) {
#pragma unroll
        for (uint32_t j = k2 >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const uint32_t rj = j / 64;
#pragma unroll
                for (uint32_t r = 0; r < uint32_t(R); ++r) {
                    if (r & rj) continue;
                    const uint32_t r2 = r | rj;
                    const bool asc = ((r * 64 + lane) & k2) == 0;
                    if (asc ? (k[r] > k[r2]) : (k[r] < k[r2])) {
                        const unsigned long long tk = k[r];
                        k[r] = k[r2];
                        k[r2] = tk;
                        const uint32_t tp = p[r];
                        p[r] = p[r2];
                        p[r2] = tp;
                    }
                }
            } else {
#pragma unroll
                for (uint32_t r = 0; r < uint32_t(R); ++r) {
                    const uint32_t lo = lane_xor(uint32_t(k[r]), j), hi = lane_xor(uint32_t(k[r] >> 32), j);
                    const uint32_t op = lane_xor(p[r], j);
                    const unsigned long long o = (static_cast<unsigned long long>(hi) << 32) | lo;
                    // The lower element of the pair keeps the minimum when the
                    // k2-block sorts ascending, the upper one otherwise.
                    const bool take_min = ((lane & j) == 0) == (((r * 64 + lane) & k2) == 0);
                    if (take_min ? (o < k[r]) : (o > k[r])) {
                        k[r] = o;
                        p[r] = op;
                    }
                }
            }
        }
    }
}

// apply_logs merge for rows that received logs (refiner/nbr_log_table.hpp:
// 127-182): one wavefront per row.  The row's entries and a run of its logs
// are sorted together by (distance, vid, origin) in registers, over the next
// power of two >= their count (64, 128 or 256); the first element of each vid
// is kept (origin 0 = the row's entry sorts first: the row wins ties) and the
// first cap survivors stay.  Logs beyond kMergeN - (survivors) take further
// rounds against the survivors, which selects the same entries as one sort of
// everything: an entry cut in a round has cap smaller unique entries that stay.
__global__ __launch_bounds__(256) void merge_kernel(Ctx c) {
    __shared__ unsigned long long s_key[4][kMergeN];
    __shared__ uint32_t s_pay[4][kMergeN];
    const uint32_t lane = lane_id(), wid = threadIdx.x / 64;
    unsigned long long* key = s_key[wid];
    uint32_t* pay = s_pay[wid];
    for (uint32_t row = blockIdx.x * 4 + wid; row < c.n; row += gridDim.x * 4) {
        const uint32_t m = c.t_cnt[row];
        if (m == 0) continue;  // uniform across the wavefront
        const uint32_t base = c.t_off[row];
        uint32_t cur = c.cnt[row];
        for (uint32_t k = lane; k < cur; k += 64) {
            const uint2 e = c.rows[size_t(row) * c.cap + k];
            key[k] = sort_key(__uint_as_float(e.y), e.x, 0);
            pay[k] = e.x;
        }
        for (uint32_t off = 0; off < m;) {
            const uint32_t nc = min(m - off, kMergeN - cur), total = cur + nc;
            for (uint32_t k = lane; k < nc; k += 64) {
                const uint2 lg = c.seg[base + off + k];
                key[cur + k] = sort_key(__uint_as_float(lg.y), lg.x, 1);
                pay[cur + k] = (lg.x & kMask) | kNew;
            }
            wave_sync();
            unsigned long long kr[4];
            uint32_t pr[4];
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) {
                const uint32_t q = r * 64 + lane;
                kr[r] = q < total ? key[q] : ~0ull;
                pr[r] = q < total ? pay[q] : 0xFFFFFFFFu;
            }
            wave_sync();
            if (total <= 64)
                wave_bitonic<1>(kr, pr);
            else if (total <= 128)
                wave_bitonic<2>(kr, pr);
            else
                wave_bitonic<4>(kr, pr);
            // Adjacent unique by vid, the first cap survivors compacted to the
            // front of the wavefront's arrays.
            uint32_t outn = 0, last = 0xFFFFFFFFu;  // last: payload of element r * 64 - 1
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) {
                const uint32_t q = r * 64 + lane;
                uint32_t prev = __shfl_up(pr[r], 1u);
                if (lane == 0) prev = last;
                last = __shfl(pr[r], 63);
                const bool keep = q < total && (q == 0 || (prev & kMask) != (pr[r] & kMask));
                const uint64_t km = ballot(keep);
                const uint32_t dst = outn + lanes_below(km);
                if (keep && dst < c.cap) {
                    key[dst] = kr[r];
                    pay[dst] = pr[r];
                }
                outn += uint32_t(__popcll(km));
            }
            wave_sync();
            cur = min(outn, c.cap);
            off += nc;
        }
        for (uint32_t k = lane; k < cur; k += 64) {
            const unsigned long long kk = key[k];
            uint32_t u = uint32_t(kk >> 32);
            u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
            c.rows[size_t(row) * c.cap + k] = make_uint2(pay[k], u);
        }
        if (lane == 0) c.cnt[row] = cur;
        wave_sync();
    }
}

// fill_single_layer_graph_from_layer (refiner/refiner_utils.hpp:60-108):
// each local row takes its vertex's row of `src` (by vid, src_cap slots),
// sorted by (distance, vid), cut to cap.  A wavefront per row.
__global__ __launch_bounds__(256) void import_rows_kernel(Ctx c, const uint2* src, const uint32_t* src_cnt,
                                                          uint32_t src_cap) {
    __shared__ unsigned long long s_key[4][kMaxCap];
    __shared__ uint32_t s_pay[4][kMaxCap];
    const uint32_t wid = threadIdx.x / 64, lane = lane_id();
    for (uint32_t row = blockIdx.x * 4 + wid; row < c.n; row += gridDim.x * 4) {
        const uint32_t v = storage_of(c, row);
        const uint32_t m = src ? min(src_cnt[v], src_cap) : 0u;
        for (uint32_t j = lane; j < kMaxCap; j += 64) {
            if (j < m) {
                const uint2 e = src[size_t(v) * src_cap + j];
                s_key[wid][j] = sort_key(__uint_as_float(e.y), e.x, 0);
                s_pay[wid][j] = e.x;
            } else {
                s_key[wid][j] = ~0ull;
                s_pay[wid][j] = 0xFFFFFFFFu;
            }
        }
        wave_sync();
        // bitonic sort of kMaxCap keys by one wavefront
        for (uint32_t k2 = 2; k2 <= kMaxCap; k2 <<= 1) {
            for (uint32_t j = k2 >> 1; j > 0; j >>= 1) {
                for (uint32_t t = lane; t < kMaxCap / 2; t += 64) {
                    const uint32_t i = 2 * j * (t / j) + (t % j), l = i + j;
                    const bool up = (i & k2) == 0;
                    const unsigned long long a = s_key[wid][i], b = s_key[wid][l];
                    if ((a > b) == up) {
                        s_key[wid][i] = b;
                        s_key[wid][l] = a;
                        const uint32_t pa = s_pay[wid][i];
                        s_pay[wid][i] = s_pay[wid][l];
                        s_pay[wid][l] = pa;
                    }
                }
                wave_sync();
            }
        }
        const uint32_t keep = min(m, c.cap);
        for (uint32_t j = lane; j < keep; j += 64) {
            uint32_t u = uint32_t(s_key[wid][j] >> 32);
            u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
            c.rows[size_t(row) * c.cap + j] = make_uint2(s_pay[wid][j], u);
        }
        if (lane == 0) c.cnt[row] = keep;
        wave_sync();
    }
}

__global__ void count_logs_kernel(const uint32_t* t_off, uint32_t n, unsigned long long* stats) {
    if (threadIdx.x == 0 && blockIdx.x == 0) stats[0] += t_off[n];
}

}  // namespace

struct artea_refine {
    Ctx c{};
    hipStream_t s = nullptr;
    uint64_t seed = 0;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    uint32_t* d_l2g = nullptr;
    uint32_t* d_g2l = nullptr;
    int dev_cus = 0;
    int join = ARTEA_JOIN_FP32;
    // routing loops since the last artea_refine_routing_stats: distances,
    // expansions, rows whose candidate list still dropped a tied entry
    unsigned long long route_stats[3] = {0, 0, 0};
};

namespace {

void refine_free(artea_refine* r) {
    Ctx& c = r->c;
    for (void* p : {(void*)c.rows, (void*)c.cnt, (void*)c.e_t, (void*)c.e_vd, (void*)c.e_cnt, (void*)c.t_cnt,
                    (void*)c.t_off, (void*)c.t_fill, (void*)c.seg, (void*)c.stats, r->scan_tmp, (void*)r->d_l2g,
                    (void*)r->d_g2l})
        if (p) (void)hipFree(p);
    if (r->s) (void)hipStreamDestroy(r->s);
}

uint32_t grid_for(const artea_refine* r, uint32_t rows_per_block, uint32_t blocks_per_cu, uint32_t rows = ~0u) {
    const uint32_t want = (std::min(rows, r->c.n) + rows_per_block - 1) / rows_per_block;
    return std::max(1u, std::min(want, uint32_t(r->dev_cus) * blocks_per_cu));
}

// Bucket, merge and account the logs emitted by the last producer kernel.
int merge_logs(artea_refine* r) {
    Ctx& c = r->c;
    hipStream_t s = r->s;
    ARTEA_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(r->scan_tmp, r->scan_bytes, c.t_cnt, c.t_off, c.n + 1, s));
    const size_t slots = size_t(c.n) * c.cap;
    const uint32_t blocks = uint32_t((slots + 255) / 256);
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(256), 0, s, c);
    hipLaunchKernelGGL(merge_kernel, dim3(grid_for(r, 4, 8)), dim3(256), 0, s, c);
    hipLaunchKernelGGL(count_logs_kernel, dim3(1), dim3(64), 0, s, c.t_off, c.n, c.stats);
    ARTEA_HIP_TRY(hipGetLastError());
    ARTEA_HIP_TRY(hipMemsetAsync(c.t_cnt, 0, sizeof(uint32_t) * (c.n + 1), s));
    ARTEA_HIP_TRY(hipMemsetAsync(c.t_fill, 0, sizeof(uint32_t) * c.n, s));
    return ARTEA_OK;
}

int do_prefill(artea_refine* r, uint32_t gen, uint32_t lo, uint32_t hi, uint32_t rb, uint32_t re, uint64_t stream) {
    Ctx& c = r->c;
    ARTEA_CHECK_ARG(lo <= hi && hi <= c.n, "prefill: id window must lie in [0, num_rows]");
    ARTEA_CHECK_ARG(gen <= c.cap, "prefill: gen_size exceeds the row capacity");
    re = std::min(re, c.n);
    if (rb >= re || gen == 0 || lo == hi) return ARTEA_OK;
    ARTEA_HIP_TRY(hipMemsetAsync(c.e_cnt, 0, sizeof(uint32_t) * c.n, r->s));
    hipLaunchKernelGGL(prefill_kernel, dim3(grid_for(r, 4, 8)), dim3(256), 0, r->s, c, gen, lo, hi, rb, re, r->seed,
                       stream);
    ARTEA_HIP_TRY(hipGetLastError());
    return merge_logs(r);
}

// Interval half-widths of the BF16 join for dimension D (each a proven
// bound times 2).  Per coordinate the split keeps |x - hi - lo| <= 2^-16 |x|,
// so |ab - (ah bh + ah bl + al bh)| <= 3.1 * 2^-16 |a||b|; the 3D products are
// summed in FP32 (<= 3.03 D u of the absolute sum, u = 2^-23 to allow any
// rounding of the accumulator); |sum a_k b_k| <= sqrt(N_a N_b) <= (N_a + N_b) / 2.
// The squared norms carry D u, forming N_a + N_b - 2G 6u, the centring
// 2^-22, and the FP32 chain itself (D/16 + 6) u of the true value.
JoinBounds join_bounds(uint32_t D) {
    const double u = 0x1p-23, d = double(D);
    const double kg = 3.1 * 0x1p-16 + 3.03 * d * u;
    JoinBounds b;
    b.l2 = float(2.0 * (kg + d * u + 6.0 * u + 0x1p-22 + 2.0 * (d / 16.0 + 6.0) * u));
    b.dot = float(2.0 * (kg + (d / 16.0 + 4.0) * u + d * u));
    b.cos = float(2.0 * (kg + (d / 16.0 + 4.0) * u + 2.0 * d * u + 24.0 * u));
    return b;
}

int do_triangle(artea_refine* r, float scale, float shift, uint32_t rb = 0, uint32_t re = ~0u) {
    Ctx c = r->c;
    c.re = std::min(re, c.n);
    c.rb = std::min(rb, c.re);
    if (c.rb > 0 || c.re < c.n) ARTEA_HIP_TRY(hipMemsetAsync(c.e_cnt, 0, sizeof(uint32_t) * c.n, r->s));
    const float inv = 1.0f / scale;
    const float shift_term = (1.0f + inv) * shift;
    int dev = 0, max_lds = 65536;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    // Rows of at most kMfmaMinCap entries run the FP32 chain in either mode:
    // their pair counts are small and the FP32 kernel's lower register count
    // keeps twice the rows per CU in flight (measured faster at cap 48).
    const bool mfma = r->join == ARTEA_JOIN_BF16 && c.cap > kMfmaMinCap &&
                      tri_layout(c.cap, c.D, true, true).total <= uint32_t(max_lds);
    const bool staged = mfma || tri_layout(c.cap, c.D, true, false).total <= uint32_t(max_lds);
    const size_t lds = tri_layout(c.cap, c.D, staged, mfma).total;
    ARTEA_CHECK_ARG(lds <= size_t(max_lds), "triangle: row capacity too large for LDS");
    const JoinBounds kb = join_bounds(c.D);
    const uint32_t per_cu = std::max<uint32_t>(1u, uint32_t(160u * 1024u / std::max<size_t>(lds, 1)));
    const dim3 grid(grid_for(r, 1, std::min<uint32_t>(per_cu, 4u), std::max(c.re - c.rb, 1u)));
    auto launch = [&](auto kernel) -> int {
        const void* fn = reinterpret_cast<const void*>(kernel);
        if (lds > 65536) ARTEA_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        hipLaunchKernelGGL(kernel, grid, dim3(256), lds, r->s, c, inv, shift_term, staged ? 1 : 0, kb);
        return ARTEA_OK;
    };
    int rc = ARTEA_OK;
    switch (c.metric * 2 + (mfma ? 1 : 0)) {
        case kL2Sqr * 2: rc = launch(triangle_kernel<false, kL2Sqr>); break;
        case kL2Sqr * 2 + 1: rc = launch(triangle_kernel<true, kL2Sqr>); break;
        case kDot * 2: rc = launch(triangle_kernel<false, kDot>); break;
        case kDot * 2 + 1: rc = launch(triangle_kernel<true, kDot>); break;
        case kCosine * 2: rc = launch(triangle_kernel<false, kCosine>); break;
        case kCosine * 2 + 1: rc = launch(triangle_kernel<true, kCosine>); break;
        case kL2 * 2: rc = launch(triangle_kernel<false, kL2>); break;
        default: rc = launch(triangle_kernel<true, kL2>); break;
    }
    if (rc) return rc;
    ARTEA_HIP_TRY(hipGetLastError());
    return merge_logs(r);
}

int do_reverse(artea_refine* r, uint32_t rb = 0, uint32_t re = ~0u) {
    Ctx c = r->c;
    c.re = std::min(re, c.n);
    c.rb = std::min(rb, c.re);
    const size_t slots = size_t(c.n) * c.cap;
    hipLaunchKernelGGL(reverse_kernel, dim3(uint32_t((slots + 255) / 256)), dim3(256), 0, r->s, c);
    ARTEA_HIP_TRY(hipGetLastError());
    return merge_logs(r);
}

int do_schedule(artea_refine* r, uint32_t loops, uint32_t triu, float scale, float shift) {
    for (uint32_t b = 0; b < loops; ++b) {
        for (uint32_t t = 0; t < triu; ++t) {
            int rc = do_triangle(r, scale, shift);
            if (rc) return rc;
            // arc (truncate): rows are already capped at `cap` by the merge.
        }
        if (b + 1 < loops) {
            int rc = do_reverse(r);
            if (rc) return rc;
        }
    }
    return ARTEA_OK;
}

// The reverse logs of a routing pass (counted by routing_finish) merged into
// the rows: merge_logs with the pass's log buffer; the merge keeps the first
// cap entries, i.e. the arc updater that follows.
int merge_routing_reverse(artea_refine* r, const RoutingPass& p, uint2* seg) {
    Ctx c = r->c;
    c.seg = seg;
    hipStream_t s = r->s;
    ARTEA_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(r->scan_tmp, r->scan_bytes, c.t_cnt, c.t_off, c.n + 1, s));
    ARTEA_HIP_TRY(routing_scatter(c, p, seg, s));
    hipLaunchKernelGGL(merge_kernel, dim3(grid_for(r, 4, 8)), dim3(256), 0, s, c);
    hipLaunchKernelGGL(count_logs_kernel, dim3(1), dim3(64), 0, s, c.t_off, c.n, c.stats);
    ARTEA_HIP_TRY(hipGetLastError());
    ARTEA_HIP_TRY(hipMemsetAsync(c.t_cnt, 0, sizeof(uint32_t) * (c.n + 1), s));
    ARTEA_HIP_TRY(hipMemsetAsync(c.t_fill, 0, sizeof(uint32_t) * c.n, s));
    return ARTEA_OK;
}

int check_routing(const artea_refine* r, uint32_t topk, uint32_t queue, float scale, const char* who) {
    ARTEA_CHECK_ARG(r, std::string(who) + ": null handle");
    ARTEA_CHECK_ARG(topk >= 1 && topk <= 128, std::string(who) + ": routing topk must be in [1, 128]");
    if (queue < topk) {
        set_error(std::string(who) + ": candidate_queue_size (" + std::to_string(queue) + ") must be >= topk (" +
                  std::to_string(topk) + ")");
        return ARTEA_EINVAL;
    }
    ARTEA_CHECK_ARG(queue <= 256, std::string(who) + ": routing queue must be <= 256");
    ARTEA_CHECK_ARG(scale > 0.0f, std::string(who) + ": scale must be positive");
    return ARTEA_OK;
}

// The routing loops of refine_layer: next(routing).next(pruning), then
// next(reverse).next(arc) between loops.  Each routing pass merges and
// prunes every row into ext (up to cap + topk entries); the rows then take
// their first cap entries (after the last loop: the writeback keeps the
// first slots anyway), and between loops the reverse logs of the full ext
// rows are merged into them.
int do_routing_loops(artea_refine* r, uint32_t loops, uint32_t topk, uint32_t queue, float scale) {
    Ctx& c = r->c;
    if (loops == 0 || c.n == 0) return ARTEA_OK;
    RoutingPass p{};
    p.topk = topk;
    p.queue = queue;
    p.inv_scale = 1.0f / scale;
    p.ext_stride = c.cap + topk;
    const size_t slots = size_t(c.n) * p.ext_stride;
    uint2* seg = nullptr;
    hipError_t e = hipMalloc(&p.ext, slots * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc(&p.ext_cnt, size_t(c.n) * 4);
    if (e == hipSuccess) e = hipMalloc(&p.stats, kRouteStatsBytes);
    if (e == hipSuccess && loops > 1) e = hipMalloc(&seg, slots * sizeof(uint2));
    struct Free {
        void* q[4];
        ~Free() {
            for (void* x : q)
                if (x) (void)hipFree(x);
        }
    } guard{{p.ext, p.ext_cnt, p.stats, seg}};
    ARTEA_HIP_TRY(e);
    for (uint32_t l = 0; l < loops; ++l) {
        const bool last = l + 1 == loops;
        uint32_t unresolved = 0;
        ARTEA_HIP_TRY(routing_pass(c, p, r->s, r->dev_cus, &unresolved));
        unsigned long long h[2];
        ARTEA_HIP_TRY(hipMemcpy(h, p.stats, sizeof h, hipMemcpyDeviceToHost));
        r->route_stats[0] += h[0];
        r->route_stats[1] += h[1];
        r->route_stats[2] += unresolved;
        ARTEA_HIP_TRY(routing_finish(c, p, !last, r->s));
        if (!last) {
            int rc = merge_routing_reverse(r, p, seg);
            if (rc) return rc;
        }
    }
    ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return ARTEA_OK;
}

int refine_make(const float* d_vecs, uint32_t n_global, uint32_t dim_pad, int metric, uint32_t cap,
                const uint32_t* l2g, uint32_t n_local, uint64_t seed, artea_refine** out) {
    ARTEA_CHECK_ARG(out, "artea_refine_create: null output");
    ARTEA_CHECK_ARG(d_vecs || n_global == 0, "artea_refine_create: null vectors");
    ARTEA_CHECK_ARG(dim_pad > 0 && dim_pad % 16 == 0, "artea_refine_create: dim_pad must be a multiple of 16");
    ARTEA_CHECK_ARG(metric >= 0 && metric <= 3, "artea_refine_create: metric must be 0..3");
    ARTEA_CHECK_ARG(cap >= 1 && cap <= kMaxCap, "artea_refine_create: cap must be in [1, 128]");
    auto r = std::make_unique<artea_refine>();
    Ctx& c = r->c;
    c.vecs = d_vecs;
    c.n_global = n_global;
    c.D = dim_pad;
    c.C = dim_pad / 16;
    c.metric = metric;
    c.cap = cap;
    c.n = l2g ? n_local : n_global;
    c.rb = 0;
    c.re = c.n;
    c.list_cap = kListCap;
    r->seed = seed;
    const size_t slots = std::max<size_t>(size_t(c.n) * cap, 1);
    auto fail = [&](hipError_t e) {
        refine_free(r.get());
        r->c = Ctx{};
        set_error(std::string("artea_refine_create: allocation failed: ") + hipGetErrorString(e));
        return ARTEA_ENOMEM;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&r->s, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.rows, slots * sizeof(uint2))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.cnt, (c.n + 1) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.e_t, slots * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.e_vd, slots * sizeof(uint2))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.e_cnt, (c.n + 1) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.t_cnt, (c.n + 2) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.t_off, (c.n + 2) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.t_fill, (c.n + 1) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.seg, slots * sizeof(uint2))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&c.stats, 4 * sizeof(unsigned long long))) != hipSuccess) return fail(e);
    if (l2g) {
        std::vector<uint32_t> g2l(std::max<uint32_t>(n_global, 1), 0xFFFFFFFFu);
        for (uint32_t i = 0; i < n_local; ++i) {
            if (l2g[i] >= n_global) {
                refine_free(r.get());
                set_error("artea_refine_create: local_to_global id out of range");
                return ARTEA_EINVAL;
            }
            g2l[l2g[i]] = i;
        }
        if ((e = hipMalloc(&r->d_l2g, std::max<size_t>(n_local, 1) * 4)) != hipSuccess) return fail(e);
        if ((e = hipMalloc(&r->d_g2l, g2l.size() * 4)) != hipSuccess) return fail(e);
        if ((e = hipMemcpy(r->d_l2g, l2g, size_t(n_local) * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
        if ((e = hipMemcpy(r->d_g2l, g2l.data(), g2l.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
        c.l2g = r->d_l2g;
        c.g2l = r->d_g2l;
    }
    (void)hipMemset(c.cnt, 0, (c.n + 1) * 4);
    (void)hipMemset(c.e_cnt, 0, (c.n + 1) * 4);
    (void)hipMemset(c.t_cnt, 0, (c.n + 2) * 4);
    (void)hipMemset(c.t_fill, 0, (c.n + 1) * 4);
    (void)hipMemset(c.stats, 0, 4 * sizeof(unsigned long long));
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, r->scan_bytes, c.t_cnt, c.t_off, c.n + 1)) != hipSuccess)
        return fail(e);
    if ((e = hipMalloc(&r->scan_tmp, std::max<size_t>(r->scan_bytes, 16))) != hipSuccess) return fail(e);
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&r->dev_cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (r->dev_cus <= 0) r->dev_cus = 256;
    (void)hipDeviceSynchronize();
    *out = r.release();
    return ARTEA_OK;
}

// FP64 centroid partial sums in a fixed order (deterministic), then argmin
// of dist(centroid, v) with ties to the smaller vid.
__global__ void centroid_partial_kernel(const float* vecs, const uint32_t* ids, uint32_t n, uint32_t D,
                                        uint32_t per_block, double* part) {
    const uint32_t b = blockIdx.x;
    const uint32_t lo = b * per_block, hi = min(n, lo + per_block);
    for (uint32_t d = threadIdx.x; d < D; d += blockDim.x) {
        double s = 0;
        for (uint32_t i = lo; i < hi; ++i) s += double(vecs[size_t(ids[i]) * D + d]);
        part[size_t(b) * D + d] = s;
    }
}
// Source level of vertex v: with level_fraction f = 1 every vertex is on
// every level 0..max_level; otherwise the highest h <= max_level with
// u(v) < f^h, u(v) a seeded uniform draw in [0, 1) (a geometric level
// assignment standing in for the r-net insertion, which is not built here).
uint32_t build_level_of(uint64_t seed, uint32_t v, uint32_t max_level, float f) {
    if (f >= 1.0f) return max_level;
    uint64_t z = seed ^ 0x6C6576656C73ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z += uint64_t(v) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = double(z >> 11) * 0x1.0p-53;
    uint32_t h = 0;
    double p = double(f);
    while (h < max_level && u < p) {
        ++h;
        p *= double(f);
    }
    return h;
}

__global__ void centroid_final_kernel(const double* part, uint32_t blocks, uint32_t n, uint32_t D, float* out) {
    for (uint32_t d = threadIdx.x; d < D; d += blockDim.x) {
        double s = 0;
        for (uint32_t b = 0; b < blocks; ++b) s += part[size_t(b) * D + d];
        out[d] = float(s * (1.0 / double(n)));
    }
}


// HierarchicalGraphCompactor::compact_graph
// (compactor/hierarchical_graph_compactor.hpp:118-327) of a dynamic graph
// given per vertex its source level (kInvalidLevel = unassigned) and per
// level l its rows (row_off[l]: n + 1 offsets into row_ids[l]), attached to
// idx.  Step 1: the first bucket from the top with >= 32 vertices
// (min_layer_cap, framework/type_traits/index_traits.hpp:79) is the new top.
// Step 2: buckets keep the order vertices were assigned (`order`, null =
// ascending vid); the top bucket holds its own vertices, then the demoted
// ones by ascending source level, each keeping its rows at levels <= top.
// Groups are laid out h = 0..top in the neighbour array, a vertex's rows from
// its highest level down.  Entry: the top-bucket vertex closest, by
// `metric`, to the FP64 centroid of the top bucket, computed on the GPU;
// ties keep the earliest bucket position.
// The compactor reads the ids of each row's first min(cnt, slot_cap)
// entries: they are packed slot_cap apart on the device and only they are
// copied back.  Packed row i is rows[map ? map[i] : i].
__global__ void pack_ids_kernel(const uint2* rows, const uint32_t* cnt, const uint32_t* map, uint32_t nrows,
                                uint32_t stride, uint32_t slot_cap, uint32_t* ids, uint32_t* out_cnt) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(nrows) * slot_cap) return;
    const uint32_t i = uint32_t(t / slot_cap), k = uint32_t(t % slot_cap);
    const uint32_t src = map ? map[i] : i;
    const uint32_t c = min(cnt[src], slot_cap);
    if (k == 0) out_cnt[i] = c;
    if (k < c) ids[t] = rows[size_t(src) * stride + k].x & kMask;
}

// Level l of the compactor's input from nrows device rows (stride entries
// apart; row of members[i] at map[i], or at i without a map): row_off over
// all n vertices and the ids back to back in vid order.  The packed ids come
// back through two pinned staging halves of kStageBytes, each half's rows
// placed into their final slots while the next half is in flight.
constexpr size_t kStageBytes = size_t(64) << 20;

int fetch_level(const uint2* d_rows, const uint32_t* d_cnt, const uint32_t* map, const std::vector<uint32_t>& members,
                uint32_t n, uint32_t stride, uint32_t slot_cap, hipStream_t s, std::vector<uint64_t>& roff,
                std::unique_ptr<uint32_t[]>& rids) {
    const uint32_t nl = uint32_t(members.size());
    roff.assign(size_t(n) + 1, 0);
    if (nl == 0) {
        rids.reset(new uint32_t[1]);
        return ARTEA_OK;
    }
    const uint32_t chunk_rows =
        std::min(nl, uint32_t(std::max<size_t>(1, kStageBytes / (size_t(std::max(slot_cap, 1u)) * 4))));
    const size_t chunk_elems = size_t(chunk_rows) * slot_cap;
    uint32_t *d_ids = nullptr, *d_out = nullptr, *d_map = nullptr, *stage = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    auto release = [&]() {
        (void)hipStreamSynchronize(s);  // no copy may still target the staging halves
        for (void* q : {(void*)d_ids, (void*)d_out, (void*)d_map})
            if (q) (void)hipFree(q);
        if (stage) (void)hipHostFree(stage);
        for (hipEvent_t x : ev)
            if (x) (void)hipEventDestroy(x);
    };
    std::vector<uint32_t> cnt(nl);
    hipError_t e = hipMalloc(&d_ids, size_t(nl) * slot_cap * 4);
    if (e == hipSuccess) e = hipMalloc(&d_out, size_t(nl) * 4);
    if (e == hipSuccess && map) e = hipMalloc(&d_map, size_t(nl) * 4);
    if (e == hipSuccess && map) e = hipMemcpyAsync(d_map, map, size_t(nl) * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&stage), 2 * chunk_elems * 4, hipHostMallocDefault);
    for (int h = 0; h < 2 && e == hipSuccess; ++h) e = hipEventCreateWithFlags(&ev[h], hipEventDisableTiming);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pack_ids_kernel, dim3(uint32_t((size_t(nl) * slot_cap + 255) / 256)), dim3(256), 0, s,
                           d_rows, d_cnt, d_map, nl, stride, slot_cap, d_ids, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_out, size_t(nl) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) {
        for (uint32_t i = 0; i < nl; ++i) roff[members[i] + 1] = cnt[i];
        for (uint32_t v = 0; v < n; ++v) roff[v + 1] += roff[v];
        rids.reset(new uint32_t[std::max<uint64_t>(roff[n], 1)]);
        const uint32_t chunks = (nl + chunk_rows - 1) / chunk_rows;
        auto issue = [&](uint32_t c) {
            const uint32_t r0 = c * chunk_rows, r1 = std::min(nl, r0 + chunk_rows);
            hipError_t x = hipMemcpyAsync(stage + (c & 1u) * chunk_elems, d_ids + size_t(r0) * slot_cap,
                                          size_t(r1 - r0) * slot_cap * 4, hipMemcpyDeviceToHost, s);
            return x == hipSuccess ? hipEventRecord(ev[c & 1u], s) : x;
        };
        auto place = [&](uint32_t c) {
            hipError_t x = hipEventSynchronize(ev[c & 1u]);
            if (x != hipSuccess) return x;
            const uint32_t r0 = c * chunk_rows, r1 = std::min(nl, r0 + chunk_rows);
            const uint32_t* src = stage + (c & 1u) * chunk_elems;
#pragma omp parallel for schedule(static)
            for (int64_t i = r0; i < int64_t(r1); ++i) {
                const uint32_t v = members[size_t(i)];
                memcpy(rids.get() + roff[v], src + size_t(i - r0) * slot_cap, size_t(roff[v + 1] - roff[v]) * 4);
            }
            return hipSuccess;
        };
        e = issue(0);
        for (uint32_t c = 1; c < chunks && e == hipSuccess; ++c) {
            e = issue(c);
            if (e == hipSuccess) e = place(c - 1);
        }
        if (e == hipSuccess) e = place(chunks - 1);
    }
    release();
    if (e != hipSuccess) {
        set_error(std::string("compactor input: ") + hipGetErrorString(e));
        return ARTEA_EHIP;
    }
    return ARTEA_OK;
}

int compact_and_attach(artea_index* idx, const uint32_t* src_h, const uint32_t* order,
                       const uint64_t* const* row_off, const uint32_t* const* row_ids, uint32_t ul_max,
                       uint32_t bl_max, int metric) {
    const uint32_t n = idx->n;
    uint32_t src_top = kInvalidLevel;
    for (uint32_t v = 0; v < n; ++v)
        if (src_h[v] != kInvalidLevel && (src_top == kInvalidLevel || src_h[v] > src_top)) src_top = src_h[v];
    HostCsr g;
    g.n = n;
    g.ul_max = ul_max;
    g.bl_max = bl_max;
    g.highest.assign(n, kInvalidLevel);
    g.local.assign(n, kNoVertex);
    if (src_top == kInvalidLevel) {  // no assigned vertex: an empty graph, no entry
        g.top = 0;
        g.offsets.assign(1, std::vector<uint64_t>(1, 0));
        std::vector<uint32_t> vi(std::max<size_t>(size_t(n) * 2, 2), kInvalidLevel);
        const uint64_t* p0 = g.offsets[0].data();
        return artea_index_set_graph_csr(idx, 0, ul_max, bl_max, kNoVertex, vi.data(), &p0, nullptr, 0);
    }
    std::vector<uint64_t> bucket_n(src_top + 1, 0);
    for (uint32_t v = 0; v < n; ++v)
        if (src_h[v] != kInvalidLevel) bucket_n[src_h[v]]++;
    uint32_t top = 0;
    for (uint32_t h = src_top;; --h) {
        if (bucket_n[h] >= 32) {
            top = h;
            break;
        }
        if (h == 0) break;
    }
    g.top = top;
    std::vector<std::vector<uint32_t>> bucket(top + 1);
    for (uint32_t sh = 0; sh <= src_top; ++sh) {
        const uint32_t grp = std::min(sh, top);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t v = order ? order[i] : i;
            if (src_h[v] == sh) {
                g.highest[v] = grp;
                g.local[v] = uint32_t(bucket[grp].size());
                bucket[grp].push_back(v);
            }
        }
    }
    // Group offsets first (rows of a vertex from its highest level down),
    // then the ids copied in parallel.
    g.offsets.resize(top + 1);
    uint64_t total = 0;
    for (uint32_t h = 0; h <= top; ++h) {
        auto& o = g.offsets[h];
        o.assign(bucket[h].size() * (h + 1) + 1, 0);
        for (size_t loc = 0; loc < bucket[h].size(); ++loc) {
            const uint32_t v = bucket[h][loc];
            for (uint32_t l = h;; --l) {
                o[loc * (h + 1) + (h - l)] = total;
                total += row_off[l][v + 1] - row_off[l][v];
                if (l == 0) break;
            }
        }
        o.back() = total;
    }
    g.nbrs.resize(total);
    for (uint32_t h = 0; h <= top; ++h) {
        const auto& o = g.offsets[h];
        const std::vector<uint32_t>& bk = bucket[h];
#pragma omp parallel for schedule(static)
        for (int64_t loc = 0; loc < int64_t(bk.size()); ++loc) {
            const uint32_t v = bk[size_t(loc)];
            for (uint32_t l = h;; --l) {
                const uint64_t a = row_off[l][v], b = row_off[l][v + 1];
                memcpy(g.nbrs.data() + o[size_t(loc) * (h + 1) + (h - l)], row_ids[l] + a, size_t(b - a) * 4);
                if (l == 0) break;
            }
        }
    }
    {
        const std::vector<uint32_t>& tb = bucket[top];
        const uint32_t per = 4096, blocks = std::max(1u, uint32_t((tb.size() + per - 1) / per));
        double* part = nullptr;
        float* cen = nullptr;
        float* dists = nullptr;
        uint32_t* d_ids = nullptr;
        auto release = [&]() {
            for (void* q : {(void*)part, (void*)cen, (void*)dists, (void*)d_ids})
                if (q) (void)hipFree(q);
        };
        hipError_t e = hipMalloc(&part, size_t(blocks) * idx->dim_pad * 8);
        if (e == hipSuccess) e = hipMalloc(&cen, size_t(idx->dim_pad) * 4);
        if (e == hipSuccess) e = hipMalloc(&dists, std::max<size_t>(n, 1) * 4);
        if (e == hipSuccess) e = hipMalloc(&d_ids, std::max<size_t>(tb.size(), 1) * 4);
        if (e == hipSuccess) e = hipMemcpyAsync(d_ids, tb.data(), tb.size() * 4, hipMemcpyHostToDevice, idx->stream);
        if (e != hipSuccess) {
            release();
            set_error(std::string("compact_graph: device buffers failed: ") + hipGetErrorString(e));
            return ARTEA_EHIP;
        }
        hipLaunchKernelGGL(centroid_partial_kernel, dim3(blocks), dim3(256), 0, idx->stream, idx->d_vecs, d_ids,
                           uint32_t(tb.size()), idx->dim_pad, per, part);
        hipLaunchKernelGGL(centroid_final_kernel, dim3(1), dim3(256), 0, idx->stream, part, blocks,
                           uint32_t(tb.size()), idx->dim_pad, cen);
        e = hipStreamSynchronize(idx->stream);
        int rc = e == hipSuccess ? artea_distance_matrix_device(metric, cen, 1, idx->d_vecs, n, idx->dim_pad, dists,
                                                                idx->stream)
                                 : ARTEA_EHIP;
        std::vector<float> h(n);
        if (rc == ARTEA_OK) {
            e = hipMemcpyAsync(h.data(), dists, size_t(n) * 4, hipMemcpyDeviceToHost, idx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(idx->stream);
            if (e != hipSuccess) rc = ARTEA_EHIP;
        }
        release();
        if (rc) {
            if (e != hipSuccess) set_error(std::string("compact_graph: centroid failed: ") + hipGetErrorString(e));
            return rc;
        }
        float best = FLT_MAX;
        uint32_t bv = kNoVertex;
        for (uint32_t v : tb)
            if (h[v] < best) {
                best = h[v];
                bv = v;
            }
        g.entry = bv;
    }
    return attach_csr(idx, std::move(g));
}
}  // namespace

extern "C" {

int artea_refine_create(const float* d_vecs, uint32_t n_global, uint32_t dim_pad, int metric, uint32_t cap,
                        const uint32_t* local_to_global, uint32_t n_local, uint64_t seed, artea_refine** out) {
    clear_error();
    return refine_make(d_vecs, n_global, dim_pad, metric, cap, local_to_global, n_local, seed, out);
}

int artea_index_compact_graph(artea_index* idx, uint32_t max_level, uint32_t ul_max, uint32_t bl_max,
                              const uint32_t* highest, const uint64_t* const* row_off, const uint32_t* const* row_ids,
                              const uint32_t* order, int metric) {
    clear_error();
    ARTEA_CHECK_ARG(idx, "artea_index_compact_graph: null index");
    ARTEA_CHECK_ARG(max_level < 64, "artea_index_compact_graph: max_level must be < 64");
    ARTEA_CHECK_ARG(metric >= 0 && metric <= 3, "artea_index_compact_graph: metric must be 0..3");
    const uint32_t n = idx->n;
    ARTEA_CHECK_ARG(n == 0 || (highest && row_off && row_ids), "artea_index_compact_graph: null arrays");
    for (uint32_t v = 0; v < n; ++v)
        if (highest[v] != kInvalidLevel && highest[v] > max_level) {
            set_error("artea_index_compact_graph: vid " + std::to_string(v) + " highest level " +
                      std::to_string(highest[v]) + " > max level " + std::to_string(max_level));
            return ARTEA_EINVAL;
        }
    if (order) {
        std::vector<uint8_t> seen(n, 0);
        for (uint32_t i = 0; i < n; ++i) {
            ARTEA_CHECK_ARG(order[i] < n && !seen[order[i]], "artea_index_compact_graph: order is not a permutation");
            seen[order[i]] = 1;
        }
    }
    for (uint32_t l = 0; l <= max_level && n; ++l) {
        ARTEA_CHECK_ARG(row_off[l] && (row_off[l][n] == 0 || row_ids[l]), "artea_index_compact_graph: null level rows");
        for (uint32_t v = 0; v < n; ++v)
            ARTEA_CHECK_ARG(row_off[l][v + 1] >= row_off[l][v], "artea_index_compact_graph: row offsets decrease");
    }
    if (n == 0) {
        const uint64_t zero = 0;
        const uint64_t* p0 = &zero;
        return artea_index_set_graph_csr(idx, 0, ul_max, bl_max, kNoVertex, nullptr, &p0, nullptr, 0);
    }
    return compact_and_attach(idx, highest, order, row_off, row_ids, ul_max, bl_max, metric);
}

int artea_refine_destroy(artea_refine* r) {
    if (!r) return ARTEA_OK;
    refine_free(r);
    delete r;
    return ARTEA_OK;
}

int artea_refine_upload_rows(artea_refine* r, const uint32_t* counts, const uint32_t* entries) {
    clear_error();
    ARTEA_CHECK_ARG(r && counts, "artea_refine_upload_rows: null argument");
    Ctx& c = r->c;
    // Every entry is checked on the host before anything reaches the device:
    // the join and the reverse pass index the vectors and the layer map by
    // these ids.
    uint64_t total = 0;
    for (uint32_t i = 0; i < c.n; ++i) {
        ARTEA_CHECK_ARG(counts[i] <= c.cap, "artea_refine_upload_rows: row longer than cap");
        total += counts[i];
    }
    ARTEA_CHECK_ARG(total == 0 || entries, "artea_refine_upload_rows: null entries");
    for (uint64_t pos = 0; pos < total; ++pos) {
        if ((entries[2 * pos] & kMask) >= c.n_global) {
            set_error("artea_refine_upload_rows: neighbour id " + std::to_string(entries[2 * pos] & kMask) +
                      " is not below the vector count " + std::to_string(c.n_global));
            return ARTEA_EINVAL;
        }
    }
    std::vector<uint2> dense(std::max<size_t>(size_t(c.n) * c.cap, 1), make_uint2(0xFFFFFFFFu, 0));
    std::vector<uint32_t> cnt(c.n);
    size_t pos = 0;
    for (uint32_t i = 0; i < c.n; ++i) {
        cnt[i] = counts[i];
        for (uint32_t k = 0; k < counts[i]; ++k, ++pos)
            dense[size_t(i) * c.cap + k] = make_uint2(entries[2 * pos], entries[2 * pos + 1]);
    }
    ARTEA_HIP_TRY(hipMemcpyAsync(c.rows, dense.data(), dense.size() * sizeof(uint2), hipMemcpyHostToDevice, r->s));
    ARTEA_HIP_TRY(hipMemcpyAsync(c.cnt, cnt.data(), size_t(c.n) * 4, hipMemcpyHostToDevice, r->s));
    ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return ARTEA_OK;
}

int artea_refine_download_rows(artea_refine* r, uint32_t* counts, uint32_t* entries) {
    clear_error();
    ARTEA_CHECK_ARG(r && counts, "artea_refine_download_rows: null argument");
    Ctx& c = r->c;
    std::vector<uint2> dense(std::max<size_t>(size_t(c.n) * c.cap, 1));
    ARTEA_HIP_TRY(hipMemcpyAsync(dense.data(), c.rows, size_t(c.n) * c.cap * sizeof(uint2), hipMemcpyDeviceToHost, r->s));
    ARTEA_HIP_TRY(hipMemcpyAsync(counts, c.cnt, size_t(c.n) * 4, hipMemcpyDeviceToHost, r->s));
    ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    if (entries) {
        size_t pos = 0;
        for (uint32_t i = 0; i < c.n; ++i)
            for (uint32_t k = 0; k < counts[i]; ++k, ++pos) {
                entries[2 * pos] = dense[size_t(i) * c.cap + k].x;
                entries[2 * pos + 1] = dense[size_t(i) * c.cap + k].y;
            }
    }
    return ARTEA_OK;
}

int artea_refine_prefill(artea_refine* r, uint32_t gen_size, uint32_t start, uint32_t end, uint32_t row_begin,
                         uint32_t row_end, uint64_t stream) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_prefill: null handle");
    int rc = do_prefill(r, gen_size, start, end, row_begin, row_end, stream);
    if (rc == ARTEA_OK) ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return rc;
}

int artea_refine_triangle(artea_refine* r, float scale, float shift) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_triangle: null handle");
    int rc = do_triangle(r, scale, shift);
    if (rc == ARTEA_OK) ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return rc;
}

int artea_refine_triangle_range(artea_refine* r, float scale, float shift, uint32_t row_begin, uint32_t row_end) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_triangle_range: null handle");
    ARTEA_CHECK_ARG(row_begin <= row_end, "artea_refine_triangle_range: row_begin > row_end");
    int rc = do_triangle(r, scale, shift, row_begin, row_end);
    if (rc == ARTEA_OK) ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return rc;
}

int artea_refine_reverse_range(artea_refine* r, uint32_t row_begin, uint32_t row_end) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_reverse_range: null handle");
    ARTEA_CHECK_ARG(row_begin <= row_end, "artea_refine_reverse_range: row_begin > row_end");
    int rc = do_reverse(r, row_begin, row_end);
    if (rc == ARTEA_OK) ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return rc;
}

int artea_refine_truncate(artea_refine* r) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_truncate: null handle");
    return ARTEA_OK;  // rows never exceed cap on the device
}

int artea_refine_reverse(artea_refine* r) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_reverse: null handle");
    int rc = do_reverse(r);
    if (rc == ARTEA_OK) ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return rc;
}

int artea_refine_schedule(artea_refine* r, uint32_t build_loops, uint32_t triu_iters, float scale, float shift) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_schedule: null handle");
    int rc = do_schedule(r, build_loops, triu_iters, scale, shift);
    if (rc == ARTEA_OK) ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return rc;
}

int artea_refine_set_join(artea_refine* r, int precision) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_set_join: null handle");
    ARTEA_CHECK_ARG(precision == ARTEA_JOIN_FP32 || precision == ARTEA_JOIN_BF16,
                    "artea_refine_set_join: precision must be ARTEA_JOIN_FP32 or ARTEA_JOIN_BF16");
    r->join = precision;
    return ARTEA_OK;
}

int artea_refine_set_recheck_queue(artea_refine* r, uint32_t entries) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_set_recheck_queue: null handle");
    ARTEA_CHECK_ARG(entries <= kListCap, "artea_refine_set_recheck_queue: at most 1024 entries");
    r->c.list_cap = entries;
    return ARTEA_OK;
}

int artea_refine_last_stats(artea_refine* r, uint64_t* logs, uint64_t* exact_pairs, uint64_t* approx_pairs) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_last_stats: null handle");
    unsigned long long h[4];
    ARTEA_HIP_TRY(hipMemcpy(h, r->c.stats, sizeof h, hipMemcpyDeviceToHost));
    if (logs) *logs = h[0];
    if (exact_pairs) *exact_pairs = h[1];
    if (approx_pairs) *approx_pairs = h[2];
    ARTEA_HIP_TRY(hipMemset(r->c.stats, 0, sizeof h));
    return ARTEA_OK;
}

int artea_refine_routing_loops(artea_refine* r, uint32_t loops, uint32_t topk, uint32_t queue, float scale_coeffs) {
    clear_error();
    int rc = check_routing(r, topk, queue, scale_coeffs, "artea_refine_routing_loops");
    if (rc) return rc;
    return do_routing_loops(r, loops, topk, queue, scale_coeffs);
}

int artea_refine_prune(artea_refine* r, float scale_coeffs) {
    return artea_refine_prune_range(r, scale_coeffs, 0, r ? r->c.n : 0);
}

int artea_refine_prune_range(artea_refine* r, float scale_coeffs, uint32_t row_begin, uint32_t row_end) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_prune: null handle");
    ARTEA_CHECK_ARG(scale_coeffs > 0.0f, "artea_refine_prune: scale must be positive");
    ARTEA_HIP_TRY(prune_rows(r->c, 1.0f / scale_coeffs, row_begin, row_end, r->s, r->dev_cus));
    ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    return ARTEA_OK;
}

int artea_refine_routing_stats(artea_refine* r, uint64_t* distances, uint64_t* expansions, uint64_t* unresolved) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_refine_routing_stats: null handle");
    if (distances) *distances = r->route_stats[0];
    if (expansions) *expansions = r->route_stats[1];
    if (unresolved) *unresolved = r->route_stats[2];
    r->route_stats[0] = r->route_stats[1] = r->route_stats[2] = 0;
    return ARTEA_OK;
}

int artea_build_full(artea_index* idx, uint32_t ul_max, uint32_t bl_max, uint32_t max_level, uint32_t build_loops,
                     uint32_t triu_iters, float prefill_ratio, float scale_coeffs, float shifted_coeffs,
                     float l0_min_distance, uint64_t seed, int join_precision, float level_fraction,
                     double* seconds_per_level) {
    return artea_build_full_routing(idx, ul_max, bl_max, max_level, build_loops, triu_iters, prefill_ratio,
                                    scale_coeffs, shifted_coeffs, l0_min_distance, seed, join_precision,
                                    level_fraction, 0, 64, 96, seconds_per_level);
}

int artea_build_full_routing(artea_index* idx, uint32_t ul_max, uint32_t bl_max, uint32_t max_level,
                             uint32_t build_loops, uint32_t triu_iters, float prefill_ratio, float scale_coeffs,
                             float shifted_coeffs, float l0_min_distance, uint64_t seed, int join_precision,
                             float level_fraction, uint32_t routing_loops, uint32_t routing_topk,
                             uint32_t routing_queue, double* seconds_per_level) {
    clear_error();
    ARTEA_CHECK_ARG(idx, "artea_build_full: null index");
    if (routing_loops > 0) {
        ARTEA_CHECK_ARG(routing_topk >= 1 && routing_topk <= 128, "artea_build_full: routing topk must be in [1, 128]");
        ARTEA_CHECK_ARG(routing_queue >= routing_topk && routing_queue <= 256,
                        "artea_build_full: routing queue must be in [topk, 256]");
    }
    ARTEA_CHECK_ARG(join_precision == ARTEA_JOIN_FP32 || join_precision == ARTEA_JOIN_BF16,
                    "artea_build_full: join_precision must be ARTEA_JOIN_FP32 or ARTEA_JOIN_BF16");
    ARTEA_CHECK_ARG(max_level < 64, "artea_build_full: max_level must be < 64");
    ARTEA_CHECK_ARG(level_fraction > 0.0f && level_fraction <= 1.0f, "artea_build_full: level_fraction must be in (0, 1]");
    ARTEA_CHECK_ARG(uint32_t(float(bl_max) * 1.5f) <= kMaxCap && uint32_t(float(ul_max) * 1.5f) <= kMaxCap,
                    "artea_build_full: refine caps (1.5 x max) must be <= 128");
    const uint32_t n = idx->n;
    // Build metric: L2 for L2SQR search (build_infra_dispatch,
    // framework/type_context/infra_dispatcher.hpp:224-230).
    const int bmetric = idx->metric == ARTEA_L2SQR ? ARTEA_L2 : idx->metric;
    if (seconds_per_level)
        for (uint32_t l = 0; l <= max_level; ++l) seconds_per_level[l] = 0.0;
    // Source levels, then the compactor's trim: the first bucket from the top
    // holding >= 32 vertices becomes the top and higher vertices are demoted
    // to it (hierarchical_graph_compactor.hpp:118-252).  Levels above the
    // new top are never refined: the compactor would drop their rows.
    std::vector<uint32_t> src_h(n);
    for (uint32_t v = 0; v < n; ++v) src_h[v] = build_level_of(seed, v, max_level, level_fraction);
    uint32_t src_top = 0;
    for (uint32_t v = 0; v < n; ++v) src_top = std::max(src_top, src_h[v]);
    std::vector<uint64_t> bucket_n(src_top + 1, 0);
    for (uint32_t v = 0; v < n; ++v) bucket_n[src_h[v]]++;
    uint32_t top = 0;
    for (uint32_t h = src_top;; --h) {
        if (bucket_n[h] >= 32) {
            top = h;
            break;
        }
        if (h == 0) break;
    }
    // Rows of every refined level, indexed by vid (members only).
    std::vector<std::vector<uint64_t>> roff(top + 1);
    std::vector<std::unique_ptr<uint32_t[]>> rids(top + 1);
    for (uint32_t l = 0; l <= top; ++l) {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<uint32_t> members;
        for (uint32_t v = 0; v < n; ++v)
            if (src_h[v] >= l) members.push_back(v);
        const uint32_t nl = uint32_t(members.size());
        const bool dense = nl == n;
        const uint32_t slot_cap = l == 0 ? bl_max : ul_max;
        const uint32_t cap = uint32_t(float(slot_cap) * 1.5f);
        artea_refine* r = nullptr;
        int rc = refine_make(idx->d_vecs, n, idx->dim_pad, bmetric, cap, dense ? nullptr : members.data(), nl, seed,
                             &r);
        if (rc) return rc;
        std::unique_ptr<artea_refine, int (*)(artea_refine*)> guard(r, artea_refine_destroy);
        r->join = join_precision;
        const uint32_t gen = uint32_t(float(cap) * prefill_ratio);
        if (gen > 0 && nl > 0) {
            rc = do_prefill(r, gen, 0, nl, 0, nl, l);
            if (rc) return rc;
        }
        const float shift = l == 0 ? shifted_coeffs * l0_min_distance : 0.0f;
        rc = do_schedule(r, build_loops, triu_iters, scale_coeffs, shift);
        if (rc) return rc;
        rc = do_routing_loops(r, routing_loops, routing_topk, routing_queue, scale_coeffs);
        if (rc) return rc;
        // writeback_layer_from_single_layer_graph: first slot_cap entries.
        rc = fetch_level(r->c.rows, r->c.cnt, nullptr, members, n, cap, slot_cap, r->s, roff[l], rids[l]);
        if (rc) return rc;
        if (seconds_per_level)
            seconds_per_level[l] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    std::vector<const uint64_t*> offp(top + 1);
    std::vector<const uint32_t*> idp(top + 1);
    for (uint32_t l = 0; l <= top; ++l) {
        offp[l] = roff[l].data();
        idp[l] = rids[l].get();
    }
    return compact_and_attach(idx, src_h.data(), nullptr, offp.data(), idp.data(), ul_max, bl_max, bmetric);
}

}  // extern "C"

extern "C" int artea_build_full_rnet(artea_index* idx, const artea_rnet_params* p, uint32_t build_loops,
                                     uint32_t triu_iters, float prefill_ratio, float shifted_coeffs, int join_precision,
                                     uint32_t routing_loops, uint32_t routing_topk, uint32_t routing_queue,
                                     double* seconds) {
    clear_error();
    ARTEA_CHECK_ARG(idx && p, "artea_build_full_rnet: null argument");
    ARTEA_CHECK_ARG(join_precision == ARTEA_JOIN_FP32 || join_precision == ARTEA_JOIN_BF16,
                    "artea_build_full_rnet: join_precision must be ARTEA_JOIN_FP32 or ARTEA_JOIN_BF16");
    ARTEA_CHECK_ARG(uint32_t(float(p->bl_max_nbr_size) * 1.5f) <= kMaxCap &&
                        uint32_t(float(p->ul_max_nbr_size) * 1.5f) <= kMaxCap,
                    "artea_build_full_rnet: refine caps (1.5 x max) must be <= 128");
    if (routing_loops > 0) {
        ARTEA_CHECK_ARG(routing_topk >= 1 && routing_topk <= 128,
                        "artea_build_full_rnet: routing topk must be in [1, 128]");
        ARTEA_CHECK_ARG(routing_queue >= routing_topk && routing_queue <= 256,
                        "artea_build_full_rnet: routing queue must be in [topk, 256]");
    }
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = idx->n;
    const int bmetric = idx->metric == ARTEA_L2SQR ? ARTEA_L2 : idx->metric;
    artea_rnet* R = nullptr;
    int rc = artea::rnet_run(idx->d_vecs, n, idx->dim_pad, bmetric, *p, idx->stream, &R);
    if (rc) return rc;
    std::unique_ptr<artea_rnet, void (*)(artea_rnet*)> rguard(R, artea::rnet_free);
    if (seconds) seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const uint32_t L = R->L;
    const uint32_t top = R->top == kInvalidLevel ? 0u : R->top;
    std::vector<std::vector<uint64_t>> roff(L + 1);
    std::vector<std::unique_ptr<uint32_t[]>> rids(L + 1);
    for (uint32_t l = 0; l <= L; ++l) {
        std::vector<uint32_t> members;
        for (uint32_t v = 0; v < n; ++v)
            if (R->highest[v] != kInvalidLevel && R->highest[v] >= l) members.push_back(v);
        const uint32_t nl = uint32_t(members.size());
        const uint32_t slot_cap = l == 0 ? p->bl_max_nbr_size : p->ul_max_nbr_size;
        roff[l].assign(size_t(n) + 1, 0);
        if (l <= top && nl >= 32) {
            const uint32_t cap = uint32_t(float(slot_cap) * 1.5f);
            artea_refine* r = nullptr;
            rc = refine_make(idx->d_vecs, n, idx->dim_pad, bmetric, cap, nl == n ? nullptr : members.data(), nl,
                             p->seed, &r);
            if (rc) return rc;
            std::unique_ptr<artea_refine, int (*)(artea_refine*)> guard(r, artea_refine_destroy);
            r->join = join_precision;
            hipLaunchKernelGGL(import_rows_kernel, dim3(grid_for(r, 4, 8)), dim3(256), 0, r->s, r->c, R->rows[l],
                               R->cnt[l], slot_cap);
            ARTEA_HIP_TRY(hipGetLastError());
            if (l == 0) {
                const uint32_t gen = uint32_t(float(cap) * prefill_ratio);
                if (gen > 0) {
                    rc = do_prefill(r, gen, 0, nl, 0, nl, 0);
                    if (rc) return rc;
                }
            }
            const float shift = l == 0 ? shifted_coeffs * p->l0_min_distance : 0.0f;
            rc = do_schedule(r, build_loops, triu_iters, p->scale_coeffs, shift);
            if (rc) return rc;
            rc = do_routing_loops(r, routing_loops, routing_topk, routing_queue, p->scale_coeffs);
            if (rc) return rc;
            rc = fetch_level(r->c.rows, r->c.cnt, nullptr, members, n, cap, slot_cap, r->s, roff[l], rids[l]);
            if (rc) return rc;
        } else if (R->rows[l] && nl > 0) {
            // not refined: the r-net rows of the members as they are (by vid)
            rc = fetch_level(R->rows[l], R->cnt[l], members.data(), members, n, slot_cap, slot_cap, R->s, roff[l],
                             rids[l]);
            if (rc) return rc;
        } else {
            rids[l].reset(new uint32_t[1]);
        }
    }
    std::vector<const uint64_t*> offp(L + 1);
    std::vector<const uint32_t*> idp(L + 1);
    for (uint32_t l = 0; l <= L; ++l) {
        offp[l] = roff[l].data();
        idp[l] = rids[l].get();
    }
    rc = compact_and_attach(idx, R->highest.data(), nullptr, offp.data(), idp.data(), p->ul_max_nbr_size,
                            p->bl_max_nbr_size, bmetric);
    if (seconds) seconds[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

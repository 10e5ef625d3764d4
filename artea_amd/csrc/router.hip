// Hierarchical graph router on MI355X: one wavefront per query.
//
// Semantics follow HierarchicalGraphRouter::_query
// (router/hierarchical_graph_router.hpp:195-248): greedy best-improvement
// walks on levels top..1 (router/detail/beam_loop.hpp:129-166), then the L0
// beam search with StdCandidateQueue (router/data_structures/
// std_candidate_queue.hpp:358-513) and results by (distance, vid).
//
// Device design:
//  * The candidate list W is held in registers: K blocks of 64 (distance,
//    vid|explored) pairs, entry i in block i/64, lane i%64, sorted by
//    (distance, vid).  Positions [0, L) are the reference's top-L max-heap T;
//    positions >= L hold evicted, still-unexplored entries whose distance
//    equals the current lower bound (the only stale members of the
//    reference's unexplored heap U that can still be popped: U's top is
//    popped only while its distance is <= lb, and lb never grows once T is
//    full).  Popping = first unexplored entry of W.
//  * try_push's distance-only rejection (`|T| >= L && d >= lb`) is applied to
//    the candidates of an expansion in CSR order; only candidates below the
//    current bound are visited serially (ballot + find-first-set).
//  * The visited table is a lossy direct-mapped cache of vids in LDS.  It
//    never reports a vertex that was not inserted, and a vertex it forgets
//    is recomputed: in the beam the recomputed candidate is either still in
//    W (dropped as a duplicate key) or has distance >= lb with T full (the
//    rule rejects it), and in a greedy walk a revisited vertex can never be
//    strictly closer than the current best.  Results therefore equal the
//    reference's for every cache size; only the distance count can grow.
//  * Distances: quads of lanes per candidate row (16 rows per pass), see
//    device.hpp quad_distance.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <memory>
#include <string>

#include "common.hpp"
#include "device.hpp"

using namespace artea;
using namespace artea::dev;

namespace {

constexpr uint32_t kExplored = 0x80000000u;
constexpr uint32_t kVidMask = 0x7FFFFFFFu;
constexpr int kWavesPerBlock = 4;
constexpr uint32_t kVisBytes = 8192;  // lossy visited cache per wave: 8192 x 8-bit or 4096 x 16-bit tags

struct RouterArgs {
    const float* vecs;
    uint32_t n, D, C;
    int metric;
    const uint32_t* ell0;
    uint32_t stride0;
    const uint2* vinfo;
    const uint32_t* ellU;
    uint32_t strideU;
    uint32_t top, entry;
    const float* queries;
    uint32_t nq;
    uint32_t topk, queue;                         // fixed per launch when arrays are null
    const uint32_t* topk_arr;                     // per query (batched router)
    const uint32_t* queue_arr;
    uint32_t out_stride;
    uint32_t flags;
    const float* thresholds;                      // device, top+1 entries
    uint32_t* out_vids;
    float* out_dists;
    unsigned long long* stats;                    // [0] ndc, [1] expansions, [2] tail overflow
    uint32_t hmask;                               // 2^B - 1, B = max(13, ceil(log2 n))
    uint32_t vis_wide;                            // 1: 4096 x 16-bit tags (B > 20), 0: 8192 x 8-bit
    uint32_t* work_counter;                       // persistent mode when non-null
};

// Visited-cache key: h = v * odd mod 2^B is a bijection of [0, 2^B), so the
// pair (slot = low bits of h, tag = 1 + the remaining bits) identifies v
// exactly; tag 0 marks an empty slot.  With n <= 2^20 the 8 KiB hold 8192
// slots of 8-bit tags, otherwise 4096 slots of 16-bit tags (n <= 2^28).
struct VisKey {
    uint32_t slot, tag;
};
__device__ __forceinline__ VisKey vis_key(uint32_t v, uint32_t hmask, uint32_t wide) {
    const uint32_t h = (v * 0x9E3779B1u) & hmask;
    return wide ? VisKey{h & 4095u, (h >> 12) + 1} : VisKey{h & 8191u, (h >> 13) + 1};
}
__device__ __forceinline__ uint32_t vis_get(const unsigned char* vis, VisKey k, uint32_t wide) {
    return wide ? uint32_t(reinterpret_cast<const uint16_t*>(vis)[k.slot]) : uint32_t(vis[k.slot]);
}
__device__ __forceinline__ void vis_put(unsigned char* vis, VisKey k, uint32_t wide) {
    if (wide) reinterpret_cast<uint16_t*>(vis)[k.slot] = uint16_t(k.tag);
    else vis[k.slot] = uint8_t(k.tag);
}

struct WaveLds {
    float* q;         // D floats
    unsigned char* vis;  // kVisBytes
    uint32_t* cand;   // 64
    float* cdist;     // 64
};

// Collect the unvisited, valid ids of a 64-id chunk in row order into
// cand[0..cnt), inserting them into the visited cache; returns cnt.
__device__ __forceinline__ uint32_t collect(const WaveLds& w, uint32_t id, uint32_t n, uint32_t hmask,
                                            uint32_t wide) {
    const bool valid = id < n;
    const VisKey key = vis_key(id, hmask, wide);
    const bool fresh = valid && vis_get(w.vis, key, wide) != key.tag;
    wave_sync();
    if (fresh) vis_put(w.vis, key, wide);
    const uint64_t m = ballot(fresh);
    if (fresh) w.cand[lanes_below(m)] = id;
    wave_sync();
    return uint32_t(__popcll(m));
}

// Distances of cand[0..cnt) into cdist[0..cnt).
template <int CT>
__device__ __forceinline__ void distances(const RouterArgs& a, const WaveLds& w, uint32_t cnt, float qn2) {
    const uint32_t lane = lane_id();
    const uint32_t sub = lane & 3u;
    for (uint32_t g = 0; g < cnt; g += 16) {
        const uint32_t r = g + (lane >> 2);
        float d = 0.f;
        if (r < cnt) {
            const uint32_t v = w.cand[r];
            d = quad_distance<CT>(w.q, a.vecs + size_t(v) * a.D, int(a.C), a.metric, sub, qn2);
        }
        if (r < cnt && sub == 0) w.cdist[r] = d;
    }
    wave_sync();
}

__device__ __forceinline__ const uint32_t* upper_row(const RouterArgs& a, uint32_t v, uint32_t level) {
    const uint2 vi = a.vinfo[v];
    if (vi.x == 0xFFFFFFFFu || level > vi.x || level == 0) return nullptr;
    return a.ellU + size_t(vi.y + (vi.x - level)) * a.strideU;
}

template <int K>
struct CandList {
    float d[K];
    uint32_t v[K];
    uint32_t n;   // entries in W
    uint32_t L;   // capacity of T
    float lb;     // distance of W[L-1] when n >= L, else +inf
    uint32_t overflow;

    __device__ __forceinline__ void init(uint32_t cap) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            d[k] = __int_as_float(0x7F800000);
            v[k] = 0xFFFFFFFFu;
        }
        n = 0;
        L = cap;
        lb = __int_as_float(0x7F800000);
        overflow = 0;
    }
    __device__ __forceinline__ bool full() const { return n >= L; }

    __device__ __forceinline__ float get_d(uint32_t idx) const {
        float r = 0.f;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (idx / 64 == uint32_t(k)) r = readlane_f(d[k], int(idx % 64));
        return r;
    }

    // try_push with the reference's rule already decided by the caller;
    // inserts (dist, vid) unless the same key is present.
    __device__ __forceinline__ void insert(float dist, uint32_t vid) {
        const uint32_t lane = lane_id();
        uint32_t p = 0;
        bool dup = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t vk = v[k] & kVidMask;
            p += uint32_t(__popcll(ballot(key_less(d[k], vk, dist, vid) && v[k] != 0xFFFFFFFFu)));
            dup |= ballot(d[k] == dist && vk == vid && v[k] != 0xFFFFFFFFu) != 0;
        }
        if (dup) return;
        if (n == uint32_t(K) * 64) overflow += 1;  // last entry falls off
        // Shift [p, n) right by one, high blocks first.
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            float pd = __shfl_up(d[k], 1);
            uint32_t pv = __shfl_up(v[k], 1);
            if (k > 0) {
                const float cd = readlane_f(d[k - 1], 63);
                const uint32_t cv = readlane_u(v[k - 1], 63);
                if (lane == 0) {
                    pd = cd;
                    pv = cv;
                }
            }
            const uint32_t idx = uint32_t(k) * 64 + lane;
            if (idx > p) {
                d[k] = pd;
                v[k] = pv;
            } else if (idx == p) {
                d[k] = dist;
                v[k] = vid;
            }
        }
        n = min(n + 1, uint32_t(K) * 64);
        if (n > L) {
            // W[L] was T's maximum: it stays only while tied with the new bound.
            const float nlb = get_d(L - 1);
            uint32_t keep = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t idx = uint32_t(k) * 64 + lane;
                keep += uint32_t(__popcll(ballot(idx >= L && idx < n && d[k] == nlb)));
            }
            const uint32_t nn = L + keep;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t idx = uint32_t(k) * 64 + lane;
                if (idx >= nn) {
                    d[k] = __int_as_float(0x7F800000);
                    v[k] = 0xFFFFFFFFu;
                }
            }
            n = nn;
        }
        lb = n >= L ? get_d(L - 1) : __int_as_float(0x7F800000);
    }

    // pop_best_unexplored_entry: first unexplored entry; marks it explored.
    // Every register block is indexed with compile-time constants only
    // (a data-dependent block index would move the list to scratch memory).
    __device__ __forceinline__ bool pop(uint32_t* vid) {
        const uint32_t lane = lane_id();
        int kk = -1, l = 0;
        uint32_t found = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint64_t b = ballot(v[k] != 0xFFFFFFFFu && (v[k] & kExplored) == 0);
            if (kk < 0 && b) {
                kk = k;
                l = __ffsll((unsigned long long)b) - 1;
                found = readlane_u(v[k], l);
            }
        }
        if (kk < 0) return false;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k == kk && int(lane) == l) v[k] |= kExplored;
        *vid = found & kVidMask;
        return true;
    }

    // The entry the next pop would return if no insertion precedes it
    // (first unexplored entry), or 0xFFFFFFFF.
    __device__ __forceinline__ uint32_t peek() const {
        uint32_t found = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint64_t b = ballot(v[k] != 0xFFFFFFFFu && (v[k] & kExplored) == 0);
            if (found == 0xFFFFFFFFu && b) found = readlane_u(v[k], __ffsll((unsigned long long)b) - 1) & kVidMask;
        }
        return found;
    }
};

struct QueryStats {
    uint32_t ndc, expansions, overflow;
};

template <int K, int CT>
__device__ __forceinline__ QueryStats route_one(const RouterArgs& a, const WaveLds& w, uint32_t qi) {
    const uint32_t lane = lane_id();
    // Per-query work items are clamped to what this launch was sized for:
    // topk <= row stride, topk <= L <= the router's maximum queue.
    const uint32_t topk = min(a.topk_arr ? a.topk_arr[qi] : a.topk, a.out_stride);
    const uint32_t L = max(min(a.queue_arr ? a.queue_arr[qi] : a.queue, a.queue), max(topk, 1u));
    uint32_t* ov = a.out_vids + size_t(qi) * a.out_stride;
    float* od = a.out_dists + size_t(qi) * a.out_stride;

    // Query row into LDS.
    const float* qg = a.queries + size_t(qi) * a.D;
    for (uint32_t i = lane; i < a.D; i += 64) w.q[i] = qg[i];
    for (uint32_t i = lane; i < kVisBytes / 4; i += 64) reinterpret_cast<uint32_t*>(w.vis)[i] = 0u;
    wave_sync();
    const float qn2 = a.metric == kCosine ? quad_norm2(w.q, int(a.C), lane & 3u) : 0.f;

    // Entry point distance (all lanes compute it; quad 0's value is used).
    uint32_t cur = a.entry;
    float cd = quad_distance<CT>(w.q, a.vecs + size_t(cur) * a.D, int(a.C), a.metric, lane & 3u, qn2);
    cd = readlane_f(cd, 0);
    uint32_t ndc = 1, expansions = 0;

    if (!(a.flags & ARTEA_L0_ONLY)) {
        for (uint32_t level = a.top; level >= 1; --level) {
            // Greedy walk on `level` (greedy_loop_body); seed marked visited.
            if (lane == 0) vis_put(w.vis, vis_key(cur, a.hmask, a.vis_wide), a.vis_wide);
            wave_sync();
            const bool early = (a.flags & ARTEA_EARLY_STOP) != 0;
            const float thr = early ? a.thresholds[level] : 0.f;
            while (true) {
                if (early && cd <= thr) break;
                const uint32_t* row = upper_row(a, cur, level);
                if (!row) break;
                uint32_t best_i = 0xFFFFFFFFu;
                float best_d = cd;
                uint32_t best_v = cur;
                for (uint32_t base = 0; base < a.strideU; base += 64) {
                    const uint32_t id = (base + lane < a.strideU) ? row[base + lane] : 0xFFFFFFFFu;
                    if (ballot(id != 0xFFFFFFFFu) == 0) break;
                    const uint32_t cnt = collect(w, id, a.n, a.hmask, a.vis_wide);
                    if (cnt == 0) continue;
                    distances<CT>(a, w, cnt, qn2);
                    ndc += cnt;
                    // First strict minimum in row order below the running best.
                    float dd = lane < cnt ? w.cdist[lane] : __int_as_float(0x7F800000);
                    uint32_t ii = lane;
                    for (int off = 32; off >= 1; off >>= 1) {
                        const float od2 = __shfl_xor(dd, off);
                        const uint32_t oi = __shfl_xor(ii, off);
                        if (od2 < dd || (od2 == dd && oi < ii)) {
                            dd = od2;
                            ii = oi;
                        }
                    }
                    dd = readlane_f(dd, 0);
                    ii = readlane_u(ii, 0);
                    if (dd < best_d) {
                        best_d = dd;
                        best_i = ii;
                        best_v = w.cand[ii];
                    }
                    wave_sync();
                }
                if (best_i == 0xFFFFFFFFu) break;  // local optimum
                cur = best_v;
                cd = best_d;
            }
            // visited.clear() after every level (:228-236).
            for (uint32_t i = lane; i < kVisBytes / 4; i += 64) reinterpret_cast<uint32_t*>(w.vis)[i] = 0u;
            wave_sync();
        }
    }

    // L0 beam search from the cursor (beam_search + beam_loop_body).
    CandList<K> W;
    W.init(L);
    W.insert(cd, cur);
    if (lane == 0) vis_put(w.vis, vis_key(cur, a.hmask, a.vis_wide), a.vis_wide);
    wave_sync();
    uint32_t pv;
    // Speculative read of the first 64 ids of the row the next pop will most
    // likely return, issued before this expansion's distance gathers so the
    // next expansion usually starts without a dependent row load.
    uint32_t pre_v = 0xFFFFFFFFu, pre_id = 0xFFFFFFFFu;
    while (W.pop(&pv)) {
        expansions++;
        const uint32_t* row = a.ell0 + size_t(pv) * a.stride0;
        const bool hit = pv == pre_v;
        const uint32_t first = hit ? pre_id : (lane < a.stride0 ? row[lane] : 0xFFFFFFFFu);
        pre_v = W.peek();
        pre_id = (pre_v != 0xFFFFFFFFu && lane < a.stride0) ? a.ell0[size_t(pre_v) * a.stride0 + lane] : 0xFFFFFFFFu;
        for (uint32_t base = 0; base < a.stride0; base += 64) {
            const uint32_t id = base == 0 ? first : ((base + lane < a.stride0) ? row[base + lane] : 0xFFFFFFFFu);
            if (ballot(id != 0xFFFFFFFFu) == 0) break;
            const uint32_t cnt = collect(w, id, a.n, a.hmask, a.vis_wide);
            if (cnt == 0) continue;
            distances<CT>(a, w, cnt, qn2);
            ndc += cnt;
            const float myd = lane < cnt ? w.cdist[lane] : 0.f;
            const uint32_t myv = lane < cnt ? w.cand[lane] : 0u;
            wave_sync();
            uint64_t pending = cnt == 64 ? ~0ull : ((1ull << cnt) - 1);
            while (pending) {
                const bool ok = ((pending >> lane) & 1ull) && (!W.full() || myd < W.lb);
                const uint64_t m = ballot(ok);
                if (!m) break;
                const int i = __ffsll((unsigned long long)m) - 1;
                pending &= ~((2ull << i) - 1ull);
                if (i == 63) pending = 0;
                W.insert(readlane_f(myd, i), readlane_u(myv, i));
            }
        }
    }

    // extract_results(min(topk, |T|)), padded with the invalid entry.
    const uint32_t nt = min(W.n, W.L);
    const uint32_t kk = min(topk, nt);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t idx = uint32_t(k) * 64 + lane;
        if (idx < kk) {
            ov[idx] = W.v[k] & kVidMask;
            od[idx] = W.d[k];
        }
    }
    for (uint32_t i = kk + lane; i < topk; i += 64) {
        ov[i] = kInvalidVid;
        od[i] = FLT_MAX;
    }
    return QueryStats{ndc, expansions, W.overflow};
}

__device__ __forceinline__ WaveLds wave_lds(const RouterArgs& a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t wid = threadIdx.x / 64;
    const size_t per_wave = size_t(a.D) * 4 + kVisBytes + 512;
    unsigned char* base = smem + wid * per_wave;
    WaveLds w;
    w.q = reinterpret_cast<float*>(base);
    w.vis = base + size_t(a.D) * 4;
    w.cand = reinterpret_cast<uint32_t*>(w.vis + kVisBytes);
    w.cdist = reinterpret_cast<float*>(w.cand + 64);
    return w;
}

__device__ __forceinline__ void add_stats(const RouterArgs& a, unsigned long long ndc, unsigned long long exps,
                                          unsigned long long ovf) {
    if (lane_id() == 0 && a.stats) {
        atomicAdd(a.stats + 0, ndc);
        atomicAdd(a.stats + 1, exps);
        if (ovf) atomicAdd(a.stats + 2, ovf);
    }
}

// Plain router: one wavefront per query, the grid covers the batch.
template <int K, int CT>
__global__ __launch_bounds__(kWavesPerBlock * 64) void router_kernel(RouterArgs a) {
    const WaveLds w = wave_lds(a);
    const uint32_t qi = blockIdx.x * kWavesPerBlock + threadIdx.x / 64;
    if (qi >= a.nq) return;
    const QueryStats st = route_one<K, CT>(a, w, qi);
    add_stats(a, st.ndc, st.expansions, st.overflow);
}

// Batched router: a persistent grid sized to the resident wave slots pulls
// query indices from a device work counter until the batch is drained, so
// queries of uneven work (mixed topk / queue, hop counts) balance across
// waves and no wave idles at a launch tail.
template <int K, int CT>
__global__ __launch_bounds__(kWavesPerBlock * 64) void router_persistent_kernel(RouterArgs a) {
    const WaveLds w = wave_lds(a);
    unsigned long long ndc = 0, exps = 0, ovf = 0;
    while (true) {
        uint32_t qi = 0;
        if (lane_id() == 0) qi = atomicAdd(a.work_counter, 1u);
        qi = readfirst_u(qi);
        if (qi >= a.nq) break;
        const QueryStats st = route_one<K, CT>(a, w, qi);
        ndc += st.ndc;
        exps += st.expansions;
        ovf += st.overflow;
    }
    add_stats(a, ndc, exps, ovf);
}

// Clears the launch counters with a kernel rather than a memset node, so a
// captured launch sequence replays correctly.
__global__ void reset_kernel(unsigned long long* stats, uint32_t* counter) {
    if (threadIdx.x < 4) stats[threadIdx.x] = 0;
    if (threadIdx.x == 0 && counter) *counter = 0;
}

// Empty-graph launch: every slot gets the invalid entry.
__global__ void fill_invalid_kernel(uint32_t* vids, float* dists, size_t count) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) {
        vids[i] = kInvalidVid;
        dists[i] = FLT_MAX;
    }
}

using KernelFn = void (*)(RouterArgs);

template <int K>
KernelFn pick_ct(uint32_t C, bool persistent) {
    switch (C) {
    case 6: return persistent ? router_persistent_kernel<K, 6> : router_kernel<K, 6>;
    case 7: return persistent ? router_persistent_kernel<K, 7> : router_kernel<K, 7>;
    case 8: return persistent ? router_persistent_kernel<K, 8> : router_kernel<K, 8>;
    default: return persistent ? router_persistent_kernel<K, 0> : router_kernel<K, 0>;
    }
}

constexpr uint32_t kMinTail = 16;

int pick_kernel(uint32_t L, uint32_t C, KernelFn* fn, uint32_t* K, bool persistent = false) {
    const uint32_t need = L + kMinTail;
    if (need <= 64) { *K = 1; *fn = pick_ct<1>(C, persistent); }
    else if (need <= 128) { *K = 2; *fn = pick_ct<2>(C, persistent); }
    else if (need <= 256) { *K = 4; *fn = pick_ct<4>(C, persistent); }
    else if (need <= 512) { *K = 8; *fn = pick_ct<8>(C, persistent); }
    else if (need <= 832) { *K = 13; *fn = pick_ct<13>(C, persistent); }
    else if (need <= 1024) { *K = 16; *fn = pick_ct<16>(C, persistent); }
    else {
        set_error("candidate_queue_size " + std::to_string(L) + " exceeds the device maximum of 1008");
        return ARTEA_EINVAL;
    }
    return ARTEA_OK;
}

size_t lds_bytes(uint32_t D) { return kWavesPerBlock * (size_t(D) * 4 + kVisBytes + 512); }

}  // namespace

struct artea_router {
    artea_index* idx = nullptr;
    uint32_t topk = 0, queue = 0;
    hipStream_t stream = nullptr;
    unsigned long long* d_stats = nullptr;
    float* d_thr = nullptr;
    uint32_t* d_counter = nullptr;
    bool persistent = false;
    const void* occ_fn = nullptr;  // cached persistent grid size for (kernel, LDS bytes)
    size_t occ_lds = 0;
    uint32_t occ_blocks = 0;
};

struct artea_batched_router {
    artea_router base;
    uint32_t max_topk = 0, max_queue = 0;
};

namespace {

int router_init(artea_router* r, artea_index* idx, uint32_t topk, uint32_t queue) {
    r->idx = idx;
    r->topk = topk;
    r->queue = queue;
    ARTEA_HIP_TRY(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    ARTEA_HIP_TRY(hipMalloc(&r->d_stats, 4 * sizeof(unsigned long long)));
    ARTEA_HIP_TRY(hipMalloc(&r->d_thr, 64 * sizeof(float)));
    ARTEA_HIP_TRY(hipMalloc(&r->d_counter, 64));
    return ARTEA_OK;
}

void router_free(artea_router* r) {
    if (r->d_stats) (void)hipFree(r->d_stats);
    if (r->d_thr) (void)hipFree(r->d_thr);
    if (r->d_counter) (void)hipFree(r->d_counter);
    if (r->stream) (void)hipStreamDestroy(r->stream);
}

// Common launch path of both routers.  The graph must be attached.
int launch(artea_router* r, const float* d_q, uint32_t nq, uint32_t topk, uint32_t queue, const uint32_t* d_topk,
           const uint32_t* d_queue, uint32_t out_stride, uint32_t flags, const float* thresholds, uint32_t* d_ov,
           float* d_od, hipStream_t s, bool persistent) {
    artea_index* idx = r->idx;
    ARTEA_CHECK_ARG(idx->dg.attached, "router: no graph attached to the index");
    ARTEA_CHECK_ARG((flags & ~uint32_t(ARTEA_EARLY_STOP | ARTEA_L0_ONLY)) == 0, "router: unknown query flags");
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, r->d_stats, persistent ? r->d_counter : nullptr);
    ARTEA_HIP_TRY(hipGetLastError());
    if (nq == 0) return ARTEA_OK;
    if (idx->n == 0) {
        const size_t cnt = size_t(nq) * out_stride;
        fill_invalid_kernel<<<dim3(uint32_t((cnt + 255) / 256)), dim3(256), 0, s>>>(d_ov, d_od, cnt);
        ARTEA_HIP_TRY(hipGetLastError());
        return ARTEA_OK;
    }
    ARTEA_CHECK_ARG(idx->dg.entry < idx->n, "router: graph has no valid entry point");
    const DeviceGraph& g = idx->dg;
    if (flags & ARTEA_EARLY_STOP) {
        ARTEA_CHECK_ARG(thresholds != nullptr, "router: early stop needs per-level thresholds");
        ARTEA_CHECK_ARG(g.top < 64, "router: too many levels");
        for (uint32_t h = 1; h <= g.top; ++h) {
            const float t = thresholds[h];
            if (!(t > 0.0f) || !(t <= FLT_MAX)) {
                set_error("Early-stop threshold is not representable in query distance units");
                return ARTEA_EOVERFLOW;
            }
        }
        ARTEA_HIP_TRY(hipMemcpyAsync(r->d_thr, thresholds, sizeof(float) * (g.top + 1), hipMemcpyHostToDevice, s));
    }
    RouterArgs a{};
    a.vecs = idx->d_vecs;
    a.n = idx->n;
    a.D = idx->dim_pad;
    a.C = idx->dim_pad / 16;
    a.metric = idx->metric;
    a.ell0 = g.ell0;
    a.stride0 = g.stride0;
    a.vinfo = g.vinfo;
    a.ellU = g.ellU;
    a.strideU = g.strideU;
    a.top = g.top;
    a.entry = g.entry;
    a.queries = d_q;
    a.nq = nq;
    a.topk = topk;
    a.queue = queue;
    a.topk_arr = d_topk;
    a.queue_arr = d_queue;
    a.out_stride = out_stride;
    a.flags = flags;
    a.thresholds = r->d_thr;
    a.out_vids = d_ov;
    a.out_dists = d_od;
    a.stats = r->d_stats;
    a.work_counter = nullptr;
    {
        uint32_t B = 13;
        while (B < 32 && (uint64_t(1) << B) < idx->n) ++B;
        ARTEA_CHECK_ARG(B <= 28, "router: more than 2^28 vectors");
        a.hmask = uint32_t((uint64_t(1) << B) - 1);
        a.vis_wide = B > 20 ? 1u : 0u;
    }
    KernelFn fn;
    uint32_t K;
    int rc = pick_kernel(queue, a.C, &fn, &K, persistent);
    if (rc) return rc;
    const size_t lds = lds_bytes(a.D);
    if (persistent) {
        a.work_counter = r->d_counter;
        if (r->occ_fn != reinterpret_cast<const void*>(fn) || r->occ_lds != lds) {
            int dev = 0, cus = 0, per_cu = 0;
            ARTEA_HIP_TRY(hipGetDevice(&dev));
            ARTEA_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            ARTEA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWavesPerBlock * 64, lds));
            r->occ_fn = reinterpret_cast<const void*>(fn);
            r->occ_lds = lds;
            r->occ_blocks = uint32_t(std::max(1, cus * std::max(per_cu, 1)));
        }
        uint32_t blocks = r->occ_blocks;
        blocks = std::min(blocks, (nq + kWavesPerBlock - 1) / kWavesPerBlock);
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(kWavesPerBlock * 64), lds, s, a);
    } else {
        const uint32_t blocks = (nq + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(kWavesPerBlock * 64), lds, s, a);
    }
    ARTEA_HIP_TRY(hipGetLastError());
    return ARTEA_OK;
}

int read_stats(artea_router* r, uint64_t* ndc, uint64_t* exps, uint64_t* ovf) {
    unsigned long long h[4];
    ARTEA_HIP_TRY(hipMemcpy(h, r->d_stats, sizeof h, hipMemcpyDeviceToHost));
    if (ndc) *ndc = h[0];
    if (exps) *exps = h[1];
    if (ovf) *ovf = h[2];
    return ARTEA_OK;
}

}  // namespace

extern "C" {

int artea_router_create(artea_index* idx, uint32_t topk, uint32_t queue, artea_router** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(idx && out, "artea_router_create: null argument");
    if (queue < topk) {
        set_error("candidate_queue_size (" + std::to_string(queue) + ") must be >= topk (" + std::to_string(topk) + ")");
        return ARTEA_EINVAL;
    }
    KernelFn fn;
    uint32_t K;
    int rc = pick_kernel(queue, idx->dim_pad / 16, &fn, &K);
    if (rc) return rc;
    auto r = std::make_unique<artea_router>();
    rc = router_init(r.get(), idx, topk, queue);
    if (rc) {
        router_free(r.get());
        return rc;
    }
    *out = r.release();
    return ARTEA_OK;
}

int artea_router_destroy(artea_router* r) {
    if (!r) return ARTEA_OK;
    router_free(r);
    delete r;
    return ARTEA_OK;
}

int artea_router_batch_query_device(artea_router* r, const float* d_q, uint32_t nq, uint32_t flags,
                                    const float* thresholds, uint32_t* d_ov, float* d_od, void* stream) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_router_batch_query_device: null router");
    ARTEA_CHECK_ARG(nq == 0 || (d_q && d_ov && d_od), "artea_router_batch_query_device: null buffer");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : r->stream;
    return launch(r, d_q, nq, r->topk, r->queue, nullptr, nullptr, r->topk, flags, thresholds, d_ov, d_od, s, false);
}

int artea_router_batch_query(artea_router* r, const float* q, uint32_t nq, uint32_t flags, const float* thresholds,
                             uint32_t* ov, float* od) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_router_batch_query: null router");
    ARTEA_CHECK_ARG(nq == 0 || (q && ov && od), "artea_router_batch_query: null buffer");
    const size_t qb = size_t(nq) * r->idx->dim_pad * 4, rb = size_t(nq) * r->topk;
    float* dq = nullptr;
    uint32_t* dv = nullptr;
    float* dd = nullptr;
    int rc = ARTEA_OK;
    auto cleanup = [&]() {
        if (dq) (void)hipFree(dq);
        if (dv) (void)hipFree(dv);
        if (dd) (void)hipFree(dd);
    };
    if (nq == 0) return ARTEA_OK;
    if (hipMalloc(&dq, qb) != hipSuccess || hipMalloc(&dv, rb * 4 + 4) != hipSuccess ||
        hipMalloc(&dd, rb * 4 + 4) != hipSuccess) {
        cleanup();
        set_error("artea_router_batch_query: device allocation failed");
        return ARTEA_ENOMEM;
    }
    if (hipMemcpyAsync(dq, q, qb, hipMemcpyHostToDevice, r->stream) != hipSuccess) {
        cleanup();
        set_error("artea_router_batch_query: query upload failed");
        return ARTEA_EHIP;
    }
    rc = launch(r, dq, nq, r->topk, r->queue, nullptr, nullptr, r->topk, flags, thresholds, dv, dd, r->stream, false);
    if (rc == ARTEA_OK) {
        if (hipMemcpyAsync(ov, dv, rb * 4, hipMemcpyDeviceToHost, r->stream) != hipSuccess ||
            hipMemcpyAsync(od, dd, rb * 4, hipMemcpyDeviceToHost, r->stream) != hipSuccess ||
            hipStreamSynchronize(r->stream) != hipSuccess) {
            set_error("artea_router_batch_query: result download failed");
            rc = ARTEA_EHIP;
        }
    } else {
        (void)hipStreamSynchronize(r->stream);
    }
    cleanup();
    return rc;
}

int artea_router_last_stats(artea_router* r, uint64_t* ndc, uint64_t* exps, uint64_t* ovf) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_router_last_stats: null router");
    return read_stats(r, ndc, exps, ovf);
}

int artea_batched_router_create(artea_index* idx, uint32_t max_topk, uint32_t max_queue, artea_batched_router** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(idx && out, "artea_batched_router_create: null argument");
    if (max_queue < max_topk) {
        set_error("candidate_queue_size (" + std::to_string(max_queue) + ") must be >= topk (" +
                  std::to_string(max_topk) + ")");
        return ARTEA_EINVAL;
    }
    KernelFn fn;
    uint32_t K;
    int rc = pick_kernel(max_queue, idx->dim_pad / 16, &fn, &K);
    if (rc) return rc;
    auto r = std::make_unique<artea_batched_router>();
    rc = router_init(&r->base, idx, max_topk, max_queue);
    if (rc) {
        router_free(&r->base);
        return rc;
    }
    r->max_topk = max_topk;
    r->max_queue = max_queue;
    *out = r.release();
    return ARTEA_OK;
}

int artea_batched_router_destroy(artea_batched_router* r) {
    if (!r) return ARTEA_OK;
    router_free(&r->base);
    delete r;
    return ARTEA_OK;
}

int artea_batched_router_query_device(artea_batched_router* r, const float* d_q, uint32_t nq, const uint32_t* d_topk,
                                      const uint32_t* d_queue, uint32_t flags, const float* thresholds,
                                      uint32_t* d_ov, float* d_od, void* stream) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_batched_router_query_device: null router");
    ARTEA_CHECK_ARG(nq == 0 || (d_q && d_ov && d_od && d_topk && d_queue),
                    "artea_batched_router_query_device: null buffer");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : r->base.stream;
    return launch(&r->base, d_q, nq, r->max_topk, r->max_queue, d_topk, d_queue, r->max_topk, flags, thresholds, d_ov,
                  d_od, s, true);
}

int artea_batched_router_last_stats(artea_batched_router* r, uint64_t* ndc, uint64_t* exps, uint64_t* ovf) {
    artea::clear_error();
    ARTEA_CHECK_ARG(r, "artea_batched_router_last_stats: null router");
    return read_stats(&r->base, ndc, exps, ovf);
}

}  // extern "C"

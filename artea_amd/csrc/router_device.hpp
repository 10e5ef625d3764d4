// Device code of the hierarchical graph router: one wavefront per query.
// Included by router_inst.hip, which instantiates the kernels of one
// (metric family, candidate-list blocks) pair per compilation.
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
//  * try_push's distance-only rejection (`|T| >= L && d >= lb`) prefilters
//    the 64 candidates of an expansion chunk against the current bound; the
//    survivors are decided and merged into W together (merge_chunk), with
//    the same outcome as pushing them one by one in CSR order.
//  * The visited table is a lossy two-choice cache of vids in LDS.  It
//    never reports a vertex that was not inserted, and a vertex it forgets
//    is recomputed: in the beam the recomputed candidate is either still in
//    W (dropped as a duplicate key) or has distance >= lb with T full (the
//    rule rejects it), and in a greedy walk a revisited vertex can never be
//    strictly closer than the current best.  Results therefore equal the
//    reference's for every cache size; only the distance count can grow.
//  * Distances: quads of lanes per candidate row, 32 rows per pass (every
//    row's loads issued before any reduction), see device.hpp
//    quad_distance2; from one-byte rows (whole-number data) in exact
//    integer arithmetic, 16-64 rows per pass (distances_u8).  A merge
//    moves every register block in one ds_bpermute gather, and the tail is
//    trimmed once per expansion chunk.
#pragma once

#include <cfloat>

#include "common.hpp"
#include "device.hpp"
#include "router_kernels.hpp"

using namespace artea;
using namespace artea::dev;
using namespace artea::rk;

namespace {

constexpr uint32_t kExplored = 0x80000000u;
constexpr uint32_t kVidMask = 0x7FFFFFFFu;



// Visited cache: two-choice hashing over vis_bytes(K, U8) bytes of codes.
// h1 = v * A1 and h2 = v * A2 mod 2^B (odd multipliers) are bijections of
// [0, 2^B), so a slot (h mod S) and a quotient (h div S) identify v exactly.
// Choice 1 stores quotient + 1 in its slot, choice 2 stores quotient + 1 +
// (the number of choice-1 codes), so a stored code also says which choice
// wrote it; 0 marks an empty slot.  Modes: 0 (n <= 2^20) S = vis_bytes(K, U8)
// slots of 8-bit codes; 1 (n <= 2^27) S = vis_bytes(K, U8) / 2 slots of 16-bit
// codes; 2 (n <= 2^28) as 1 with choice 1 only, and a vertex whose code
// exceeds 16 bits is never stored.  The choice-2 codes fill what the code
// width leaves above the choice-1 codes, so vertices with a larger choice-2
// quotient use choice 1 only.  A lookup hits only on v's own code, so the
// cache never reports a vertex that was not inserted.  An insertion takes
// choice 1 when empty, else choice 2 when empty, else overwrites choice 1.
struct VisKey {
    uint32_t s1, t1, s2, t2;  // t2 == 0: no second choice
};
template <int K, bool U8>
__device__ __forceinline__ VisKey vis_key(uint32_t v, uint32_t hmask, uint32_t mode) {
    constexpr uint32_t S8 = vis_bytes(K, U8), S16 = vis_bytes(K, U8) / 2;
    constexpr uint32_t C8 = ((1u << 20) - 1) / S8 + 1, C16 = ((1u << 27) - 1) / S16 + 1;  // choice-1 codes
    static_assert(C8 < 255 && C16 < 65535, "visited cache too small for its code width");
    const uint32_t h1 = (v * 0x9E3779B1u) & hmask, h2 = (v * 0x85EBCA77u) & hmask;
    if (mode == 0) {
        const uint32_t q1 = h1 / S8, q2 = h2 / S8;
        return VisKey{h1 - q1 * S8, q1 + 1, h2 - q2 * S8, q2 + C8 + 1 <= 255u ? q2 + C8 + 1 : 0u};
    }
    const uint32_t q1 = h1 / S16, q2 = h2 / S16;
    return VisKey{h1 - q1 * S16, q1 + 1, h2 - q2 * S16, (mode == 1 && q2 + C16 + 1 <= 65535u) ? q2 + C16 + 1 : 0u};
}
__device__ __forceinline__ uint32_t vis_load(const unsigned char* vis, uint32_t slot, uint32_t mode) {
    return mode ? uint32_t(reinterpret_cast<const uint16_t*>(vis)[slot]) : uint32_t(vis[slot]);
}
__device__ __forceinline__ void vis_store(unsigned char* vis, uint32_t slot, uint32_t code, uint32_t mode) {
    if (mode) {
        if (code <= 0xFFFFu) reinterpret_cast<uint16_t*>(vis)[slot] = uint16_t(code);
    } else {
        vis[slot] = uint8_t(code);
    }
}

struct WaveLds {
    float* q;         // D floats
    unsigned char* vis;  // vis_bytes(K, U8)
    uint32_t* cand;   // 64 (= wv[0..64): staging between collect and a merge)
    float* cdist;     // 64 (= wd[0..64))
    float* wd;        // 64 K: candidate-list distances by position (merge_chunk)
    uint32_t* wv;     // 64 K: candidate-list vids by position
};

// Collect the unvisited, valid ids of a 64-id chunk in row order into
// cand[0..cnt), inserting them into the visited cache; returns cnt.
template <int K, bool U8>
__device__ __forceinline__ uint32_t collect(const WaveLds& w, uint32_t id, uint32_t n, uint32_t hmask,
                                            uint32_t mode) {
    const bool valid = id < n;
    const VisKey key = vis_key<K, U8>(id, hmask, mode);
    const uint32_t c1 = vis_load(w.vis, key.s1, mode), c2 = vis_load(w.vis, key.s2, mode);
    const bool fresh = valid && c1 != key.t1 && (key.t2 == 0 || c2 != key.t2);
    wave_sync();
    if (fresh) {
        const bool second = c1 != 0 && key.t2 != 0 && c2 == 0;
        vis_store(w.vis, second ? key.s2 : key.s1, second ? key.t2 : key.t1, mode);
    }
    const uint64_t m = ballot(fresh);
    if (fresh) w.cand[lanes_below(m)] = id;
    wave_sync();
    return uint32_t(__popcll(m));
}

// Distances of cand[0..cnt) into cdist[0..cnt).
// The chain a kernel of metric family MF evaluates: 0 = L2SQR (and L2, whose
// square root is taken in a pass of its own, l2_pass: selected per row inside
// the chains, the correctly rounded sqrt would be computed for every L2SQR
// row too), 1 = DOT, 2 = COSINE.  Fixing the family at compile time keeps the
// other families' register demand out of the kernel.
template <int MF>
__device__ __forceinline__ int family_metric(int) {
    return MF == 0 ? int(kL2Sqr) : (MF == 1 ? int(kDot) : int(kCosine));
}

// L2 = sqrt of the L2SQR chain, bit for bit as dist_func_t computes it.
template <int MF>
__device__ __forceinline__ void l2_pass(const RouterArgs& a, const WaveLds& w, uint32_t cnt) {
    if (MF == 0 && a.metric == kL2) {
        const uint32_t lane = lane_id();
        if (lane < cnt) w.cdist[lane] = __builtin_sqrtf(w.cdist[lane]);
        wave_sync();
    }
}

template <int CT, bool WIDE, int MF>
__device__ __forceinline__ void distances(const RouterArgs& a, const WaveLds& w, uint32_t cnt, float qn2) {
    const uint32_t lane = lane_id();
    const int metric = family_metric<MF>(a.metric);
    const uint32_t sub = lane & 3u;
    uint32_t g = 0;
    // WIDE: 32 rows per pass while more than 16 remain: rows g + lane/4 and
    // g + 16 + lane/4; a quad past the end re-reads its first row (a cache
    // hit) instead of branching around the loads.  Otherwise 16 rows per pass.
    for (; WIDE && g + 16 < cnt; g += 32) {
        const uint32_t ra = g + (lane >> 2), rb = ra + 16;
        const uint32_t va = w.cand[ra];
        const uint32_t vb = rb < cnt ? w.cand[rb] : va;
        float da, db;
        quad_distance2<CT>(w.q, a.vecs + size_t(va) * a.D, a.vecs + size_t(vb) * a.D, int(a.C), metric, sub, qn2,
                           &da, &db);
        if (sub == 0) {
            w.cdist[ra] = da;
            if (rb < cnt) w.cdist[rb] = db;
        }
    }
    for (; g < cnt; g += 16) {
        const uint32_t r = g + (lane >> 2);
        if (r < cnt) {
            const uint32_t v = w.cand[r];
            float d;
            if constexpr (CT > 0) {
                // every chunk's load issued before the chain starts
                const float* x = a.vecs + size_t(v) * a.D;
                float4 xv[CT];
#pragma unroll
                for (int c = 0; c < CT; ++c) xv[c] = *reinterpret_cast<const float4*>(x + c * 16 + sub * 4);
                d = quad_finish<CT>(w.q, xv, metric, sub, qn2);
            } else {
                d = quad_distance<CT>(w.q, a.vecs + size_t(v) * a.D, int(a.C), metric, sub, qn2);
            }
            if (sub == 0) w.cdist[r] = d;
        }
    }
    wave_sync();
}

// One-byte state of a query (device.hpp, "one-byte rows"): this lane's quad
// slice of the query as bytes and q.q, used when the index keeps one-byte
// rows and every value of the query is a whole number in [0, 255].
template <int CT>
struct U8Query {
    bool on;
    uint32_t qn;
    uint32_t qw[CT > 0 ? CT : 1];
};

template <int CT, bool U8>
__device__ __forceinline__ U8Query<CT> u8_query(const RouterArgs& a, const WaveLds& w) {
    U8Query<CT> u;
    u.on = false;
    u.qn = 0;
    if constexpr (CT > 0 && U8) {
        if (a.vecs_u8) {
            const uint32_t lane = lane_id(), sub = lane & 3u;
            bool bad = false;
            for (uint32_t i = lane; i < a.D; i += 64) bad |= !byte_valued(w.q[i]);
            u.on = ballot(bad) == 0;
            if (u.on) {
                uint32_t s = 0;
#pragma unroll
                for (int k = 0; k < CT; ++k) {
                    const float4 f = *reinterpret_cast<const float4*>(w.q + sub * 4 * CT + 4 * k);
                    u.qw[k] = uint32_t(f.x) | (uint32_t(f.y) << 8) | (uint32_t(f.z) << 16) | (uint32_t(f.w) << 24);
                    s = __builtin_amdgcn_udot4(u.qw[k], u.qw[k], s, false);
                }
                u.qn = quad_sum_u(s);
            }
        }
    }
    return u;
}

// Distances of cand[0..cnt) into cdist[0..cnt) from the one-byte rows: a
// quad per row, 16 rows per group, R groups' loads issued before any sum.
// Equal to distances() bit for bit on such queries (exact arithmetic).
template <int CT, int R>
__device__ __forceinline__ void distances_u8(const RouterArgs& a, const WaveLds& w, uint32_t cnt, const U8Query<CT>& u,
                                             int metric) {
    const uint32_t lane = lane_id(), sub = lane & 3u;
    for (uint32_t g = 0; g < cnt; g += 16 * R) {
        uint32_t x[R][CT];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (g + 16u * r >= cnt) break;
            const uint32_t row = min(g + (lane >> 2) + 16u * r, cnt - 1);  // past the end: re-read the last row
            load_u8_slice<CT>(a.vecs_u8 + size_t(w.cand[row]) * (4 * CT) + sub * CT, x[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (g + 16u * r >= cnt) break;
            uint32_t qx, xx;
            u8_dots<CT>(u.qw, x[r], &qx, &xx);
            const uint32_t row = g + (lane >> 2) + 16u * r;
            if (sub == 0 && row < cnt) w.cdist[row] = u8_finish(metric, u.qn, qx, xx);
        }
    }
    wave_sync();
}

// The query's distances to cand[0..cnt): one-byte rows when it can use them.
template <int CT, bool WIDE, int MF, int R>
__device__ __forceinline__ void any_distances(const RouterArgs& a, const WaveLds& w, uint32_t cnt, const U8Query<CT>& u,
                                              int metric, float qn2) {
    bool done = false;
    if constexpr (CT > 0) {
        if (u.on) {
            distances_u8<CT, R>(a, w, cnt, u, metric);
            done = true;
        }
    }
    if (!done) distances<CT, WIDE, MF>(a, w, cnt, qn2);
    l2_pass<MF>(a, w, cnt);
}

// Distance of one row (every lane of quad 0 returns it).
template <int CT>
__device__ __forceinline__ float row_distance(const RouterArgs& a, const WaveLds& w, const U8Query<CT>& u, uint32_t v,
                                              int metric, float qn2) {
    if constexpr (CT > 0) {
        if (u.on) {
            uint32_t x[CT], qx, xx;
            load_u8_slice<CT>(a.vecs_u8 + size_t(v) * (4 * CT) + (lane_id() & 3u) * CT, x);
            u8_dots<CT>(u.qw, x, &qx, &xx);
            return u8_finish(metric, u.qn, qx, xx);
        }
    }
    return quad_distance<CT>(w.q, a.vecs + size_t(v) * a.D, int(a.C), metric, lane_id() & 3u, qn2);
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
    uint32_t lb_block, lb_lane;  // position L-1
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
        lb_block = (cap - 1) / 64;
        lb_lane = (cap - 1) % 64;
        lb = __int_as_float(0x7F800000);
        overflow = 0;
    }
    __device__ __forceinline__ bool full() const { return n >= L; }

    // Distance of W[L-1] (block and lane of L-1 fixed per query).
    __device__ __forceinline__ float bound_d() const {
        float r = 0.f;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (lb_block == uint32_t(k)) r = readlane_f(d[k], int(lb_lane));
        return r;
    }

    __device__ __forceinline__ float get_d(uint32_t idx) const {
        float r = 0.f;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (idx / 64 == uint32_t(k)) r = readlane_f(d[k], int(idx % 64));
        return r;
    }

    // Tail trimming: W[L..n) keeps only the entries tied with the current
    // bound.  Trimming after every insertion or once after a run of
    // insertions leaves the same list (the bound never grows, and a tail
    // entry tied with the final bound was tied with every bound in between),
    // so the beam trims once per expansion chunk, and before an insertion
    // that would otherwise push an entry off the end.
    __device__ __forceinline__ void trim() {
        if (n <= L) return;
        const uint32_t lane = lane_id();
        uint32_t keep = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t idx = uint32_t(k) * 64 + lane;
            keep += uint32_t(__popcll(ballot(idx >= L && idx < n && d[k] == lb)));
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

    // try_push with the reference's rule already decided by the caller;
    // inserts (dist, vid) unless the same key is present and returns whether
    // it did.  Empty slots hold (+inf, 0xFFFFFFFF): never below a candidate
    // key, never equal to one.  Position p = #{keys below (dist, vid)}: the
    // distance ballots decide it unless some entry has the same distance;
    // only then are vids compared (ties order by vid, and an equal vid means
    // the key is already present).
    __device__ __forceinline__ bool insert(float dist, uint32_t vid) {
        const uint32_t lane = lane_id();
        uint32_t p = 0;
        uint64_t eq[K], any_eq = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            eq[k] = ballot(d[k] == dist);
            any_eq |= eq[k];
            p += uint32_t(__popcll(ballot(d[k] < dist)));
        }
        if (any_eq) {
            uint64_t dup = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (!eq[k]) continue;
                const uint32_t vk = v[k] & kVidMask;
                p += uint32_t(__popcll(eq[k] & ballot(vk < vid)));
                dup |= eq[k] & ballot(vk == vid);
            }
            if (dup) return false;
        }
        if (n == uint32_t(K) * 64) {
            trim();
            if (n == uint32_t(K) * 64) overflow += 1;  // last entry falls off
        }
        // Shift [p, n] right by one (DPP wave shift; lane 0 takes the previous
        // block's lane 63), high blocks first; blocks wholly below p or past
        // n are unchanged.  The new entry is written into lane p % 64 of
        // block p / 64.
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            if (p >= uint32_t(k + 1) * 64 || uint32_t(k) * 64 > n) continue;
            const int cd = k > 0 ? __builtin_amdgcn_readlane(__float_as_int(d[k - 1]), 63) : 0;
            const int cv = k > 0 ? __builtin_amdgcn_readlane(int(v[k - 1]), 63) : 0;
            const float pd = __int_as_float(__builtin_amdgcn_update_dpp(cd, __float_as_int(d[k]), 0x138, 0xF, 0xF, false));
            const uint32_t pv = uint32_t(__builtin_amdgcn_update_dpp(cv, int(v[k]), 0x138, 0xF, 0xF, false));
            const bool shifted = uint32_t(k) * 64 + lane > p;
            d[k] = shifted ? pd : d[k];
            v[k] = shifted ? pv : v[k];
            if (p >= uint32_t(k) * 64) {
                const bool at = lane == p - uint32_t(k) * 64;
                d[k] = at ? dist : d[k];
                v[k] = at ? vid : v[k];
            }
        }
        n = min(n + 1, uint32_t(K) * 64);
        if (n >= L) lb = bound_d();
        return true;
    }

    // One expansion chunk at once: lanes in P hold the candidates that pass
    // the distance prefilter, in CSR order.  try_push's rule for the p-th
    // candidate, |T| < L or d < lb, is the count test
    //   c_p = min(#{list entries with distance <= d_p}, L)
    //         + #{earlier accepted candidates with distance <= d_p} < L
    // (T is the L smallest keys of the list plus the accepted ones, and the
    // bound exceeds d_p exactly when fewer than L of them are <= d_p); a
    // candidate whose key is already in the list, or equals an earlier
    // accepted one, is dropped as insert() would drop it.
    //
    // The test needs no serial pass: a candidate i before p with
    // d_i <= d_p that the test rejects has c_p >= c_i >= L, so p is accepted
    // exactly when the count over ALL earlier candidates not in the list
    // keeps c_p below L.  Only a vid repeated inside the chunk (a row that
    // lists a neighbour twice) takes the serial route.  The accepted
    // candidates are merged in one gather: the accepted key of rank r lands
    // at its list position plus r, an old entry moves up by the accepted
    // keys below it.  Entries pushed past the register capacity are dropped:
    // they lie after the new bound and trim() would drop them as well,
    // unless one ties with that bound, which is counted as an overflow.
    // Returns whether anything was accepted; before the merge on_first(d,
    // vid) receives the smallest accepted key.
    template <class OnFirst>
    __device__ bool merge_chunk(float myd, uint32_t myv, uint64_t P, const WaveLds& w, OnFirst&& on_first) {
        constexpr uint32_t CAP = uint32_t(K) * 64;
        constexpr uint32_t POW2 = K <= 1 ? 64 : (K <= 2 ? 128 : (K <= 4 ? 256 : (K <= 8 ? 512 : 1024)));
        const uint32_t lane = lane_id();
        // Every candidate's place in the list at once, from an LDS mirror of
        // (distance, vid): a binary search for the entries below its
        // distance, then a scan of the entries tied with it (usually none).
#pragma unroll
        for (int k = 0; k < K; ++k) {
            w.wd[k * 64 + lane] = d[k];
            w.wv[k * 64 + lane] = v[k] & kVidMask;
        }
        wave_sync();
        uint32_t lt = 0;
#pragma unroll
        for (uint32_t st = POW2 / 2; st >= 1; st >>= 1) {
            const uint32_t q = lt + st;
            if (q <= CAP && w.wd[q - 1] < myd) lt = q;
        }
        uint32_t le = lt, mypos = lt;  // entries <= d; list keys below (d, vid)
        bool dupl = false;
        if ((P >> lane) & 1ull) {
            while (le < CAP && w.wd[le] == myd) {
                const uint32_t vj = w.wv[le];
                mypos += vj < myv ? 1u : 0u;
                dupl |= vj == myv;
                ++le;
            }
        }
        const uint64_t Q = P & ~ballot(dupl);
        // Per lane: candidates of Q before it with distance <= its own,
        // candidates of Q with a smaller key, and the lanes that repeat the
        // vid of an earlier candidate of Q (same vid, same distance).
        uint32_t cnt = 0, rank = 0;
        uint64_t rep = 0;
        for (uint64_t it = Q; it; it &= it - 1) {
            const int i = __ffsll((unsigned long long)it) - 1;
            const float di = readlane_f(myd, i);
            const uint32_t vi = readlane_u(myv, i);
            cnt += (int(lane) > i && di <= myd) ? 1u : 0u;
            rank += key_less(di, vi, myd, myv) ? 1u : 0u;
            rep |= ballot(vi == myv) & ((~0ull << i) << 1);
        }
        uint64_t A = 0;
        if (!(rep & Q)) {
            A = Q & ballot(cnt < L - min(le, L));
            // rank among the accepted: take back the rejected keys below
            for (uint64_t it = Q & ~A; it; it &= it - 1) {
                const int r = __ffsll((unsigned long long)it) - 1;
                rank -= key_less(readlane_f(myd, r), readlane_u(myv, r), myd, myv) ? 1u : 0u;
            }
        } else {
            for (uint64_t it = Q; it; it &= it - 1) {
                const int p = __ffsll((unsigned long long)it) - 1;
                const float dp = readlane_f(myd, p);
                const uint32_t vp = readlane_u(myv, p);
                const uint32_t c = min(readlane_u(le, p), L) + uint32_t(__popcll(A & ballot(myd <= dp)));
                if ((A & ballot(myv == vp)) == 0 && c < L) A |= 1ull << p;
            }
            rank = 0;
            for (uint64_t it = A; it; it &= it - 1) {
                const int a = __ffsll((unsigned long long)it) - 1;
                rank += key_less(readlane_f(myd, a), readlane_u(myv, a), myd, myv) ? 1u : 0u;
            }
        }
        if (!A) return false;
        const bool inA = (A >> lane) & 1ull;
        const uint32_t m = uint32_t(__popcll(A));
        const uint32_t f = mypos + rank;  // final position of an accepted lane
        // Smallest accepted key first: the caller's next-pop row read can
        // then overlap the merge.
        const int r0 = __ffsll((unsigned long long)ballot(inA && rank == 0)) - 1;
        on_first(readlane_f(myd, r0), readlane_u(myv, r0));
        const uint32_t fmin = readlane_u(f, r0);
        // Accepted entries into the mirror at their final positions, marked
        // by the top vid bit (mirror vids have it clear).
        wave_sync();
        if (inA && f < CAP) {
            w.wd[f] = myd;
            w.wv[f] = myv | kExplored;
        }
        wave_sync();
        // Gather, high blocks first (block k reads the old blocks k and k-1).
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            if (uint32_t(k) * 64 + 63 < fmin) continue;
            const float nd = w.wd[k * 64 + lane];
            const uint32_t nv = w.wv[k * 64 + lane];
            const bool isnew = (nv & kExplored) != 0;
            const uint32_t below = k == 0 ? 0u : uint32_t(__popcll(ballot(inA && f < uint32_t(k) * 64)));
            const uint32_t c = below + lanes_below(ballot(isnew));  // new entries before this position
            const uint32_t t = uint32_t(k) * 64 + lane - c;
            const int addr = int((t & 63u) << 2);
            float od = __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(d[k])));
            uint32_t ov = uint32_t(__builtin_amdgcn_ds_bpermute(addr, int(v[k])));
            if (k > 0) {
                const float od1 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(d[k - 1])));
                const uint32_t ov1 = uint32_t(__builtin_amdgcn_ds_bpermute(addr, int(v[k - 1])));
                const bool here = t >= uint32_t(k) * 64;
                od = here ? od : od1;
                ov = here ? ov : ov1;
            }
            d[k] = isnew ? nd : od;
            v[k] = isnew ? (nv & kVidMask) : ov;
        }
        const bool spill = n + m > CAP;
        n = min(n + m, CAP);
        if (n >= L) lb = bound_d();
        if (spill && readlane_f(d[K - 1], 63) == lb) overflow += 1;
        trim();
        return true;
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

    // peek() that also returns the entry's position in W.
    __device__ __forceinline__ uint32_t peek_pos(float* dist, uint32_t* pos) const {
        uint32_t found = 0xFFFFFFFFu, at = 0;
        float fd = __int_as_float(0x7F800000);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint64_t b = ballot(v[k] != 0xFFFFFFFFu && (v[k] & kExplored) == 0);
            if (found == 0xFFFFFFFFu && b) {
                const int l = __ffsll((unsigned long long)b) - 1;
                found = readlane_u(v[k], l) & kVidMask;
                fd = readlane_f(d[k], l);
                at = uint32_t(k) * 64 + uint32_t(l);
            }
        }
        *dist = fd;
        *pos = at;
        return found;
    }

    // Marks the entry at position pos explored (the pop of a peeked entry).
    __device__ __forceinline__ void mark_explored(uint32_t pos) {
        const uint32_t lane = lane_id();
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (pos / 64 == uint32_t(k) && lane == pos % 64) v[k] |= kExplored;
    }

    // The entry the next pop would return if no insertion precedes it
    // (first unexplored entry), or 0xFFFFFFFF.
    __device__ __forceinline__ uint32_t peek(float* dist) const {
        uint32_t found = 0xFFFFFFFFu;
        float fd = __int_as_float(0x7F800000);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint64_t b = ballot(v[k] != 0xFFFFFFFFu && (v[k] & kExplored) == 0);
            if (found == 0xFFFFFFFFu && b) {
                const int l = __ffsll((unsigned long long)b) - 1;
                found = readlane_u(v[k], l) & kVidMask;
                fd = readlane_f(d[k], l);
            }
        }
        *dist = fd;
        return found;
    }
};

struct QueryStats {
    uint32_t ndc, expansions, overflow;
};

// U8R > 0: the kernel also carries the one-byte path (used by queries that
// can), with U8R groups of 16 rows per distance pass in the beam.
template <int K, int CT, int MF, bool WIDE, int U8R>
__device__ __forceinline__ QueryStats route_one(const RouterArgs& a, const WaveLds& w, uint32_t qi) {
    const uint32_t lane = lane_id();
    const int metric = family_metric<MF>(a.metric);
    // Per-query work items are clamped to what this launch was sized for:
    // topk <= row stride, topk <= L <= the router's maximum queue.
    const uint32_t topk = readfirst_u(min(a.topk_arr ? a.topk_arr[qi] : a.topk, a.out_stride));
    const uint32_t L = readfirst_u(max(min(a.queue_arr ? a.queue_arr[qi] : a.queue, a.queue), max(topk, 1u)));
    uint32_t* ov = a.out_vids + size_t(qi) * a.out_stride;
    float* od = a.out_dists + size_t(qi) * a.out_stride;

    // Query row into LDS.
    const float* qg = a.queries + size_t(qi) * a.D;
    for (uint32_t i = lane; i < a.D; i += 64) w.q[i] = qg[i];
    for (uint32_t i = lane; i < vis_bytes(K, (U8R > 0)) / 4; i += 64) reinterpret_cast<uint32_t*>(w.vis)[i] = 0u;
    wave_sync();
    const float qn2 = MF == 2 ? quad_norm2(w.q, int(a.C), lane & 3u) : 0.f;
    const U8Query<CT> u = u8_query<CT, (U8R > 0)>(a, w);

    // Entry point distance (all lanes compute it; quad 0's value is used).
    uint32_t cur = a.entry;
    float cd = readlane_f(row_distance<CT>(a, w, u, cur, metric, qn2), 0);
    if (MF == 0 && a.metric == kL2) cd = __builtin_sqrtf(cd);
    uint32_t ndc = 1, expansions = 0;

    if (!(a.flags & ARTEA_L0_ONLY)) {
        for (uint32_t level = a.top; level >= 1; --level) {
            // Greedy walk on `level` (greedy_loop_body); seed marked visited.
            if (lane == 0) {
                const VisKey k0 = vis_key<K, (U8R > 0)>(cur, a.hmask, a.vis_mode);
                vis_store(w.vis, k0.s1, k0.t1, a.vis_mode);
            }
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
                    const uint32_t cnt = collect<K, (U8R > 0)>(w, id, a.n, a.hmask, a.vis_mode);
                    if (cnt == 0) continue;
                    any_distances<CT, false, MF, 1>(a, w, cnt, u, metric, qn2);
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
            for (uint32_t i = lane; i < vis_bytes(K, (U8R > 0)) / 4; i += 64) reinterpret_cast<uint32_t*>(w.vis)[i] = 0u;
            wave_sync();
        }
    }

    // L0 beam search from the cursor (beam_search + beam_loop_body).
    CandList<K> W;
    W.init(L);
    W.insert(cd, cur);
    if (lane == 0) {
        const VisKey k0 = vis_key<K, (U8R > 0)>(cur, a.hmask, a.vis_mode);
        vis_store(w.vis, k0.s1, k0.t1, a.vis_mode);
    }
    wave_sync();
    uint32_t pv;
    // Speculative read of the first 64 ids of the row the next pop will most
    // likely return, issued before this expansion's distance gathers so the
    // next expansion usually starts without a dependent row load.
    // An insertion whose key precedes that entry becomes the next pop (unless
    // a later one of the same expansion precedes it too), so its row read is
    // issued at once and overlaps the rest of the expansion.
    uint32_t pre_v = 0xFFFFFFFFu, pre_id = 0xFFFFFFFFu;
    float pre_d = 0.f;
    // The read is issued after this row's ids are consumed and before its
    // distance gathers, whose in-order waits then cover it; issued earlier,
    // the wait on this row's ids would also wait for it.
    auto on_first = [&](float d0, uint32_t v0) {
        if (pre_v == 0xFFFFFFFFu || key_less(d0, v0, pre_d, pre_v)) {
            pre_v = v0;
            pre_d = d0;
            pre_id = lane < a.stride0 ? a.ell0[size_t(v0) * a.stride0 + lane] : 0xFFFFFFFFu;
        }
    };
    auto predict = [&](uint32_t* pos) {
        pre_v = W.peek_pos(&pre_d, pos);
        pre_id = (pre_v != 0xFFFFFFFFu && lane < a.stride0) ? a.ell0[size_t(pre_v) * a.stride0 + lane] : 0xFFFFFFFFu;
    };
    // Rows of at most 64 ids (one chunk): chains of expansions where the
    // order allows.  After an expansion's distances, the predicted next pop e
    // (the first unexplored entry) is certainly the next pop when it lies in
    // T and no candidate held since the last merge (prefiltered) precedes it:
    // an insertion lands after e, and the tail trim only touches positions at
    // or after L.  e is then expanded at once, and the chain's candidates are
    // merged together in expansion order, which is the same sequence of
    // try_push calls (merge_chunk's count test holds for any run of
    // candidates in order; a later expansion's prefilter uses the bound
    // before the merge, which only lets more candidates reach the exact
    // test).  A chain ends when its next pop is not certain or when its
    // candidates would not fit one merge (64 lanes; the run that did not fit
    // is merged after the held ones); every link marks an unexplored entry of
    // T explored and none is added before the merge, so a chain has at most L
    // links.  The loop has one merge site: each inlined copy of merge_chunk
    // costs code size and registers (two and three sites measured slower).
    if (a.stride0 <= 64) {
        // Candidates held for the chain's merge, in order: the lanes of
        // `held` (one expansion's prefiltered lanes as they came, or, once a
        // second run joins, lanes 0 .. m - 1); `carry` holds a run that did
        // not fit, merged next.
        float accd = 0.f, crd = 0.f;
        uint32_t accv = 0u, crv = 0u;
        uint64_t held = 0, carry = 0;
        while (true) {
            if (carry) {
                accd = crd;
                accv = crv;
                held = carry;
                carry = 0;
            } else {
                if (!W.pop(&pv)) break;
                uint32_t ids = pv == pre_v ? pre_id : (lane < a.stride0 ? a.ell0[size_t(pv) * a.stride0 + lane]
                                                                        : 0xFFFFFFFFu);
                held = 0;
                while (true) {
                    expansions++;
                    const bool rows = ballot(ids != 0xFFFFFFFFu) != 0;
                    const uint32_t cnt = rows ? collect<K, (U8R > 0)>(w, ids, a.n, a.hmask, a.vis_mode) : 0u;
                    uint32_t epos;
                    predict(&epos);
                    if (cnt) {
                        any_distances<CT, WIDE, MF, (U8R > 0 ? U8R : 1)>(a, w, cnt, u, metric, qn2);
                        ndc += cnt;
                        const float myd = lane < cnt ? w.cdist[lane] : 0.f;
                        const uint32_t myv = lane < cnt ? w.cand[lane] : 0u;
                        wave_sync();
                        const uint64_t P =
                            (cnt == 64 ? ~0ull : ((1ull << cnt) - 1)) & (W.full() ? ballot(myd < W.lb) : ~0ull);
                        if (P && !held) {
                            accd = myd;
                            accv = myv;
                            held = P;
                        } else if (P) {
                            const uint32_t m0 = uint32_t(__popcll(held)), m = m0 + uint32_t(__popcll(P));
                            if (m > 64) {
                                crd = myd;
                                crv = myv;
                                carry = P;
                                break;
                            }
                            // The held run, then this one, through the
                            // staging slots into lanes 0 .. m - 1.
                            if ((held >> lane) & 1ull) {
                                w.cdist[lanes_below(held)] = accd;
                                w.cand[lanes_below(held)] = accv;
                            }
                            if ((P >> lane) & 1ull) {
                                w.cdist[m0 + lanes_below(P)] = myd;
                                w.cand[m0 + lanes_below(P)] = myv;
                            }
                            wave_sync();
                            accd = lane < m ? w.cdist[lane] : 0.f;
                            accv = lane < m ? w.cand[lane] : 0u;
                            held = m == 64 ? ~0ull : ((1ull << m) - 1);
                            wave_sync();
                        }
                    }
                    if (pre_v == 0xFFFFFFFFu || epos >= W.L ||
                        ballot(((held >> lane) & 1ull) && key_less(accd, accv, pre_d, pre_v)) != 0)
                        break;
                    W.mark_explored(epos);
                    ids = pre_id;
                }
                if (!held) continue;
            }
            W.merge_chunk(accd, accv, held, w, on_first);
        }
    }
    const bool chained = a.stride0 <= 64;
    while (!chained && W.pop(&pv)) {
        expansions++;
        const uint32_t* row = a.ell0 + size_t(pv) * a.stride0;
        const bool hit = pv == pre_v;
        const uint32_t first = hit ? pre_id : (lane < a.stride0 ? row[lane] : 0xFFFFFFFFu);
        bool predicted = false;
        for (uint32_t base = 0; base < a.stride0; base += 64) {
            const uint32_t id = base == 0 ? first : ((base + lane < a.stride0) ? row[base + lane] : 0xFFFFFFFFu);
            if (ballot(id != 0xFFFFFFFFu) == 0) break;
            const uint32_t cnt = collect<K, (U8R > 0)>(w, id, a.n, a.hmask, a.vis_mode);
            if (!predicted) {
                predicted = true;
                pre_v = W.peek(&pre_d);
                pre_id = (pre_v != 0xFFFFFFFFu && lane < a.stride0) ? a.ell0[size_t(pre_v) * a.stride0 + lane]
                                                                     : 0xFFFFFFFFu;
            }
            if (cnt == 0) continue;
            any_distances<CT, WIDE, MF, (U8R > 0 ? U8R : 1)>(a, w, cnt, u, metric, qn2);
            ndc += cnt;
            const float myd = lane < cnt ? w.cdist[lane] : 0.f;
            const uint32_t myv = lane < cnt ? w.cand[lane] : 0u;
            wave_sync();
            const uint64_t P = (cnt == 64 ? ~0ull : ((1ull << cnt) - 1)) & (W.full() ? ballot(myd < W.lb) : ~0ull);
            if (!P) continue;
            W.merge_chunk(myd, myv, P, w, [&](float d0, uint32_t v0) {
                if (pre_v == 0xFFFFFFFFu || key_less(d0, v0, pre_d, pre_v)) {
                    pre_v = v0;
                    pre_d = d0;
                    pre_id = lane < a.stride0 ? a.ell0[size_t(v0) * a.stride0 + lane] : 0xFFFFFFFFu;
                }
            });
        }
        if (!predicted) pre_v = 0xFFFFFFFFu;  // empty row: no read in flight
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

template <int K, bool U8>
__device__ __forceinline__ WaveLds wave_lds(const RouterArgs& a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t wid = threadIdx.x / 64;
    const size_t per_wave = wave_lds_bytes(a.D, K, U8);
    unsigned char* base = smem + wid * per_wave;
    WaveLds w;
    w.wd = reinterpret_cast<float*>(base);
    w.wv = reinterpret_cast<uint32_t*>(w.wd + 64 * K);
    w.cdist = w.wd;
    w.cand = w.wv;
    w.vis = reinterpret_cast<unsigned char*>(w.wv + 64 * K);
    w.q = reinterpret_cast<float*>(w.vis + vis_bytes(K, U8));
    return w;
}

__device__ __forceinline__ void add_stats(const RouterArgs& a, uint32_t qi, unsigned long long ndc,
                                          unsigned long long exps, unsigned long long ovf) {
    if (lane_id() == 0 && a.stats) {
        atomicAdd(a.stats + 0, ndc);
        atomicAdd(a.stats + 1, exps);
        if (ovf) {
            atomicAdd(a.stats + 2, ovf);
            uint32_t* list = reinterpret_cast<uint32_t*>(a.stats + 4);
            const uint32_t at = atomicAdd(list, 1u);
            if (at < kOverflowList) list[1 + at] = qi;
        }
    }
}

// Plain router: one wavefront per query, the grid covers the batch.
template <int K, int CT, int MF, int U8R>
__global__ __launch_bounds__(kWavesPerBlock * 64) void router_kernel(RouterArgs a) {
    const WaveLds w = wave_lds<K, (U8R > 0)>(a);
    const uint32_t qi = blockIdx.x * kWavesPerBlock + threadIdx.x / 64;
    if (qi >= a.nq) return;
    const QueryStats st = route_one<K, CT, MF, (U8R == 0), U8R>(a, w, qi);
    add_stats(a, qi, st.ndc, st.expansions, st.overflow);
}

// Batched router: a persistent grid sized to the resident wave slots pulls
// query indices from a device work counter until the batch is drained, so
// queries of uneven work (mixed topk / queue, hop counts) balance across
// waves and no wave idles at a launch tail.  WAVES = the minimum resident
// waves per SIMD the register allocation must allow (1 = no constraint);
// WIDE = 32 rows per FP32 distance pass instead of 16; U8R > 0 = with the
// one-byte path, 16 * U8R rows per pass.
template <int K, int CT, int MF, int WAVES, bool WIDE, int U8R>
__global__ __launch_bounds__(kWavesPerBlock * 64, WAVES) void router_persistent_kernel(RouterArgs a) {
    const WaveLds w = wave_lds<K, (U8R > 0)>(a);
    while (true) {
        uint32_t qi = 0;
        if (lane_id() == 0) qi = atomicAdd(a.work_counter, 1u);
        qi = readfirst_u(qi);
        if (qi >= a.nq) break;
        const QueryStats st = route_one<K, CT, MF, WIDE, U8R>(a, w, qi);
        add_stats(a, qi, st.ndc, st.expansions, st.overflow);
    }
}

}  // namespace

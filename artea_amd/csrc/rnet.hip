// R-net insertion (index/stacked_rgraph/index_factory.hpp:192-583) on MI355X.
//
// The reference inserts vertex after vertex (threads race on shared rows).
// Here the vertices are inserted in waves: every vertex of a wave plans its
// insertion against the graph as it stood when the wave began (one
// wavefront per vertex: the descent from the top level with the search
// queue, the covering rule for its level, and per level the select search
// and the RNG pruning of the results), then the wave is applied in
// insertion order: levels, forward rows, and the reverse edges, which are
// sorted by (level, target, inserting vertex) and applied target by target
// in that order (an append while the target's row has room, else the row
// plus the new entry re-pruned).  A wave ends after its first vertex that
// grows the hierarchy; the rest plan again in the next wave.  Waves of one
// vertex are the reference's serial insertion.
//
// The descent keeps one visited set across its levels (a vertex seen on a
// higher level is not evaluated again lower down), so it uses an exact
// bitmap per wavefront slot in HBM; the select searches start from a clear
// set each and use the router's LDS cache (results do not depend on it).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "rnet.hpp"
#include "router_device.hpp"

using namespace artea;
using namespace artea::dev;

namespace {

constexpr uint32_t kInvalidLevelR = 0xFFFFFFFFu;
constexpr uint32_t kNewR = 0x80000000u;
constexpr uint32_t kMaskR = 0x7FFFFFFFu;
constexpr uint64_t kRnetStream = 0x726E6574ull;  // entry draws
constexpr uint32_t kTouched = 16384;            // bitmap words remembered for clearing
constexpr uint32_t kCandMax = 256;              // select results per level
constexpr uint32_t kRowMax = 128;               // row capacity
constexpr uint32_t kRowSlots = kRowMax / 64;    // row entries per lane
constexpr uint32_t kMaxLevels = 64;
constexpr uint32_t kRnetOverflowList = 1024;    // vertices to plan again with a longer list
constexpr size_t kRnetStatsBytes = 4 * sizeof(unsigned long long) + kRnetOverflowList * sizeof(uint32_t);

__device__ __forceinline__ uint64_t mix64r(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// The oracle's counter-based draw in [lo, hi).
__device__ __forceinline__ uint32_t draw(uint64_t seed, uint64_t stream, uint32_t row, uint32_t j, uint32_t lo,
                                         uint32_t hi) {
    const uint64_t x = mix64r(mix64r(mix64r(seed ^ mix64r(stream)) ^ uint64_t(row)) + uint64_t(j));
    return lo + uint32_t((((x >> 32) * uint64_t(hi - lo)) >> 32));
}

struct RnetArgs {
    const float* vecs;
    uint32_t D, C;
    int metric;
    uint2* const* rows;     // [L + 1] -> n * cap(l) (null where unused)
    uint32_t* const* cnt;   // [L + 1] -> n
    const float* radius;    // [L + 1]
    uint32_t L, ul_max, bl_max;
    uint32_t search_qs, ul_sel, bl_sel;
    float inv_scale;
    uint32_t insert_on_l0;
    uint64_t seed;
    uint32_t top;           // kInvalidLevelR: empty graph
    const uint32_t* top_bucket;
    uint32_t top_size;
    uint32_t first;         // the wave's vertices: first + list[i] or first + i
    const uint32_t* list;
    uint32_t count;
    uint32_t* work;
    uint32_t* out_h;
    uint32_t* out_grow;
    uint2* fwd;             // count * fstride
    uint32_t* fcnt;         // count * (L + 1)
    uint32_t fstride;
    uint2* stash;           // slots * (L + 1) * search_qs
    uint32_t* stash_n;      // slots * (L + 1)
    uint32_t* bitmap;       // slots * bm_words
    uint32_t bm_words;
    uint32_t* touched;      // slots * kTouched
    unsigned long long* stats;  // [0] distances, [1] expansions, [2] overflowed vertices, [3] list count, list
};

__device__ __forceinline__ uint32_t cap_of(const RnetArgs& r, uint32_t l) { return l == 0 ? r.bl_max : r.ul_max; }
__device__ __forceinline__ uint32_t fwd_off(const RnetArgs& r, uint32_t l) {
    return l == 0 ? 0u : r.bl_max + (l - 1) * r.ul_max;
}

struct RnetLds {
    float min_d[kMaxLevels];
    uint32_t cv[kCandMax];
    float cd[kCandMax];
    uint32_t kv[kRowMax];
    float kd[kRowMax];
};

// HierarchicalPruningUpdater::update_impl (refiner/updaters/
// hierarchical_pruning_updater.hpp:96-138) on av/ad[0..m): the first entry
// is kept; each later one, while fewer than max_nbr are kept, is kept iff
// no kept entry (pairs of two old entries skipped) is closer to it than its
// distance * inv_scale.  Kept entries go to bv/bd, marked old.  The kept
// entries are tested together, one per lane.
template <int CT = 0>
__device__ uint32_t hier_prune(const float* vecs, uint32_t D, uint32_t C, int metric, const uint32_t* av,
                               const float* ad, uint32_t m, uint32_t* bv, float* bd, uint32_t max_nbr,
                               float inv_scale, unsigned long long* pairs) {
    const uint32_t lane = lane_id();
    if (m == 0) return 0;
    if (lane == 0) {
        bv[0] = av[0];
        bd[0] = ad[0];
    }
    wave_sync();
    uint32_t nk = 1, nk_new = (av[0] & kNewR) ? 1u : 0u;
    unsigned long long np = 0;
    for (uint32_t i = 1; i < m && nk < max_nbr; ++i) {
        const uint32_t ci = av[i];
        const float di = ad[i];
        const bool old_i = (ci & kNewR) == 0;
        const float thr = di * inv_scale;
        bool pruned = false;
        if (!(old_i && nk_new == 0)) {
            const float* cvec = vecs + size_t(ci & kMaskR) * D;
            for (uint32_t base = 0; base < nk && !pruned; base += 64) {
                const uint32_t j = base + lane;
                bool hit = false, need = false;
                if (j < nk) {
                    const uint32_t rj = bv[j];
                    need = !(old_i && (rj & kNewR) == 0);
                    if (need) hit = lane_distance<CT>(cvec, vecs + size_t(rj & kMaskR) * D, C, metric) < thr;
                }
                np += uint64_t(__popcll(ballot(need)));
                pruned = ballot(hit) != 0;
            }
        }
        if (!pruned) {
            if (lane == 0) {
                bv[nk] = ci;
                bd[nk] = di;
            }
            ++nk;
            nk_new += old_i ? 0u : 1u;
            wave_sync();
        }
    }
    for (uint32_t j = lane; j < nk; j += 64) bv[j] &= kMaskR;
    wave_sync();
    *pairs += np;
    return nk;
}

// Exact test-and-set of a chunk of ids in the slot's bitmap; the fresh ids
// go to w.cand in lane order.  Words that were empty are remembered for the
// clear at the end of the insertion.
__device__ __forceinline__ uint32_t exact_collect(const WaveLds& w, uint32_t id, uint32_t n, uint32_t* bm,
                                                  uint32_t* touched, uint32_t& tcount) {
    bool fresh = false, rec = false;
    const uint32_t word = id >> 5;
    if (id < n) {
        const uint32_t bit = 1u << (id & 31u);
        const uint32_t old = atomicOr(bm + word, bit);
        fresh = (old & bit) == 0;
        rec = old == 0;
    }
    const uint64_t rm = ballot(rec);
    if (rec) {
        const uint32_t at = tcount + lanes_below(rm);
        if (at < kTouched) touched[at] = word;
    }
    tcount += uint32_t(__popcll(rm));
    const uint64_t m = ballot(fresh);
    if (fresh) w.cand[lanes_below(m)] = id;
    wave_sync();
    return uint32_t(__popcll(m));
}

// reset_exploration (std_candidate_queue.hpp:421-426): the list keeps T
// (its first min(n, L) entries), every entry unexplored.
template <int K>
__device__ __forceinline__ void reset_exploration(CandList<K>& W) {
    const uint32_t lane = lane_id();
    const uint32_t keep = min(W.n, W.L);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t idx = uint32_t(k) * 64 + lane;
        if (idx >= keep) {
            W.d[k] = __int_as_float(0x7F800000);
            W.v[k] = 0xFFFFFFFFu;
        } else {
            W.v[k] &= kVidMask;
        }
    }
    W.n = keep;
}

// Beam loop over level l's rows from W (already reset and marked).  CT > 0:
// rows of CT 16-float chunks (unrolled distance chains, 32 rows per pass).
template <int K, int MF, bool EXACT, int CT>
__device__ void beam_level(const RnetArgs& r, const RouterArgs& a, const WaveLds& w, CandList<K>& W, uint32_t l,
                           uint32_t* bm, uint32_t* touched, uint32_t& tcount, float qn2, unsigned long long& ndc,
                           unsigned long long& exps) {
    const uint32_t lane = lane_id();
    const int metric = family_metric<MF>(a.metric);
    const U8Query<CT> u{};
    const uint32_t capl = cap_of(r, l);
    const uint2* rows = r.rows[l];
    const uint32_t* cnt = r.cnt[l];
    uint32_t cur;
    while (W.pop(&cur)) {
        exps++;
        const uint32_t nr = cnt[cur];
        const uint2* row = rows + size_t(cur) * capl;
        for (uint32_t base = 0; base < nr; base += 64) {
            const uint32_t id = base + lane < nr ? (row[base + lane].x & kMaskR) : 0xFFFFFFFFu;
            const uint32_t c = EXACT ? exact_collect(w, id, a.n, bm, touched, tcount)
                                     : collect<K, false>(w, id, a.n, a.hmask, a.vis_mode);
            if (c == 0) continue;
            any_distances<CT, (CT > 0), MF, 1>(a, w, c, u, metric, qn2);
            ndc += c;
            const float myd = lane < c ? w.cdist[lane] : 0.f;
            const uint32_t myv = lane < c ? w.cand[lane] : 0u;
            wave_sync();
            const uint64_t P = (c == 64 ? ~0ull : ((1ull << c) - 1)) & (W.full() ? ballot(myd < W.lb) : ~0ull);
            if (!P) continue;
            W.merge_chunk(myd, myv, P, w, [](float, uint32_t) {});
        }
    }
}

// Push sorted (vid, dist) entries into W in order (try_push), marking them
// in the LDS cache.
template <int K>
__device__ void seed_list(const RouterArgs& a, const WaveLds& w, CandList<K>& W, const uint2* e, uint32_t m) {
    const uint32_t lane = lane_id();
    for (uint32_t base = 0; base < m; base += 64) {
        const uint32_t j = base + lane;
        const uint32_t id = j < m ? e[j].x : 0xFFFFFFFFu;
        const float dd = j < m ? __uint_as_float(e[j].y) : 0.f;
        const VisKey key = vis_key<K, false>(id, a.hmask, a.vis_mode);
        const uint32_t c1 = vis_load(w.vis, key.s1, a.vis_mode), c2 = vis_load(w.vis, key.s2, a.vis_mode);
        const bool fresh = id < a.n && c1 != key.t1 && (key.t2 == 0 || c2 != key.t2);
        wave_sync();
        if (fresh) {
            const bool second = c1 != 0 && key.t2 != 0 && c2 == 0;
            vis_store(w.vis, second ? key.s2 : key.s1, second ? key.t2 : key.t1, a.vis_mode);
        }
        wave_sync();
        const uint64_t P = ballot(j < m) & (W.full() ? ballot(dd < W.lb) : ~0ull);
        if (P) W.merge_chunk(dd, id, P, w, [](float, uint32_t) {});
    }
}

template <int K, int MF, int CT>
__device__ void plan_vertex(const RnetArgs& r, const RouterArgs& a, const WaveLds& w, RnetLds& S, uint32_t slot,
                            uint32_t i) {
    const uint32_t lane = lane_id();
    const uint32_t q = r.first + (r.list ? r.list[i] : i);
    const uint32_t oi = r.list ? r.list[i] : i;  // output slot
    const float* qv = a.vecs + size_t(q) * a.D;
    for (uint32_t j = lane; j < a.D; j += 64) w.q[j] = qv[j];
    for (uint32_t l = lane; l <= r.L; l += 64) S.min_d[l] = FLT_MAX;
    wave_sync();
    const float qn2 = MF == 2 ? quad_norm2(w.q, int(a.C), lane & 3u) : 0.f;
    const U8Query<CT> u{};
    uint2* stash = r.stash + size_t(slot) * (r.L + 1) * r.search_qs;
    uint32_t* stash_n = r.stash_n + size_t(slot) * (r.L + 1);
    for (uint32_t l = lane; l <= r.L; l += 64) stash_n[l] = 0;
    uint32_t* bm = r.bitmap + size_t(slot) * r.bm_words;
    uint32_t* touched = r.touched + size_t(slot) * kTouched;
    uint32_t tcount = 0;
    unsigned long long ndc = 0, exps = 0, pairs = 0;
    uint32_t overflow = 0;
    const uint32_t top = r.top;

    // Step A: descent top .. 1 from one apex vertex, one visited set.
    if (top != kInvalidLevelR && top >= 1) {
        CandList<K> W;
        W.init(r.search_qs);
        const uint32_t e = r.top_bucket[draw(r.seed, kRnetStream, q, 0, 0, r.top_size)];
        if (lane == 0) w.cand[0] = e;
        wave_sync();
        any_distances<CT, false, MF, 1>(a, w, 1, u, family_metric<MF>(a.metric), qn2);
        const float ed = w.cdist[0];
        wave_sync();
        ndc++;
        W.insert(ed, e);
        for (uint32_t l = top; l >= 1; --l) {
            reset_exploration(W);
            if (W.n > 0) {
                // beam_search marks the queue's entries visited
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint32_t idx = uint32_t(k) * 64 + lane;
                    const uint32_t id = idx < W.n ? (W.v[k] & kVidMask) : 0xFFFFFFFFu;
                    exact_collect(w, id, a.n, bm, touched, tcount);
                }
                beam_level<K, MF, true, CT>(r, a, w, W, l, bm, touched, tcount, qn2, ndc, exps);
            }
            const uint32_t nt = min(W.n, W.L);
            const float best = W.get_d(0);
            if (nt > 0 && lane == 0) S.min_d[l] = best;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t idx = uint32_t(k) * 64 + lane;
                if (idx < nt) stash[size_t(l) * r.search_qs + idx] = make_uint2(W.v[k] & kVidMask, __float_as_uint(W.d[k]));
            }
            if (lane == 0) stash_n[l] = nt;
            overflow += W.overflow;
            W.overflow = 0;
            wave_sync();
            if (l == 1) break;
        }
    }
    if (r.insert_on_l0 && stash_n[1] > 0) {
        const uint32_t m = stash_n[1];
        for (uint32_t j = lane; j < m; j += 64) stash[j] = stash[r.search_qs + j];
        if (lane == 0) stash_n[0] = m;
        wave_sync();
    }
    // Step B: the covering rule.
    uint32_t h;
    if (top == kInvalidLevelR) {
        h = 1;
    } else {
        h = min(top + 1, r.L);
        for (uint32_t k = 1; k <= top; ++k)
            if (S.min_d[k] <= r.radius[k]) {
                h = k - 1;
                break;
            }
    }
    const bool grows = top == kInvalidLevelR || h > top;
    // Step D's searches: per level, the select search from the stashed
    // queue with the select capacity, then the pruning of its results.
    uint32_t* fc = r.fcnt + size_t(oi) * (r.L + 1);
    for (uint32_t l = lane; l <= r.L; l += 64) fc[l] = 0;
    wave_sync();
    for (uint32_t l = r.insert_on_l0 ? 0u : 1u; l <= h; ++l) {
        const uint32_t sm = stash_n[l];
        if (sm == 0) continue;
        CandList<K> W;
        W.init(l == 0 ? r.bl_sel : r.ul_sel);
        for (uint32_t j = lane; j < vis_bytes(K, false) / 4; j += 64) reinterpret_cast<uint32_t*>(w.vis)[j] = 0u;
        wave_sync();
        seed_list<K>(a, w, W, stash + size_t(l) * r.search_qs, sm);
        beam_level<K, MF, false, CT>(r, a, w, W, l, nullptr, nullptr, tcount, qn2, ndc, exps);
        overflow += W.overflow;
        const uint32_t nt = min(W.n, W.L);
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t idx = uint32_t(k) * 64 + lane;
            const uint32_t v = W.v[k] & kVidMask;
            const bool keep = idx < nt && v != q;
            const uint64_t km = ballot(keep);
            if (keep) {
                S.cv[m + lanes_below(km)] = v | kNewR;
                S.cd[m + lanes_below(km)] = W.d[k];
            }
            m += uint32_t(__popcll(km));
        }
        wave_sync();
        const uint32_t nk = hier_prune<CT>(a.vecs, a.D, a.C, a.metric, S.cv, S.cd, m, S.kv, S.kd, cap_of(r, l),
                                           r.inv_scale, &pairs);
        uint2* out = r.fwd + size_t(oi) * r.fstride + fwd_off(r, l);
        for (uint32_t j = lane; j < nk; j += 64) out[j] = make_uint2(S.kv[j], __float_as_uint(S.kd[j]));
        if (lane == 0) fc[l] = nk;
        wave_sync();
    }
    if (lane == 0) {
        r.out_h[oi] = h;
        r.out_grow[oi] = grows ? 1u : 0u;
        atomicAdd(r.stats + 0, ndc + pairs);
        atomicAdd(r.stats + 1, exps);
        if (overflow) {
            atomicAdd(r.stats + 2, 1ull);
            const uint32_t at = uint32_t(atomicAdd(r.stats + 3, 1ull));
            if (at < kRnetOverflowList) reinterpret_cast<uint32_t*>(r.stats + 4)[at] = oi;
        }
    }
    // Clear the exact visited set for the slot's next vertex.
    if (tcount <= kTouched) {
        for (uint32_t j = lane; j < tcount; j += 64) atomicExch(bm + touched[j], 0u);
    } else {
        for (uint32_t j = lane; j < r.bm_words; j += 64) atomicExch(bm + j, 0u);
    }
    __threadfence();
    wave_sync();
}

template <int K, int MF, int CT>
__global__ __launch_bounds__(kWavesPerBlock * 64) void rnet_plan_kernel(RnetArgs r, RouterArgs a) {
    const WaveLds w = wave_lds<K, false>(a);
    extern __shared__ __align__(16) unsigned char smem[];
    RnetLds& S = *reinterpret_cast<RnetLds*>(smem + kWavesPerBlock * wave_lds_bytes(a.D, K, false) +
                                              (threadIdx.x / 64) * sizeof(RnetLds));
    const uint32_t slot = blockIdx.x * kWavesPerBlock + threadIdx.x / 64;
    while (true) {
        uint32_t i = 0;
        if (lane_id() == 0) i = atomicAdd(r.work, 1u);
        i = readfirst_u(i);
        if (i >= r.count) break;
        plan_vertex<K, MF, CT>(r, a, w, S, slot, i);
    }
}

// Forward rows of the applied vertices (their rows are empty before).
__global__ void forward_kernel(RnetArgs r, uint32_t applied) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(applied) * r.fstride) return;
    const uint32_t i = uint32_t(t / r.fstride), o = uint32_t(t % r.fstride);
    uint32_t l, k;
    if (o < r.bl_max) {
        l = 0;
        k = o;
    } else {
        l = 1 + (o - r.bl_max) / r.ul_max;
        k = (o - r.bl_max) % r.ul_max;
    }
    if (l > r.L || !r.rows[l]) return;
    const uint32_t m = r.fcnt[size_t(i) * (r.L + 1) + l];
    const uint32_t q = r.first + i;
    if (k == 0) r.cnt[l][q] = m;
    if (k < m) r.rows[l][size_t(q) * cap_of(r, l) + k] = r.fwd[size_t(i) * r.fstride + o];
}

// Reverse-edge records: per applied vertex i, its total over the levels.
__global__ void rec_count_kernel(RnetArgs r, uint32_t applied, uint32_t* counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= applied) return;
    uint32_t s = 0;
    for (uint32_t l = 0; l <= r.L; ++l) s += r.fcnt[size_t(i) * (r.L + 1) + l];
    counts[i] = s;
}

// key = level:6 | target:31 | vertex index:27, value = distance bits.
__global__ void rec_emit_kernel(RnetArgs r, uint32_t applied, const uint32_t* offs, unsigned long long* keys,
                                uint32_t* vals) {
    const uint32_t i = blockIdx.x;
    if (i >= applied) return;
    uint32_t pos = offs[i];
    for (uint32_t l = 0; l <= r.L; ++l) {
        const uint32_t m = r.fcnt[size_t(i) * (r.L + 1) + l];
        const uint2* f = r.fwd + size_t(i) * r.fstride + fwd_off(r, l);
        for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) {
            keys[pos + k] = (static_cast<unsigned long long>(l) << 58) |
                            (static_cast<unsigned long long>(f[k].x & kMaskR) << 27) | i;
            vals[pos + k] = f[k].y;
        }
        pos += m;
    }
}

__global__ void seg_head_kernel(const unsigned long long* keys, uint32_t m, uint32_t* flags) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    flags[j] = (j == 0 || (keys[j] >> 27) != (keys[j - 1] >> 27)) ? 1u : 0u;
}

// _write_reverse_edges (index_factory.hpp:528-580) for one target row: the
// records in vertex order; an append while the row has room, else the row
// plus the new entry ordered by distance (stable) and re-pruned.
//
// A full row whose entries are all old (every row after its first re-prune)
// takes a shortcut with the same result: old/old pairs are never compared,
// so the old entries before the new one stay, the new one stays iff none of
// them is closer to it than its distance / scale, and a later old entry
// goes iff the new one stayed and is closer to it than its own distance /
// scale; the first cap survivors are kept.  The new entry's distances to
// the row are computed in one pass, one lane per entry.
template <int CT>
__global__ __launch_bounds__(256) void reverse_apply_kernel(RnetArgs r, const unsigned long long* keys,
                                                            const uint32_t* vals, const uint32_t* heads,
                                                            const uint32_t* n_heads, uint32_t m, uint32_t* seg_work) {
    __shared__ uint32_t s_rv[4][kRowMax + 1], s_tv[4][kRowMax + 1];
    __shared__ float s_rd[4][kRowMax + 1], s_td[4][kRowMax + 1];
    const uint32_t wid = threadIdx.x / 64, lane = lane_id();
    const uint32_t nh = *n_heads;
    unsigned long long pairs = 0;
    while (true) {
        uint32_t sgi = 0;
        if (lane == 0) sgi = atomicAdd(seg_work, 1u);
        sgi = readfirst_u(sgi);
        if (sgi >= nh) break;
        const uint32_t b = heads[sgi], e = sgi + 1 < nh ? heads[sgi + 1] : m;
        const unsigned long long k0 = keys[b];
        const uint32_t l = uint32_t(k0 >> 58), t = uint32_t((k0 >> 27) & kMaskR);
        const uint32_t cap = cap_of(r, l);
        uint2* row = r.rows[l] + size_t(t) * cap;
        uint32_t n = r.cnt[l][t];
        uint32_t* rv = s_rv[wid];
        float* rd = s_rd[wid];
        for (uint32_t j = lane; j < n; j += 64) {
            const uint2 x = row[j];
            rv[j] = x.x;
            rd[j] = __uint_as_float(x.y);
        }
        wave_sync();
        for (uint32_t rec = b; rec < e; ++rec) {
            const uint32_t qv = (r.first + uint32_t(keys[rec] & ((1ull << 27) - 1))) | kNewR;
            const float qd = __uint_as_float(vals[rec]);
            if (n < cap) {
                if (lane == 0) {
                    rv[n] = qv;
                    rd[n] = qd;
                }
                ++n;
                wave_sync();
                continue;
            }
            uint32_t* tv = s_tv[wid];
            float* td = s_td[wid];
            bool irregular = false;  // a new entry, or a row out of distance order
            uint32_t pn = 0;         // the new entry's place: after every entry not farther
            for (uint32_t j = lane; j < cap; j += 64) {
                irregular |= (rv[j] & kNewR) != 0 || (j > 0 && rd[j - 1] > rd[j]);
                pn += rd[j] <= qd ? 1u : 0u;
            }
            if (ballot(irregular) == 0) {
                for (int off = 32; off >= 1; off >>= 1) pn += __shfl_xor(pn, off);
                if (pn >= cap) continue;  // the new entry is past the cap: nothing changes
                const float* qvec = r.vecs + size_t(qv & kMaskR) * r.D;
                const float thr_new = qd * r.inv_scale;
                float dn[kRowSlots];
                bool hit_before = false;
#pragma unroll
                for (uint32_t s = 0; s < kRowSlots; ++s) {
                    const uint32_t j = s * 64 + lane;
                    dn[s] = 0.f;
                    if (j < cap) {
                        dn[s] =
                            lane_distance<CT>(qvec, r.vecs + size_t(rv[j] & kMaskR) * r.D, r.C, r.metric);
                        hit_before |= j < pn && dn[s] < thr_new;
                    }
                }
                pairs += cap;
                if (pn > 0 && ballot(hit_before) != 0) continue;  // the new entry is pruned: nothing changes
                // The new entry stays; the old entries from pn on stay unless
                // it is closer to them than their distance / scale.
                uint32_t out = pn + 1;
#pragma unroll
                for (uint32_t s = 0; s < kRowSlots; ++s) {
                    const uint32_t j = s * 64 + lane;
                    const bool keep = j >= pn && j < cap && !(dn[s] < rd[j] * r.inv_scale);
                    const uint64_t km = ballot(keep);
                    const uint32_t at = out + lanes_below(km);
                    if (keep && at < cap) {
                        tv[at] = rv[j];
                        td[at] = rd[j];
                    }
                    out += uint32_t(__popcll(km));
                }
                wave_sync();
                const uint32_t nn = min(out, cap);
                for (uint32_t j = pn + lane; j < nn; j += 64) {
                    rv[j] = j == pn ? (qv & kMaskR) : tv[j];
                    rd[j] = j == pn ? qd : td[j];
                }
                n = nn;
                wave_sync();
                continue;
            }
            // general case: cap + 1 entries ordered by distance, ties in
            // place order, then the hierarchical pruning
            for (uint32_t p = lane; p <= cap; p += 64) {
                const float dp = p < cap ? rd[p] : qd;
                uint32_t rank = 0;
                for (uint32_t j = 0; j <= cap; ++j) {
                    const float dj = j < cap ? rd[j] : qd;
                    rank += (dj < dp || (dj == dp && j < p)) ? 1u : 0u;
                }
                tv[rank] = p < cap ? rv[p] : qv;
                td[rank] = dp;
            }
            wave_sync();
            n = hier_prune<CT>(r.vecs, r.D, r.C, r.metric, tv, td, cap + 1, rv, rd, cap, r.inv_scale, &pairs);
        }
        for (uint32_t j = lane; j < n; j += 64) row[j] = make_uint2(rv[j], __float_as_uint(rd[j]));
        if (lane == 0) r.cnt[l][t] = n;
        wave_sync();
    }
    if (lane_id() == 0 && pairs) atomicAdd(r.stats + 0, pairs);
}

}  // namespace

namespace {

using PlanFn = void (*)(RnetArgs, RouterArgs);

// L2 rows of 96, 112 or 128 floats (6, 7 or 8 chunks) get the unrolled
// chains; other widths and the DOT / cosine families the generic ones.
template <int K>
PlanFn plan_fn(int metric, uint32_t chunks) {
    if (metric == kDot) return rnet_plan_kernel<K, 1, 0>;
    if (metric == kCosine) return rnet_plan_kernel<K, 2, 0>;
    switch (chunks) {
    case 6: return rnet_plan_kernel<K, 0, 6>;
    case 7: return rnet_plan_kernel<K, 0, 7>;
    case 8: return rnet_plan_kernel<K, 0, 8>;
    default: return rnet_plan_kernel<K, 0, 0>;
    }
}
PlanFn plan_fn_k(int K, int metric, uint32_t chunks) {
    return K == 2 ? plan_fn<2>(metric, chunks) : (K == 4 ? plan_fn<4>(metric, chunks) : plan_fn<8>(metric, chunks));
}
using ReverseFn = void (*)(RnetArgs, const unsigned long long*, const uint32_t*, const uint32_t*, const uint32_t*,
                           uint32_t, uint32_t*);
ReverseFn reverse_fn(int metric, uint32_t chunks) {
    if (metric == kDot || metric == kCosine) return reverse_apply_kernel<0>;
    switch (chunks) {
    case 6: return reverse_apply_kernel<6>;
    case 7: return reverse_apply_kernel<7>;
    case 8: return reverse_apply_kernel<8>;
    default: return reverse_apply_kernel<0>;
    }
}
size_t plan_lds(uint32_t D, int K) { return kWavesPerBlock * (wave_lds_bytes(D, K, false) + sizeof(RnetLds)); }

// Device buffers of one build, freed together.
struct Bufs {
    std::vector<void*> ptrs;
    ~Bufs() {
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    template <typename T>
    hipError_t alloc(T** p, size_t count) {
        *p = nullptr;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count * sizeof(T), 16));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

#define RNET_TRY(expr)                                                   \
    do {                                                                 \
        const hipError_t rnet_e_ = (expr);                               \
        if (rnet_e_ != hipSuccess) {                                     \
            set_error(std::string("rnet: ") + hipGetErrorString(rnet_e_)); \
            return ARTEA_EHIP;                                           \
        }                                                                \
    } while (0)

}  // namespace

namespace artea {

void rnet_free(artea_rnet* r) {
    if (!r) return;
    for (uint2* p : r->rows)
        if (p) (void)hipFree(p);
    for (uint32_t* p : r->cnt)
        if (p) (void)hipFree(p);
    if (r->own_stream && r->s) (void)hipStreamDestroy(r->s);
    delete r;
}

int rnet_run(const float* d_vecs, uint32_t n, uint32_t dim_pad, int metric, const artea_rnet_params& p, hipStream_t s,
             artea_rnet** out) {
    ARTEA_CHECK_ARG(out, "rnet: null output");
    ARTEA_CHECK_ARG(d_vecs || n == 0, "rnet: null vectors");
    ARTEA_CHECK_ARG(dim_pad > 0 && dim_pad % 16 == 0, "rnet: dim_pad must be a positive multiple of 16");
    ARTEA_CHECK_ARG(metric >= 0 && metric <= 3, "rnet: metric must be 0..3");
    ARTEA_CHECK_ARG(std::isfinite(p.rnet_beta) && p.rnet_beta > 1.0f, "rnet: rnet_beta must be finite and > 1");
    ARTEA_CHECK_ARG(std::isfinite(p.tau_k) && p.tau_k >= 0.0f, "rnet: tau_k must be finite and >= 0");
    ARTEA_CHECK_ARG(p.scale_coeffs > 0.0f, "rnet: scale must be positive");
    ARTEA_CHECK_ARG(p.ul_max_nbr_size >= 1 && p.ul_max_nbr_size <= kRowMax && p.bl_max_nbr_size >= 1 && p.bl_max_nbr_size <= kRowMax,
                    "rnet: max_nbr_size must be in [1, 128]");
    ARTEA_CHECK_ARG(p.search_nn_qs >= 1 && p.ul_select_nbrs_qs >= p.search_nn_qs && p.bl_select_nbrs_qs >= p.search_nn_qs &&
                        p.ul_select_nbrs_qs <= 224 && p.bl_select_nbrs_qs <= 224,
                    "rnet: queues must satisfy 1 <= search_nn_qs <= select queues <= 224");
    ARTEA_CHECK_ARG(p.wave_max >= 1, "rnet: wave_max must be at least 1");
    ARTEA_CHECK_ARG(n < (1u << 31), "rnet: more than 2^31 vectors");
    std::unique_ptr<artea_rnet, void (*)(artea_rnet*)> R(new artea_rnet(), rnet_free);
    R->n = n;
    R->D = dim_pad;
    R->C = dim_pad / 16;
    R->metric = metric;
    R->d_vecs = d_vecs;
    R->p = p;
    uint32_t L = 1;  // compute_max_allowed_level_id (stacked_rgraph/configs.hpp:241-249)
    if (n > 1000) L = std::max<uint32_t>(uint32_t(std::ceil(std::log(double(n) / 1000.0))), 1u);
    ARTEA_CHECK_ARG(L < kMaxLevels, "rnet: too many levels");
    R->L = L;
    R->highest.assign(n, kInvalidLevelR);
    if (s) {
        R->s = s;
    } else {
        RNET_TRY(hipStreamCreateWithFlags(&R->s, hipStreamNonBlocking));
        R->own_stream = true;
    }
    hipStream_t st = R->s;
    R->rows.assign(L + 1, nullptr);
    R->cnt.assign(L + 1, nullptr);
    for (uint32_t l = p.insert_on_l0 ? 0u : 1u; l <= L; ++l) {
        const uint32_t cap = l == 0 ? p.bl_max_nbr_size : p.ul_max_nbr_size;
        RNET_TRY(hipMalloc(&R->rows[l], std::max<size_t>(size_t(n) * cap * sizeof(uint2), 16)));
        RNET_TRY(hipMalloc(&R->cnt[l], std::max<size_t>(size_t(n) * 4, 16)));
        RNET_TRY(hipMemsetAsync(R->cnt[l], 0, size_t(n) * 4, st));
    }
    *out = nullptr;
    if (n == 0) {
        *out = R.release();
        return ARTEA_OK;
    }
    Bufs B;
    uint2** d_rows = nullptr;
    uint32_t** d_cnt = nullptr;
    float* d_radius = nullptr;
    RNET_TRY(B.alloc(&d_rows, L + 1));
    RNET_TRY(B.alloc(&d_cnt, L + 1));
    RNET_TRY(B.alloc(&d_radius, L + 1));
    std::vector<float> radius(L + 1);
    for (uint32_t h = 0; h <= L; ++h)
        radius[h] = h == 0 ? p.l0_min_distance
                           : float(double(p.l0_min_distance) * (1.0 + double(p.tau_k)) *
                                   std::pow(double(p.rnet_beta), double(h - 1)));
    RNET_TRY(hipMemcpyAsync(d_rows, R->rows.data(), (L + 1) * sizeof(uint2*), hipMemcpyHostToDevice, st));
    RNET_TRY(hipMemcpyAsync(d_cnt, R->cnt.data(), (L + 1) * sizeof(uint32_t*), hipMemcpyHostToDevice, st));
    RNET_TRY(hipMemcpyAsync(d_radius, radius.data(), (L + 1) * 4, hipMemcpyHostToDevice, st));

    const uint32_t qmax = std::max({p.search_nn_qs, p.ul_select_nbrs_qs, p.bl_select_nbrs_qs});
    const int K0 = qmax <= 96 ? 2 : 4;
    int dev = 0, cus = 0;
    RNET_TRY(hipGetDevice(&dev));
    RNET_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // Wave slots: enough blocks for the occupancy of the widest kernel used.
    uint32_t blocks = 0;
    for (int K : {K0, 8}) {
        const PlanFn fn = plan_fn_k(K, metric, dim_pad / 16);
        const size_t lds = plan_lds(dim_pad, K);
        RNET_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(lds)));
        int per_cu = 0;
        RNET_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWavesPerBlock * 64, lds));
        blocks = std::max(blocks, uint32_t(std::max(cus, 1) * std::max(per_cu, 1)));
    }
    const uint32_t slots = blocks * kWavesPerBlock;
    const uint32_t wmax = std::min(p.wave_max, n);
    const uint32_t fstride = p.bl_max_nbr_size + L * p.ul_max_nbr_size;
    const uint32_t bm_words = (n + 31) / 32;
    RnetArgs a{};
    a.vecs = d_vecs;
    a.D = dim_pad;
    a.C = dim_pad / 16;
    a.metric = metric;
    a.rows = d_rows;
    a.cnt = d_cnt;
    a.radius = d_radius;
    a.L = L;
    a.ul_max = p.ul_max_nbr_size;
    a.bl_max = p.bl_max_nbr_size;
    a.search_qs = p.search_nn_qs;
    a.ul_sel = p.ul_select_nbrs_qs;
    a.bl_sel = p.bl_select_nbrs_qs;
    a.inv_scale = 1.0f / p.scale_coeffs;
    a.insert_on_l0 = p.insert_on_l0 ? 1u : 0u;
    a.seed = p.seed;
    a.fstride = fstride;
    a.bm_words = bm_words;
    uint32_t* top_bucket = nullptr;
    RNET_TRY(B.alloc(&top_bucket, n));
    RNET_TRY(B.alloc(&a.work, 1));
    RNET_TRY(B.alloc(&a.out_h, wmax));
    RNET_TRY(B.alloc(&a.out_grow, wmax));
    RNET_TRY(B.alloc(&a.fwd, size_t(wmax) * fstride));
    RNET_TRY(B.alloc(&a.fcnt, size_t(wmax) * (L + 1)));
    RNET_TRY(B.alloc(&a.stash, size_t(slots) * (L + 1) * p.search_nn_qs));
    RNET_TRY(B.alloc(&a.stash_n, size_t(slots) * (L + 1)));
    RNET_TRY(B.alloc(&a.bitmap, size_t(slots) * bm_words));
    RNET_TRY(B.alloc(&a.touched, size_t(slots) * kTouched));
    RNET_TRY(B.alloc(&a.stats, kRnetStatsBytes / sizeof(unsigned long long)));
    RNET_TRY(hipMemsetAsync(a.bitmap, 0, size_t(slots) * bm_words * 4, st));
    a.top_bucket = top_bucket;
    // reverse-edge records
    const size_t max_rec = size_t(wmax) * fstride;
    unsigned long long *keys = nullptr, *keys2 = nullptr;
    uint32_t *vals = nullptr, *vals2 = nullptr, *rc = nullptr, *roff = nullptr, *flags = nullptr, *heads = nullptr,
             *n_heads = nullptr;
    RNET_TRY(B.alloc(&keys, max_rec));
    RNET_TRY(B.alloc(&keys2, max_rec));
    RNET_TRY(B.alloc(&vals, max_rec));
    RNET_TRY(B.alloc(&vals2, max_rec));
    RNET_TRY(B.alloc(&rc, wmax + 1));
    RNET_TRY(B.alloc(&roff, wmax + 1));
    RNET_TRY(B.alloc(&flags, max_rec));
    RNET_TRY(B.alloc(&heads, max_rec));
    RNET_TRY(B.alloc(&n_heads, 1));
    uint32_t* seg_work = nullptr;
    RNET_TRY(B.alloc(&seg_work, 1));
    size_t tmp_sort = 0, tmp_scan = 0, tmp_sel = 0;
    RNET_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, keys, keys2, vals, vals2, int(max_rec), 0, 64, st));
    RNET_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan, rc, roff, int(wmax + 1), st));
    hipcub::CountingInputIterator<uint32_t> iota(0);
    RNET_TRY(hipcub::DeviceSelect::Flagged(nullptr, tmp_sel, iota, flags, heads, n_heads, int(max_rec), st));
    void* tmp = nullptr;
    const size_t tmp_bytes = std::max({tmp_sort, tmp_scan, tmp_sel});
    RNET_TRY(B.alloc(reinterpret_cast<unsigned char**>(&tmp), tmp_bytes));

    RouterArgs ra{};
    ra.vecs = d_vecs;
    ra.vecs_u8 = nullptr;
    ra.n = n;
    ra.D = dim_pad;
    ra.C = dim_pad / 16;
    ra.metric = metric;
    {
        uint32_t Bb = 13;
        while (Bb < 32 && (uint64_t(1) << Bb) < n) ++Bb;
        ra.hmask = uint32_t((uint64_t(1) << Bb) - 1);
        ra.vis_mode = Bb <= 20 ? 0u : (Bb <= 27 ? 1u : 2u);
    }
    auto launch_plan = [&](int K, const uint32_t* list, uint32_t count) -> hipError_t {
        a.list = list;
        a.count = count;
        hipError_t e = hipMemsetAsync(a.work, 0, 4, st);
        if (e != hipSuccess) return e;
        const PlanFn fn = plan_fn_k(K, metric, dim_pad / 16);
        const uint32_t nb = std::max(1u, std::min(blocks, (count + kWavesPerBlock - 1) / kWavesPerBlock));
        hipLaunchKernelGGL(fn, dim3(nb), dim3(kWavesPerBlock * 64), plan_lds(dim_pad, K), st, a, ra);
        return hipGetLastError();
    };

    std::vector<uint32_t> h_h(wmax), h_g(wmax);
    uint32_t top_size = 0;
    uint32_t pos = 0;
    while (pos < n) {
        uint32_t w = pos < 10 ? 1u : std::max(1u, p.wave_div ? pos / p.wave_div : wmax);
        w = std::min({w, wmax, n - pos});
        a.top = R->top;
        a.top_size = top_size;
        a.first = pos;
        RNET_TRY(hipMemsetAsync(a.stats, 0, kRnetStatsBytes, st));
        RNET_TRY(launch_plan(K0, nullptr, w));
        unsigned long long hs[4];
        RNET_TRY(hipMemcpyAsync(hs, a.stats, sizeof hs, hipMemcpyDeviceToHost, st));
        RNET_TRY(hipStreamSynchronize(st));
        R->distances += hs[0];
        R->expansions += hs[1];
        if (hs[2]) {
            // vertices whose list dropped an entry tied with its bound: plan
            // again with 512 entries (all of the wave when the list overflowed)
            const uint32_t nov = uint32_t(hs[2]);
            const bool all = nov > kRnetOverflowList;
            uint32_t* lst = nullptr;
            if (!all) {
                RNET_TRY(B.alloc(&lst, nov));
                RNET_TRY(hipMemcpyAsync(lst, reinterpret_cast<uint32_t*>(a.stats + 4), size_t(nov) * 4,
                                        hipMemcpyDeviceToDevice, st));
            }
            RNET_TRY(hipMemsetAsync(a.stats, 0, kRnetStatsBytes, st));
            RNET_TRY(launch_plan(8, lst, all ? w : nov));
            RNET_TRY(hipMemcpyAsync(hs, a.stats, sizeof hs, hipMemcpyDeviceToHost, st));
            RNET_TRY(hipStreamSynchronize(st));
            R->distances += hs[0];
            R->expansions += hs[1];
            R->unresolved += hs[2];
        }
        RNET_TRY(hipMemcpyAsync(h_h.data(), a.out_h, size_t(w) * 4, hipMemcpyDeviceToHost, st));
        RNET_TRY(hipMemcpyAsync(h_g.data(), a.out_grow, size_t(w) * 4, hipMemcpyDeviceToHost, st));
        RNET_TRY(hipStreamSynchronize(st));
        uint32_t end = w;
        for (uint32_t i = 0; i < w; ++i)
            if (h_g[i]) {
                end = i + 1;
                break;
            }
        // forward rows, then the reverse edges sorted by (level, target, vertex)
        const size_t ft = size_t(end) * fstride;
        hipLaunchKernelGGL(forward_kernel, dim3(uint32_t((ft + 255) / 256)), dim3(256), 0, st, a, end);
        RNET_TRY(hipGetLastError());
        hipLaunchKernelGGL(rec_count_kernel, dim3((end + 255) / 256), dim3(256), 0, st, a, end, rc);
        RNET_TRY(hipMemsetAsync(rc + end, 0, 4, st));
        RNET_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_scan, rc, roff, int(end + 1), st));
        uint32_t m = 0;
        RNET_TRY(hipMemcpyAsync(&m, roff + end, 4, hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL(rec_emit_kernel, dim3(end), dim3(64), 0, st, a, end, roff, keys, vals);
        RNET_TRY(hipGetLastError());
        RNET_TRY(hipStreamSynchronize(st));
        if (m > 0) {
            size_t tb = tmp_bytes;
            RNET_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, keys2, vals, vals2, int(m), 0, 64, st));
            hipLaunchKernelGGL(seg_head_kernel, dim3((m + 255) / 256), dim3(256), 0, st, keys2, m, flags);
            tb = tmp_bytes;
            RNET_TRY(hipcub::DeviceSelect::Flagged(tmp, tb, iota, flags, heads, n_heads, int(m), st));
            RNET_TRY(hipMemsetAsync(seg_work, 0, 4, st));
            hipLaunchKernelGGL(reverse_fn(metric, dim_pad / 16),
                               dim3(std::max(1u, std::min((m + 3) / 4, uint32_t(cus) * 8u))),
                               dim3(256), 0, st, a, keys2, vals2, heads, n_heads, m, seg_work);
            RNET_TRY(hipGetLastError());
        }
        // levels and the top bucket (assignment order)
        std::vector<uint32_t> add;
        for (uint32_t i = 0; i < end; ++i) {
            const uint32_t q = pos + i, h = h_h[i];
            R->highest[q] = h;
            if (h_g[i]) {
                R->top = h;
                top_size = 0;
                add.clear();
            }
            if (h == R->top) add.push_back(q);
        }
        if (!add.empty()) {
            RNET_TRY(hipMemcpyAsync(top_bucket + top_size, add.data(), add.size() * 4, hipMemcpyHostToDevice, st));
            top_size += uint32_t(add.size());
        }
        RNET_TRY(hipStreamSynchronize(st));
        pos += end;
        R->waves++;
    }
    *out = R.release();
    return ARTEA_OK;
}

}  // namespace artea

extern "C" {

int artea_rnet_build(artea_index* idx, int metric, const artea_rnet_params* p, artea_rnet** out) {
    clear_error();
    ARTEA_CHECK_ARG(idx && p && out, "artea_rnet_build: null argument");
    return artea::rnet_run(idx->d_vecs, idx->n, idx->dim_pad, metric, *p, idx->stream, out);
}

int artea_rnet_info(const artea_rnet* r, uint32_t* max_level, uint32_t* top, uint32_t* waves, uint64_t* distances,
                    uint64_t* unresolved) {
    clear_error();
    ARTEA_CHECK_ARG(r, "artea_rnet_info: null handle");
    if (max_level) *max_level = r->L;
    if (top) *top = r->top;
    if (waves) *waves = r->waves;
    if (distances) *distances = r->distances;
    if (unresolved) *unresolved = r->unresolved;
    return ARTEA_OK;
}

int artea_rnet_levels(const artea_rnet* r, uint32_t* highest) {
    clear_error();
    ARTEA_CHECK_ARG(r && highest, "artea_rnet_levels: null argument");
    std::copy(r->highest.begin(), r->highest.end(), highest);
    return ARTEA_OK;
}

int artea_rnet_level_rows(const artea_rnet* r, uint32_t level, uint32_t* counts, uint32_t* entries,
                          uint64_t max_entries, uint64_t* total) {
    clear_error();
    ARTEA_CHECK_ARG(r && counts, "artea_rnet_level_rows: null argument");
    ARTEA_CHECK_ARG(level <= r->L, "artea_rnet_level_rows: level above the maximum");
    const uint32_t n = r->n;
    if (!r->rows[level]) {
        std::fill(counts, counts + n, 0u);
        if (total) *total = 0;
        return ARTEA_OK;
    }
    const uint32_t cap = level == 0 ? r->p.bl_max_nbr_size : r->p.ul_max_nbr_size;
    std::vector<uint2> rows(std::max<size_t>(size_t(n) * cap, 1));
    ARTEA_HIP_TRY(hipMemcpyAsync(counts, r->cnt[level], size_t(n) * 4, hipMemcpyDeviceToHost, r->s));
    ARTEA_HIP_TRY(hipMemcpyAsync(rows.data(), r->rows[level], size_t(n) * cap * sizeof(uint2), hipMemcpyDeviceToHost,
                                 r->s));
    ARTEA_HIP_TRY(hipStreamSynchronize(r->s));
    uint64_t pos = 0;
    for (uint32_t v = 0; v < n; ++v)
        for (uint32_t k = 0; k < counts[v]; ++k, ++pos)
            if (entries && pos < max_entries) {
                entries[2 * pos] = rows[size_t(v) * cap + k].x;
                entries[2 * pos + 1] = rows[size_t(v) * cap + k].y;
            }
    if (total) *total = pos;
    return ARTEA_OK;
}

int artea_rnet_destroy(artea_rnet* r) {
    clear_error();
    artea::rnet_free(r);
    return ARTEA_OK;
}

}  // extern "C"

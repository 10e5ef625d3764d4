// Routing loops of refine_layer (index/artea_graph/index_factory.hpp:378-403)
// on MI355X: RoutingUpdater (refiner/updaters/routing_updater.hpp:84-114)
// with its warm-start SingleLayerRouter query
// (router/single_layer_router.hpp:259-282), the merge of its logs, and
// PruningUpdater (refiner/updaters/pruning_updater.hpp:63-177).
//
// The routing updater only logs into the pivot's own row, so one wavefront
// per pivot does all three steps: the beam search (the router's candidate
// list, visited cache and distance passes, router_device.hpp) over the
// layer's current rows, the merge of the pivot's row with its results in
// LDS, and the pruning of the merged row.  The pass reads the rows and
// writes the pruned rows to a separate buffer, so every pivot routes over
// the graph as it was before the pass, as in the reference.
#include "common.hpp"
#include "device.hpp"
#include "refine_ctx.hpp"
#include "router_device.hpp"

using namespace artea;
using namespace artea::dev;
using namespace artea::rf;

namespace {

constexpr uint32_t kExtMax = 256;  // cap + topk, entries of a merged row
constexpr uint32_t kTopMax = 128;  // router results per pivot

struct RouteLds {
    uint32_t av[kExtMax];
    float ad[kExtMax];
    uint32_t bv[kExtMax];
    float bd[kExtMax];
    uint32_t gv[kTopMax];
    float gd[kTopMax];
};

// Position of (d, v) among the first n keys of (kd, kv) sorted by (distance,
// vid): the number of keys below it, plus the keys equal to it when
// `after_equal`.
__device__ __forceinline__ uint32_t rank_in(const float* kd, const uint32_t* kv, uint32_t n, float d, uint32_t v,
                                            bool after_equal) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        const float md = kd[mid];
        const uint32_t mv = kv[mid] & kMask;
        const bool before = key_less(md, mv, d, v) || (after_equal && md == d && mv == v);
        if (before) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// PruningUpdater on the entries a[0..m) (row order): each is dropped when a
// retained entry (pairs of two old entries skipped) lies closer to it than
// its distance / scale; the retained entries go to b, marked old.  The
// reference stops at the first such entry; the decision does not depend on
// the order, so the retained entries are tested together, one per lane
// (CT > 0: rows of CT chunks, the lane's chain unrolled).
template <int CT>
__device__ uint32_t prune_list(const Ctx& c, const uint32_t* av, const float* ad, uint32_t m, uint32_t* bv,
                               float* bd, float inv_scale, unsigned long long* pairs) {
    const uint32_t lane = lane_id();
    uint32_t nr = 0, nr_new = 0;
    unsigned long long np = 0;
    for (uint32_t i = 0; i < m; ++i) {
        const uint32_t ci = av[i];
        const float di = ad[i];
        const bool old_i = (ci & kNew) == 0;
        const float thr = di * inv_scale;
        bool pruned = false;
        if (!(old_i && nr_new == 0)) {
            const float* cv = c.vecs + size_t(ci & kMask) * c.D;
            for (uint32_t base = 0; base < nr && !pruned; base += 64) {
                const uint32_t j = base + lane;
                bool hit = false, need = false;
                if (j < nr) {
                    const uint32_t rj = bv[j];
                    need = !(old_i && (rj & kNew) == 0);
                    if (need) hit = lane_distance<CT>(cv, c.vecs + size_t(rj & kMask) * c.D, c.C, c.metric) < thr;
                }
                np += uint64_t(__popcll(ballot(need)));
                pruned = ballot(hit) != 0;
            }
        }
        if (!pruned) {
            if (lane == 0) {
                bv[nr] = ci;
                bd[nr] = di;
            }
            ++nr;
            nr_new += old_i ? 0u : 1u;
            wave_sync();
        }
    }
    for (uint32_t j = lane; j < nr; j += 64) bv[j] &= kMask;  // mark_as_old
    wave_sync();
    *pairs += np;
    return nr;
}

// One pivot: warm-start beam search, filter, merge with the row, prune.
// CT > 0: the distance chains unrolled for rows of CT 16-float chunks, and
// 32 rows per distance pass.
template <int K, int MF, int CT>
__device__ void route_row(const Ctx& c, const RouterArgs& a, const RoutingPass& p, const WaveLds& w, RouteLds& R,
                          uint32_t row) {
    const uint32_t lane = lane_id();
    const int metric = family_metric<MF>(a.metric);
    const uint32_t sv = storage_of(c, row);
    const float* pv = c.vecs + size_t(sv) * c.D;
    for (uint32_t i = lane; i < c.D; i += 64) w.q[i] = pv[i];
    for (uint32_t i = lane; i < vis_bytes(K, false) / 4; i += 64) reinterpret_cast<uint32_t*>(w.vis)[i] = 0u;
    wave_sync();
    const float qn2 = MF == 2 ? quad_norm2(w.q, int(c.C), lane & 3u) : 0.f;
    const U8Query<CT> u{};
    const uint32_t n0 = c.cnt[row];
    const uint2* rp = c.rows + size_t(row) * c.cap;
    CandList<K> W;
    W.init(max(p.topk, p.queue));
    unsigned long long ndc = 0, exps = 0;

    // Seeds: the row in order with its cached distances; each marked
    // visited (test_and_set) and pushed by try_push's rule.
    for (uint32_t base = 0; base < n0; base += 64) {
        const uint32_t j = base + lane;
        const uint2 e = j < n0 ? rp[j] : make_uint2(0xFFFFFFFFu, 0u);
        const uint32_t id = j < n0 ? (e.x & kMask) : 0xFFFFFFFFu;
        const VisKey key = vis_key<K, false>(id, a.hmask, a.vis_mode);
        const uint32_t c1 = vis_load(w.vis, key.s1, a.vis_mode), c2 = vis_load(w.vis, key.s2, a.vis_mode);
        const bool fresh = id < a.n && c1 != key.t1 && (key.t2 == 0 || c2 != key.t2);
        wave_sync();
        if (fresh) {
            const bool second = c1 != 0 && key.t2 != 0 && c2 == 0;
            vis_store(w.vis, second ? key.s2 : key.s1, second ? key.t2 : key.t1, a.vis_mode);
        }
        wave_sync();
        const float myd = __uint_as_float(e.y);
        const uint64_t P = ballot(fresh) & (W.full() ? ballot(myd < W.lb) : ~0ull);
        if (P) W.merge_chunk(myd, id, P, w, [](float, uint32_t) {});
    }

    // beam_loop_body over the layer's rows (DynamicFlatRange).  The row of
    // the entry the next pop will most likely return (its count and first 64
    // ids) is read after this row's first ids are consumed and before their
    // distance gathers, whose waits then cover it; a merge whose smallest
    // accepted key precedes that entry reads the new first entry's row
    // instead.  The rows do not change during a routing pass, so a read row
    // is the row the pop finds.
    uint32_t pre_v = 0xFFFFFFFFu, pre_n = 0, pre_id = 0xFFFFFFFFu;
    float pre_d = 0.f;
    const uint32_t head = min(c.cap, 64u);  // the row slots a first chunk may read
    auto read_row = [&](uint32_t v, uint32_t* nr_out, uint32_t* id_out) {
        const uint32_t lr = local_of(c, v);
        const bool ok = lr != 0xFFFFFFFFu;
        const uint32_t nr = ok ? c.cnt[lr] : 0u;
        const uint32_t x = ok && lane < head ? c.rows[size_t(lr) * c.cap + lane].x : 0u;
        *nr_out = nr;
        *id_out = lane < min(nr, head) ? (x & kMask) : 0xFFFFFFFFu;
    };
    auto on_first = [&](float d0, uint32_t v0) {
        if (pre_v == 0xFFFFFFFFu || key_less(d0, v0, pre_d, pre_v)) {
            pre_v = v0;
            pre_d = d0;
            read_row(v0, &pre_n, &pre_id);
        }
    };
    uint32_t cur;
    while (W.pop(&cur)) {
        exps++;
        uint32_t nr, first;
        if (cur == pre_v) {
            nr = pre_n;
            first = pre_id;
        } else {
            read_row(cur, &nr, &first);
        }
        pre_v = 0xFFFFFFFFu;
        bool predicted = false;
        for (uint32_t base = 0; base < nr; base += 64) {
            uint32_t id = first;
            if (base > 0) {
                const uint2* nrow = c.rows + size_t(local_of(c, cur)) * c.cap;
                id = base + lane < nr ? (nrow[base + lane].x & kMask) : 0xFFFFFFFFu;
            }
            const uint32_t cnt = collect<K, false>(w, id, a.n, a.hmask, a.vis_mode);
            if (!predicted) {
                predicted = true;
                pre_v = W.peek(&pre_d);
                if (pre_v != 0xFFFFFFFFu) read_row(pre_v, &pre_n, &pre_id);
            }
            if (cnt == 0) continue;
            any_distances<CT, (CT > 0), MF, 1>(a, w, cnt, u, metric, qn2);
            ndc += cnt;
            const float myd = lane < cnt ? w.cdist[lane] : 0.f;
            const uint32_t myv = lane < cnt ? w.cand[lane] : 0u;
            wave_sync();
            const uint64_t P = (cnt == 64 ? ~0ull : ((1ull << cnt) - 1)) & (W.full() ? ballot(myd < W.lb) : ~0ull);
            if (!P) continue;
            W.merge_chunk(myd, myv, P, w, on_first);
        }
    }

    // extract_results(topk); drop the pivot itself and, when the row is
    // full, results not closer than its entry max_nbr_size - 1.
    const uint32_t kk = min(p.topk, min(W.n, W.L));
    const bool full = n0 >= c.cap;
    const float last = full ? __uint_as_float(rp[c.cap - 1].y) : 0.f;
    uint32_t g = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t idx = uint32_t(k) * 64 + lane;
        const uint32_t v = W.v[k] & kVidMask;
        const float d = W.d[k];
        const bool keep = idx < kk && v != sv && !(full && last <= d);
        const uint64_t km = ballot(keep);
        if (keep) {
            R.gv[g + lanes_below(km)] = v | kNew;
            R.gd[g + lanes_below(km)] = d;
        }
        g += uint32_t(__popcll(km));
    }
    // merge_logs: the row and the logs by (distance, vid), the row entry
    // first on equal keys, then adjacent-unique by vid (nbr_log_table.hpp
    // :127-182), with no degree cap.
    for (uint32_t j = lane; j < n0; j += 64) {
        const uint2 e = rp[j];
        R.av[j] = e.x;
        R.ad[j] = __uint_as_float(e.y);
    }
    wave_sync();
    for (uint32_t j = lane; j < n0; j += 64) {
        const uint32_t at = j + rank_in(R.gd, R.gv, g, R.ad[j], R.av[j] & kMask, false);
        R.bv[at] = R.av[j];
        R.bd[at] = R.ad[j];
    }
    for (uint32_t t = lane; t < g; t += 64) {
        const uint32_t at = t + rank_in(R.ad, R.av, n0, R.gd[t], R.gv[t] & kMask, true);
        R.bv[at] = R.gv[t];
        R.bd[at] = R.gd[t];
    }
    wave_sync();
    const uint32_t total = n0 + g;
    uint32_t m = 0;
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t i = base + lane;
        const bool keep = i < total && (i == 0 || (R.bv[i - 1] & kMask) != (R.bv[i] & kMask));
        const uint64_t km = ballot(keep);
        uint32_t kv = 0;
        float kd = 0.f;
        if (keep) {
            kv = R.bv[i];
            kd = R.bd[i];
        }
        wave_sync();  // av/ad are free again: the row was consumed above
        if (keep) {
            R.av[m + lanes_below(km)] = kv;
            R.ad[m + lanes_below(km)] = kd;
        }
        m += uint32_t(__popcll(km));
    }
    wave_sync();
    unsigned long long pairs = 0;
    const uint32_t nr = prune_list<CT>(c, R.av, R.ad, m, R.bv, R.bd, p.inv_scale, &pairs);
    uint2* out = p.ext + size_t(row) * p.ext_stride;
    for (uint32_t j = lane; j < nr; j += 64) out[j] = make_uint2(R.bv[j], __float_as_uint(R.bd[j]));
    if (lane == 0) {
        p.ext_cnt[row] = nr;
        atomicAdd(p.stats + 0, ndc + pairs);
        atomicAdd(p.stats + 1, exps);
        if (W.overflow) {
            atomicAdd(p.stats + 2, 1ull);
            const uint32_t at = uint32_t(atomicAdd(p.stats + 3, 1ull));
            if (at < kRouteOverflowList) reinterpret_cast<uint32_t*>(p.stats + 4)[at] = row;
        }
    }
    wave_sync();
}

// Persistent pass over rows (or the listed rows), a wavefront per pivot.
template <int K, int MF, int CT>
__global__ __launch_bounds__(kWavesPerBlock * 64) void routing_kernel(Ctx c, RouterArgs a, RoutingPass p,
                                                                      const uint32_t* list, uint32_t n_rows,
                                                                      uint32_t* work) {
    const WaveLds w = wave_lds<K, false>(a);
    extern __shared__ __align__(16) unsigned char smem[];
    RouteLds& R = *reinterpret_cast<RouteLds*>(smem + kWavesPerBlock * wave_lds_bytes(a.D, K, false) +
                                                (threadIdx.x / 64) * sizeof(RouteLds));
    while (true) {
        uint32_t i = 0;
        if (lane_id() == 0) i = atomicAdd(work, 1u);
        i = readfirst_u(i);
        if (i >= n_rows) break;
        route_row<K, MF, CT>(c, a, p, w, R, list ? list[i] : i);
    }
}

// Rows <- first cap entries of the ext rows; with `reverse`, count the
// ReverseUpdater logs of every ext entry into its target's bucket.
__global__ void finish_kernel(Ctx c, RoutingPass p, int reverse) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(c.n) * p.ext_stride) return;
    const uint32_t row = uint32_t(i / p.ext_stride), j = uint32_t(i % p.ext_stride);
    const uint32_t nr = p.ext_cnt[row];
    if (j == 0) c.cnt[row] = min(nr, c.cap);
    if (j >= nr) return;
    const uint2 e = p.ext[i];
    if (j < c.cap) c.rows[size_t(row) * c.cap + j] = e;
    if (reverse) {
        uint32_t t = local_of(c, e.x & kMask);
        if (t == 0xFFFFFFFFu) t = c.n;  // outside the layer: the dead bucket, never merged
        atomicAdd(c.t_cnt + t, 1u);
    }
}

__global__ void reverse_scatter_kernel(Ctx c, RoutingPass p, uint2* seg) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(c.n) * p.ext_stride) return;
    const uint32_t row = uint32_t(i / p.ext_stride), j = uint32_t(i % p.ext_stride);
    if (j >= p.ext_cnt[row]) return;
    const uint2 e = p.ext[i];
    const uint32_t t = local_of(c, e.x & kMask);
    if (t == 0xFFFFFFFFu) return;
    const uint32_t pos = c.t_off[t] + atomicAdd(c.t_fill + t, 1u);
    seg[pos] = make_uint2(storage_of(c, row), e.y);
}

// PruningUpdater over every row in place, a wavefront per row.
__global__ __launch_bounds__(256) void prune_kernel(Ctx c, float inv_scale, uint32_t rb, uint32_t re) {
    __shared__ uint32_t s_av[4][kMaxCap], s_bv[4][kMaxCap];
    __shared__ float s_ad[4][kMaxCap], s_bd[4][kMaxCap];
    const uint32_t wid = threadIdx.x / 64, lane = lane_id();
    unsigned long long pairs = 0;
    for (uint32_t row = rb + blockIdx.x * 4 + wid; row < re; row += gridDim.x * 4) {
        const uint32_t n0 = c.cnt[row];
        uint2* rp = c.rows + size_t(row) * c.cap;
        for (uint32_t j = lane; j < n0; j += 64) {
            const uint2 e = rp[j];
            s_av[wid][j] = e.x;
            s_ad[wid][j] = __uint_as_float(e.y);
        }
        wave_sync();
        const uint32_t nr = prune_list<0>(c, s_av[wid], s_ad[wid], n0, s_bv[wid], s_bd[wid], inv_scale, &pairs);
        for (uint32_t j = lane; j < nr; j += 64) rp[j] = make_uint2(s_bv[wid][j], __float_as_uint(s_bd[wid][j]));
        if (lane == 0) c.cnt[row] = nr;
        wave_sync();
    }
    if (lane_id() == 0 && pairs) atomicAdd(c.stats + 1, pairs);
}

using RoutingFn = void (*)(Ctx, RouterArgs, RoutingPass, const uint32_t*, uint32_t, uint32_t*);

// L2 rows of 96, 112 or 128 floats (6, 7 or 8 chunks) get the unrolled
// chains; other widths and the DOT / cosine families the generic ones.
template <int K>
RoutingFn routing_fn(int metric, uint32_t chunks) {
    if (metric == kDot) return routing_kernel<K, 1, 0>;
    if (metric == kCosine) return routing_kernel<K, 2, 0>;
    switch (chunks) {
    case 6: return routing_kernel<K, 0, 6>;
    case 7: return routing_kernel<K, 0, 7>;
    case 8: return routing_kernel<K, 0, 8>;
    default: return routing_kernel<K, 0, 0>;
    }
}

size_t routing_lds(uint32_t D, int K) { return kWavesPerBlock * (wave_lds_bytes(D, K, false) + sizeof(RouteLds)); }

hipError_t launch_routing(const Ctx& c, const RouterArgs& a, const RoutingPass& p, int K, const uint32_t* list,
                          uint32_t n_rows, uint32_t* work, hipStream_t s, int dev_cus) {
    const RoutingFn fn = K == 2 ? routing_fn<2>(c.metric, c.C)
                                : (K == 4 ? routing_fn<4>(c.metric, c.C) : routing_fn<8>(c.metric, c.C));
    const size_t lds = routing_lds(c.D, K);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(lds));
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWavesPerBlock * 64, lds);
    if (e != hipSuccess) return e;
    const uint32_t blocks = std::max(1u, std::min(uint32_t(dev_cus * std::max(per_cu, 1)),
                                                  (n_rows + kWavesPerBlock - 1) / kWavesPerBlock));
    e = hipMemsetAsync(work, 0, sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(kWavesPerBlock * 64), lds, s, c, a, p, list, n_rows, work);
    return hipGetLastError();
}

}  // namespace

namespace artea {
namespace rf {

hipError_t routing_pass(const Ctx& c, const RoutingPass& p, hipStream_t s, int dev_cus, uint32_t* unresolved) {
    *unresolved = 0;
    if (c.n == 0) return hipSuccess;
    RouterArgs a{};
    a.vecs = c.vecs;
    a.vecs_u8 = nullptr;
    a.n = c.n_global;
    a.D = c.D;
    a.C = c.C;
    a.metric = c.metric;
    uint32_t B = 13;
    while (B < 32 && (uint64_t(1) << B) < c.n_global) ++B;
    a.hmask = uint32_t((uint64_t(1) << B) - 1);
    a.vis_mode = B <= 20 ? 0u : (B <= 27 ? 1u : 2u);
    uint32_t* work = nullptr;
    hipError_t e = hipMalloc(&work, sizeof(uint32_t));
    if (e != hipSuccess) return e;
    struct Free {
        uint32_t* p;
        ~Free() { (void)hipFree(p); }
    } guard{work};
    if ((e = hipMemsetAsync(p.stats, 0, kRouteStatsBytes, s)) != hipSuccess) return e;
    // A list of 128 entries for queues up to 96 (32 entries of tie tail), 256 above.
    const int K0 = std::max(p.topk, p.queue) <= 96 ? 2 : 4;
    if ((e = launch_routing(c, a, p, K0, nullptr, c.n, work, s, dev_cus)) != hipSuccess) return e;
    // Rows whose list dropped an entry tied with its bound: again with 512
    // entries (the whole pass when the list of them overflowed).
    unsigned long long h[4];
    if ((e = hipMemcpyAsync(h, p.stats, sizeof h, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (h[2] == 0) return hipSuccess;
    const uint32_t n_over = uint32_t(h[2]);
    const bool all = n_over > kRouteOverflowList;
    uint32_t* list = nullptr;
    if (!all) {
        if ((e = hipMalloc(&list, size_t(n_over) * 4)) != hipSuccess) return e;
        e = hipMemcpyAsync(list, reinterpret_cast<uint32_t*>(p.stats + 4), size_t(n_over) * 4,
                           hipMemcpyDeviceToDevice, s);
    }
    struct Free2 {
        uint32_t* p;
        ~Free2() {
            if (p) (void)hipFree(p);
        }
    } guard2{list};
    if (e != hipSuccess) return e;
    if ((e = hipMemsetAsync(p.stats + 2, 0, 2 * sizeof(unsigned long long), s)) != hipSuccess) return e;
    if ((e = launch_routing(c, a, p, 8, list, all ? c.n : n_over, work, s, dev_cus)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(h, p.stats, sizeof h, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    *unresolved = uint32_t(h[2]);
    return hipSuccess;
}

hipError_t routing_finish(const Ctx& c, const RoutingPass& p, bool reverse, hipStream_t s) {
    const size_t slots = size_t(c.n) * p.ext_stride;
    if (slots == 0) return hipSuccess;
    hipLaunchKernelGGL(finish_kernel, dim3(uint32_t((slots + 255) / 256)), dim3(256), 0, s, c, p, reverse ? 1 : 0);
    return hipGetLastError();
}

hipError_t routing_scatter(const Ctx& c, const RoutingPass& p, uint2* seg, hipStream_t s) {
    const size_t slots = size_t(c.n) * p.ext_stride;
    if (slots == 0) return hipSuccess;
    hipLaunchKernelGGL(reverse_scatter_kernel, dim3(uint32_t((slots + 255) / 256)), dim3(256), 0, s, c, p, seg);
    return hipGetLastError();
}

hipError_t prune_rows(const Ctx& c, float inv_scale, uint32_t rb, uint32_t re, hipStream_t s, int dev_cus) {
    re = std::min(re, c.n);
    if (rb >= re) return hipSuccess;
    const uint32_t blocks = std::max(1u, std::min((re - rb + 3) / 4, uint32_t(dev_cus) * 8u));
    hipLaunchKernelGGL(prune_kernel, dim3(blocks), dim3(256), 0, s, c, inv_scale, rb, re);
    return hipGetLastError();
}

}  // namespace rf
}  // namespace artea

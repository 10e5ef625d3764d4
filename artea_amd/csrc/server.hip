// Continuous-batching request server over a resident index: requests of
// mixed topk and queue size are submitted from the host as they arrive and
// their results are collected by polling; no call blocks on the GPU.
//
// Requests wait in one FIFO per queue-size class (the router's register
// layouts).  `lanes` streams each hold a launch's requests in flight.  A pump
// step, run by every submit and poll call, (1) harvests the launches whose
// completion event has fired, and (2) gives every lane with nothing in
// flight the class of the oldest queued request: up to that class's
// admission limit of its requests, oldest first, staged in pinned memory and
// run by that class's batched router, followed by the result copies and an
// event.
//
// Admission limits.  artea_server_create gives every class the same request
// count.  artea_server_create_budget derives them from memory: the HBM
// budget is shared by the lanes' device buffers, so a lane holds
// hbm_bytes / (lanes * per-request bytes) requests (query row, result row of
// max_topk ids and distances, topk and queue words); and a class admits at
// most lds_rounds times the wavefronts its kernel keeps resident, which its
// LDS per wavefront (candidate-list mirror, visited cache, query row) and
// registers decide.  A large-queue class, with more LDS per wavefront and
// fewer resident wavefronts, therefore launches fewer requests at a time,
// and a launch never holds more requests than lds_rounds passes of the GPU.  Every
// request is routed by its own wavefront with no state shared with the rest
// of its launch, so its result is what batch_query returns for it alone.
//
// Host copies per request: one into the pending queue (submit), one into the
// lane's pinned staging (launch), and on the way out one from a pinned
// result block to the caller (poll: its first topk slots, the rest of the
// row filled with the invalid entry).  poll_into writes each result into the
// caller's own per-request table, so a caller needs no scatter of its own.  Result blocks come from a pool of two
// per lane; a completed segment stays in its block until polled, and a lane
// waits for a free block before it launches again.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "artea_hip.h"
#include "common.hpp"

using artea::set_error;

namespace {

// Candidate-list capacities of the router's register layouts (K blocks of 64
// entries less 16 entries of tie headroom); a request runs on the smallest
// class that holds max(queue, max_topk).
constexpr uint32_t kClasses[] = {48, 112, 240, 496, 816, 1008};
constexpr int kNumClasses = 6;
constexpr int kBlocksPerLane = 2;

int class_of(uint32_t q) {
    for (int c = 0; c < kNumClasses; ++c)
        if (q <= kClasses[c]) return c;
    return -1;
}

// Requests of one queue-size class from one submit call, not yet launched
// (from `pos`).
struct Chunk {
    std::vector<float> q;
    std::vector<uint32_t> k, l;
    std::vector<uint64_t> t;
    size_t pos = 0;
};

// Pinned result rows of one lane launch, with the tickets of its rows.
struct Block {
    uint32_t* v = nullptr;
    float* d = nullptr;
    uint32_t* ovf = nullptr;  // the launch's overflow list (count, then rows)
    std::vector<uint64_t> t;
    std::vector<uint32_t> k;   // topk of each row
    uint32_t outstanding = 0;  // rows launched or completed but not yet polled
};

struct Segment {
    Block* b;
    uint32_t row0, count;
    hipEvent_t ev;
    int cls;  // queue-size class it ran in
};

struct Lane {
    hipStream_t stream = nullptr;
    artea_batched_router* routers[kNumClasses] = {};
    hipEvent_t events[kNumClasses] = {};
    float* hq = nullptr;  // pinned staging: queries, topk, queue
    uint32_t* hk = nullptr;
    uint32_t* hl = nullptr;
    float* dq = nullptr;
    uint32_t* dk = nullptr;
    uint32_t* dl = nullptr;
    uint32_t* dv = nullptr;
    float* dd = nullptr;
    std::deque<Segment> inflight;
};

}  // namespace

struct artea_server {
    artea_index* idx = nullptr;
    uint32_t dim = 0, max_topk = 0, max_queue = 0, budget = 0;  // budget: requests a lane's buffers hold
    uint32_t admit[kNumClasses] = {};                             // requests per launch, per class
    uint32_t resident[kNumClasses] = {};                          // resident wavefronts per class (budget form)
    uint64_t hbm_bytes = 0;                                       // device bytes of the lanes' buffers
    int lo_class = 0, hi_class = 0;
    std::vector<Lane> lanes;
    std::vector<Block> blocks;
    std::vector<Block*> free_blocks;
    std::deque<Chunk> pending[kNumClasses];  // per queue-size class, oldest first
    uint64_t queued_class[kNumClasses] = {};
    std::vector<Chunk> spare;   // consumed chunks kept for their storage (no fresh pages per submit)
    std::deque<Segment> ready;  // completed, in harvest order; row0/count shrink as rows are polled
    uint64_t next_ticket = 0, queued = 0, in_flight = 0, ready_rows = 0;
    uint64_t rerun = 0, unresolved = 0;  // tie-tail overflows re-run on a larger list / left as they were
};

namespace {

void lane_free(Lane& ln) {
    if (ln.stream) (void)hipStreamSynchronize(ln.stream);
    for (int c = 0; c < kNumClasses; ++c) {
        if (ln.routers[c]) artea_batched_router_destroy(ln.routers[c]);
        if (ln.events[c]) (void)hipEventDestroy(ln.events[c]);
    }
    for (void* p : {(void*)ln.hq, (void*)ln.hk, (void*)ln.hl})
        if (p) (void)hipHostFree(p);
    for (void* p : {(void*)ln.dq, (void*)ln.dk, (void*)ln.dl, (void*)ln.dv, (void*)ln.dd})
        if (p) (void)hipFree(p);
    if (ln.stream) (void)hipStreamDestroy(ln.stream);
    ln = Lane{};
}

void server_free(artea_server* s) {
    for (Lane& ln : s->lanes) lane_free(ln);  // synchronises every stream first
    for (Block& b : s->blocks) {
        if (b.v) (void)hipHostFree(b.v);
        if (b.d) (void)hipHostFree(b.d);
        if (b.ovf) (void)hipHostFree(b.ovf);
    }
    delete s;
}

template <class T>
int host_alloc(T** p, size_t count) {
    ARTEA_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)));
    return ARTEA_OK;
}

template <class T>
int dev_alloc(T** p, size_t count) {
    ARTEA_HIP_TRY(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)));
    return ARTEA_OK;
}

int lane_init(artea_server* s, Lane& ln) {
    ARTEA_HIP_TRY(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
    for (int c = s->lo_class; c <= s->hi_class; ++c) {
        int rc = artea_batched_router_create(s->idx, s->max_topk, kClasses[c], &ln.routers[c]);
        if (rc) return rc;
        ARTEA_HIP_TRY(hipEventCreateWithFlags(&ln.events[c], hipEventDisableTiming));
    }
    const size_t B = s->budget, D = s->dim, K = s->max_topk;
    int rc = 0;
    if ((rc = host_alloc(&ln.hq, B * D)) || (rc = host_alloc(&ln.hk, B)) || (rc = host_alloc(&ln.hl, B)) ||
        (rc = dev_alloc(&ln.dq, B * D)) || (rc = dev_alloc(&ln.dk, B)) || (rc = dev_alloc(&ln.dl, B)) ||
        (rc = dev_alloc(&ln.dv, B * K)) || (rc = dev_alloc(&ln.dd, B * K)))
        return rc;
    return ARTEA_OK;
}

// Requests of a completed segment whose candidate list overflowed its tie
// tail (artea_batched_router_last_overflows) run again, synchronously, on
// the next larger list class until none overflows, and their rows of the
// block are replaced; one still overflowing in the largest class is counted
// as unresolved.  The lane's staging still holds the segment's requests: a
// lane launches again only after its segment was harvested.
int rerun_overflows(artea_server* s, Lane& ln, const Segment& seg) {
    const uint32_t listed = seg.b->ovf[0];
    std::vector<uint32_t> rows;
    if (listed <= artea::rk_overflow_list()) {
        rows.assign(seg.b->ovf + 1, seg.b->ovf + 1 + listed);
    } else {
        for (uint32_t r = 0; r < seg.count; ++r) rows.push_back(seg.row0 + r);
    }
    std::sort(rows.begin(), rows.end());
    s->rerun += rows.size();
    const size_t D = s->dim, K = s->max_topk;
    std::vector<float> q;
    std::vector<uint32_t> k, l, ov, list(artea::rk_overflow_list());
    std::vector<float> od;
    for (int c = seg.cls + 1; !rows.empty(); ++c) {
        if (c >= kNumClasses) {
            s->unresolved += rows.size();
            break;
        }
        if (!ln.routers[c]) {
            int rc = artea_batched_router_create(s->idx, s->max_topk, kClasses[c], &ln.routers[c]);
            if (rc) return rc;
        }
        const size_t m = rows.size();
        q.resize(m * D);
        k.resize(m);
        l.resize(m);
        for (size_t i = 0; i < m; ++i) {
            std::memcpy(q.data() + i * D, ln.hq + size_t(rows[i]) * D, D * sizeof(float));
            k[i] = ln.hk[rows[i]];
            l[i] = ln.hl[rows[i]];
        }
        ARTEA_HIP_TRY(hipMemcpyAsync(ln.dq, q.data(), m * D * sizeof(float), hipMemcpyHostToDevice, ln.stream));
        ARTEA_HIP_TRY(hipMemcpyAsync(ln.dk, k.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice, ln.stream));
        ARTEA_HIP_TRY(hipMemcpyAsync(ln.dl, l.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice, ln.stream));
        int rc = artea_batched_router_query_device(ln.routers[c], ln.dq, uint32_t(m), ln.dk, ln.dl, 0, nullptr, ln.dv,
                                                   ln.dd, ln.stream);
        if (rc) return rc;
        ov.resize(m * K);
        od.resize(m * K);
        ARTEA_HIP_TRY(hipMemcpyAsync(ov.data(), ln.dv, m * K * sizeof(uint32_t), hipMemcpyDeviceToHost, ln.stream));
        ARTEA_HIP_TRY(hipMemcpyAsync(od.data(), ln.dd, m * K * sizeof(float), hipMemcpyDeviceToHost, ln.stream));
        ARTEA_HIP_TRY(hipStreamSynchronize(ln.stream));
        for (size_t i = 0; i < m; ++i) {
            std::memcpy(seg.b->v + size_t(rows[i]) * K, ov.data() + i * K, K * sizeof(uint32_t));
            std::memcpy(seg.b->d + size_t(rows[i]) * K, od.data() + i * K, K * sizeof(float));
        }
        uint32_t again = 0;
        rc = artea_batched_router_last_overflows(ln.routers[c], uint32_t(list.size()), list.data(), &again);
        if (rc) return rc;
        std::vector<uint32_t> next;
        if (again > list.size()) {
            next = rows;
        } else {
            for (uint32_t i = 0; i < again; ++i) next.push_back(rows[list[i]]);
        }
        rows.swap(next);
    }
    return ARTEA_OK;
}

// Moves the completed launches of one lane to the ready queue.
int harvest(artea_server* s, Lane& ln) {
    while (!ln.inflight.empty()) {
        const Segment seg = ln.inflight.front();
        const hipError_t e = hipEventQuery(seg.ev);
        if (e == hipErrorNotReady) break;
        ARTEA_HIP_TRY(e);
        if (seg.b->ovf[0]) {
            int rc = rerun_overflows(s, ln, seg);
            if (rc) return rc;
        }
        s->ready.push_back(seg);
        s->in_flight -= seg.count;
        s->ready_rows += seg.count;
        ln.inflight.pop_front();
    }
    return ARTEA_OK;
}

// Launches queued requests of one queue-size class on an idle lane: the class
// of the oldest queued request, up to the budget of its requests, oldest
// first.  A launch of short searches then never waits behind long ones, and
// no request is passed over by a later one of its own class.
int launch(artea_server* s, Lane& ln) {
    if (s->queued == 0 || s->free_blocks.empty()) return ARTEA_OK;
    int c = -1;
    uint64_t oldest = ~0ull;
    for (int cc = 0; cc < kNumClasses; ++cc) {
        if (!s->queued_class[cc]) continue;
        const Chunk& f = s->pending[cc].front();
        if (f.t[f.pos] < oldest) {
            oldest = f.t[f.pos];
            c = cc;
        }
    }
    const uint32_t m = uint32_t(std::min<uint64_t>(s->queued_class[c], s->admit[c]));
    const size_t D = s->dim, K = s->max_topk;
    Block* b = s->free_blocks.back();
    s->free_blocks.pop_back();
    // The class's chunks from the front hold at least m requests; each
    // chunk's rows are consecutive, so they are staged in block copies.
    std::deque<Chunk>& q = s->pending[c];
    for (uint32_t i = 0; i < m;) {
        Chunk& ch = q.front();
        const size_t take = std::min<size_t>(ch.k.size() - ch.pos, m - i);
        std::memcpy(ln.hq + size_t(i) * D, ch.q.data() + ch.pos * D, take * D * sizeof(float));
        std::memcpy(ln.hk + i, ch.k.data() + ch.pos, take * sizeof(uint32_t));
        std::memcpy(ln.hl + i, ch.l.data() + ch.pos, take * sizeof(uint32_t));
        std::memcpy(b->t.data() + i, ch.t.data() + ch.pos, take * sizeof(uint64_t));
        std::memcpy(b->k.data() + i, ch.k.data() + ch.pos, take * sizeof(uint32_t));
        ch.pos += take;
        i += uint32_t(take);
        if (ch.pos == ch.k.size()) {
            if (s->spare.size() < 8) s->spare.push_back(std::move(ch));
            q.pop_front();
        }
    }
    s->queued_class[c] -= m;
    s->queued -= m;
    b->outstanding = m;
    ARTEA_HIP_TRY(hipMemcpyAsync(ln.dq, ln.hq, m * D * sizeof(float), hipMemcpyHostToDevice, ln.stream));
    ARTEA_HIP_TRY(hipMemcpyAsync(ln.dk, ln.hk, m * sizeof(uint32_t), hipMemcpyHostToDevice, ln.stream));
    ARTEA_HIP_TRY(hipMemcpyAsync(ln.dl, ln.hl, m * sizeof(uint32_t), hipMemcpyHostToDevice, ln.stream));
    int rc = artea_batched_router_query_device(ln.routers[c], ln.dq, m, ln.dk, ln.dl, 0, nullptr, ln.dv, ln.dd,
                                               ln.stream);
    if (rc) return rc;
    ARTEA_HIP_TRY(hipMemcpyAsync(b->v, ln.dv, m * K * sizeof(uint32_t), hipMemcpyDeviceToHost, ln.stream));
    ARTEA_HIP_TRY(hipMemcpyAsync(b->d, ln.dd, m * K * sizeof(float), hipMemcpyDeviceToHost, ln.stream));
    rc = artea::router_overflow_list_async(ln.routers[c], b->ovf, ln.stream);
    if (rc) return rc;
    ARTEA_HIP_TRY(hipEventRecord(ln.events[c], ln.stream));
    ln.inflight.push_back(Segment{b, 0, m, ln.events[c], c});
    s->in_flight += m;
    return ARTEA_OK;
}

int pump(artea_server* s) {
    for (Lane& ln : s->lanes) {
        int rc = harvest(s, ln);
        if (rc) return rc;
        if (ln.inflight.empty()) {
            rc = launch(s, ln);
            if (rc) return rc;
        }
    }
    return ARTEA_OK;
}

// Device bytes one request holds in a lane: its query row, its result row
// (max_topk ids and distances) and its topk and queue words.
uint64_t request_bytes(uint32_t dim_pad, uint32_t max_topk) {
    return uint64_t(dim_pad) * 4 + uint64_t(max_topk) * 8 + 8;
}

// lds_rounds == 0: every class admits `budget` requests per launch; else a
// class admits min(budget, lds_rounds * its resident wavefronts).
int server_create(artea_index* idx, uint32_t max_topk, uint32_t max_queue, uint32_t lanes, uint32_t budget,
                  uint32_t lds_rounds, artea_server** out) {
    ARTEA_CHECK_ARG(idx && out, "artea_server_create: null argument");
    ARTEA_CHECK_ARG(max_topk >= 1 && max_topk <= max_queue, "artea_server_create: need 1 <= max_topk <= max_queue");
    ARTEA_CHECK_ARG(class_of(max_queue) >= 0, "artea_server_create: max_queue exceeds the router's largest queue");
    ARTEA_CHECK_ARG(lanes >= 1 && lanes <= 16, "artea_server_create: lanes must be in [1, 16]");
    artea_server* s = new artea_server();
    s->idx = idx;
    s->dim = idx->dim_pad;
    s->max_topk = max_topk;
    s->max_queue = max_queue;
    s->budget = budget;
    s->hbm_bytes = uint64_t(lanes) * budget * request_bytes(idx->dim_pad, max_topk);
    s->lo_class = class_of(max_topk);
    s->hi_class = class_of(max_queue);
    s->lanes.resize(lanes);
    s->blocks.resize(size_t(lanes) * kBlocksPerLane);
    int rc = 0;
    for (Lane& ln : s->lanes)
        if (!rc) rc = lane_init(s, ln);
    for (Block& b : s->blocks) {
        if (!rc) rc = host_alloc(&b.v, size_t(budget) * max_topk);
        if (!rc) rc = host_alloc(&b.d, size_t(budget) * max_topk);
        if (!rc) rc = host_alloc(&b.ovf, size_t(1) + artea::rk_overflow_list());
        b.t.assign(budget, 0);
        b.k.assign(budget, 0);
        s->free_blocks.push_back(&b);
    }
    for (int c = 0; c < kNumClasses; ++c) s->admit[c] = budget;
    if (!rc && lds_rounds) {
        for (int c = s->lo_class; c <= s->hi_class && !rc; ++c) {
            rc = artea::batched_router_resident_waves(s->lanes[0].routers[c], &s->resident[c]);
            if (!rc) s->admit[c] = uint32_t(std::min<uint64_t>(budget, uint64_t(lds_rounds) * s->resident[c]));
        }
    }
    if (rc) {
        server_free(s);
        return rc;
    }
    *out = s;
    return ARTEA_OK;
}

}  // namespace

extern "C" {

int artea_server_create(artea_index* idx, uint32_t max_topk, uint32_t max_queue, uint32_t lanes, uint32_t budget,
                        artea_server** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(budget >= 1 && budget <= (1u << 20), "artea_server_create: budget must be in [1, 2^20]");
    return server_create(idx, max_topk, max_queue, lanes, budget, 0, out);
}

int artea_server_create_budget(artea_index* idx, uint32_t max_topk, uint32_t max_queue, uint32_t lanes,
                               uint64_t hbm_bytes, uint32_t lds_rounds, artea_server** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(idx && out, "artea_server_create_budget: null argument");
    ARTEA_CHECK_ARG(lanes >= 1 && lanes <= 16, "artea_server_create_budget: lanes must be in [1, 16]");
    ARTEA_CHECK_ARG(lds_rounds >= 1 && lds_rounds <= 1024, "artea_server_create_budget: lds_rounds must be in [1, 1024]");
    ARTEA_CHECK_ARG(max_topk >= 1 && max_topk <= max_queue,
                    "artea_server_create_budget: need 1 <= max_topk <= max_queue");
    const uint64_t per_request = request_bytes(idx->dim_pad, max_topk);
    const uint64_t cap = hbm_bytes / (uint64_t(lanes) * per_request);
    if (cap < 1) {
        set_error("artea_server_create_budget: " + std::to_string(hbm_bytes) + " bytes hold less than one request per lane (" +
                  std::to_string(per_request) + " bytes each, " + std::to_string(lanes) + " lanes)");
        return ARTEA_EINVAL;
    }
    return server_create(idx, max_topk, max_queue, lanes, uint32_t(std::min<uint64_t>(cap, 1u << 20)), lds_rounds,
                         out);
}

int artea_server_admission(artea_server* s, uint32_t* admit, uint32_t* resident, uint32_t* lane_requests,
                           uint64_t* hbm_bytes) {
    artea::clear_error();
    ARTEA_CHECK_ARG(s, "artea_server_admission: null server");
    for (int c = 0; c < kNumClasses; ++c) {
        if (admit) admit[c] = s->admit[c];
        if (resident) resident[c] = s->resident[c];
    }
    if (lane_requests) *lane_requests = s->budget;
    if (hbm_bytes) *hbm_bytes = s->hbm_bytes;
    return ARTEA_OK;
}

int artea_server_destroy(artea_server* s) {
    if (s) server_free(s);
    return ARTEA_OK;
}

int artea_server_submit(artea_server* s, const float* queries, uint32_t n, const uint32_t* topk,
                        const uint32_t* queue, uint64_t* first_ticket) {
    artea::clear_error();
    ARTEA_CHECK_ARG(s, "artea_server_submit: null server");
    ARTEA_CHECK_ARG(n == 0 || (queries && topk && queue), "artea_server_submit: null buffer");
    for (uint32_t i = 0; i < n; ++i) {
        if (topk[i] < 1 || topk[i] > s->max_topk || queue[i] > s->max_queue) {
            set_error("artea_server_submit: request " + std::to_string(i) + " has topk " + std::to_string(topk[i]) +
                      ", queue " + std::to_string(queue[i]) + " (server: topk 1.." + std::to_string(s->max_topk) +
                      ", queue <= " + std::to_string(s->max_queue) + ")");
            return ARTEA_EINVAL;
        }
        if (queue[i] < topk[i]) {
            set_error("candidate_queue_size (" + std::to_string(queue[i]) + ") must be >= topk (" +
                      std::to_string(topk[i]) + ")");
            return ARTEA_EINVAL;
        }
    }
    if (first_ticket) *first_ticket = s->next_ticket;
    if (n) {
        const size_t D = s->dim;
        std::vector<uint8_t> cls(n);
        uint32_t cnt[kNumClasses] = {};
        for (uint32_t i = 0; i < n; ++i) {
            cls[i] = uint8_t(class_of(std::max(queue[i], s->max_topk)));
            ++cnt[cls[i]];
        }
        Chunk* dst[kNumClasses] = {};
        for (int c = 0; c < kNumClasses; ++c) {
            if (!cnt[c]) continue;
            Chunk ch;
            if (!s->spare.empty()) {
                ch = std::move(s->spare.back());
                s->spare.pop_back();
            }
            ch.pos = 0;
            ch.q.clear();
            ch.k.clear();
            ch.l.clear();
            ch.t.clear();
            ch.q.reserve(size_t(cnt[c]) * D);
            s->pending[c].push_back(std::move(ch));
            dst[c] = &s->pending[c].back();
            s->queued_class[c] += cnt[c];
        }
        for (uint32_t i = 0; i < n;) {
            // runs of consecutive requests of one class are appended at once
            uint32_t j = i + 1;
            while (j < n && cls[j] == cls[i]) ++j;
            Chunk& ch = *dst[cls[i]];
            ch.q.insert(ch.q.end(), queries + size_t(i) * D, queries + size_t(j) * D);
            ch.k.insert(ch.k.end(), topk + i, topk + j);
            ch.l.insert(ch.l.end(), queue + i, queue + j);
            for (uint32_t r = i; r < j; ++r) ch.t.push_back(s->next_ticket + r);
            i = j;
        }
        s->next_ticket += n;
        s->queued += n;
    }
    return pump(s);
}

}  // extern "C"

namespace {

// Copies row r of a completed block to (v, d): the router's first topk
// slots, then the invalid entry up to max_topk.
inline void copy_row(const Block* b, uint32_t r, size_t K, uint32_t* v, float* d) {
    const size_t k = std::min<size_t>(b->k[r], K);
    std::memcpy(v, b->v + r * K, k * sizeof(uint32_t));
    std::memcpy(d, b->d + r * K, k * sizeof(float));
    std::fill(v + k, v + K, ARTEA_INVALID_VID);
    std::fill(d + k, d + K, FLT_MAX);
}

// Common body of both polls.  table_rows == 0: compact output (row i of the
// outputs = i-th returned request); else row (ticket - base) of the tables.
int poll_common(artea_server* s, uint32_t max_results, uint64_t* tickets, uint32_t* out_vids, float* out_dists,
                uint32_t* n_out, uint64_t base, uint64_t table_rows) {
    *n_out = 0;
    int rc = pump(s);
    if (rc) return rc;
    const size_t K = s->max_topk;
    uint32_t got = 0;
    while (got < max_results && !s->ready.empty()) {
        Segment& seg = s->ready.front();
        const uint32_t m = std::min(seg.count, max_results - got);
        if (table_rows) {
            for (uint32_t i = 0; i < m; ++i) {
                const uint64_t t = seg.b->t[seg.row0 + i];
                if (t < base || t - base >= table_rows) {
                    set_error("artea_server_poll_into: ticket " + std::to_string(t) + " is outside the table");
                    *n_out = got;
                    return ARTEA_EINVAL;
                }
            }
        }
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t r = seg.row0 + i;
            const uint64_t t = seg.b->t[r];
            const size_t row = table_rows ? size_t(t - base) : size_t(got + i);
            copy_row(seg.b, r, K, out_vids + row * K, out_dists + row * K);
            tickets[got + i] = t;
        }
        got += m;
        seg.row0 += m;
        seg.count -= m;
        s->ready_rows -= m;
        seg.b->outstanding -= m;
        if (seg.b->outstanding == 0) s->free_blocks.push_back(seg.b);
        if (seg.count == 0) s->ready.pop_front();
    }
    *n_out = got;
    // a freed block may let an idle lane launch
    if (got) rc = pump(s);
    return rc;
}

}  // namespace

extern "C" {

int artea_server_poll(artea_server* s, uint32_t max_results, uint64_t* tickets, uint32_t* out_vids,
                      float* out_dists, uint32_t* n_out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(s && n_out, "artea_server_poll: null argument");
    ARTEA_CHECK_ARG(max_results == 0 || (tickets && out_vids && out_dists), "artea_server_poll: null buffer");
    return poll_common(s, max_results, tickets, out_vids, out_dists, n_out, 0, 0);
}

int artea_server_poll_into(artea_server* s, uint64_t base_ticket, uint64_t table_rows, uint32_t* table_vids,
                           float* table_dists, uint32_t max_results, uint64_t* tickets, uint32_t* n_out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(s && n_out, "artea_server_poll_into: null argument");
    ARTEA_CHECK_ARG(table_rows > 0 && table_vids && table_dists, "artea_server_poll_into: empty table");
    ARTEA_CHECK_ARG(max_results == 0 || tickets, "artea_server_poll_into: null buffer");
    return poll_common(s, max_results, tickets, table_vids, table_dists, n_out, base_ticket, table_rows);
}

int artea_server_overflows(artea_server* s, uint64_t* rerun, uint64_t* unresolved) {
    artea::clear_error();
    ARTEA_CHECK_ARG(s, "artea_server_overflows: null server");
    if (rerun) *rerun = s->rerun;
    if (unresolved) *unresolved = s->unresolved;
    return ARTEA_OK;
}

int artea_server_status(artea_server* s, uint64_t* queued, uint64_t* in_flight, uint64_t* ready) {
    artea::clear_error();
    ARTEA_CHECK_ARG(s, "artea_server_status: null server");
    if (queued) *queued = s->queued;
    if (in_flight) *in_flight = s->in_flight;
    if (ready) *ready = s->ready_rows;
    return ARTEA_OK;
}

}  // extern "C"

// HGRA v1 on the device (persistence/hierarchical_graph_file_manager.hpp:
// :20-40 wire format, :88-145 snapshot, :152-324 restore).  The loader parses
// the file's body in HBM straight into the router's ELL layout; the writer
// lays the records out in HBM from that layout and streams them to the file.
// The header checks and the failure texts are the host reader's (hgra.cpp):
// a file is accepted or rejected exactly as HierarchicalGraphFileManager::
// restore does, with the same text.
//
// Loader.  A record's length follows from its own words (highest level, then
// per level a count and that many ids), so the records are found by a
// segmented parse.  The body is cut into segments of S words.  Each segment's
// first record start is guessed: the first word from which `trial` records
// parse.  Each segment is parsed from its guess up to the first record start
// at or past its end (its exit).  The host then walks the segments in order
// and takes segment j's parse only when its guess equals segment j-1's exit;
// otherwise segment j is parsed again from that exit.  Segment 0 starts at
// word 0, so the walk is the sequential parse whatever the guesses were.  A
// parse stops at the first structural fault (truncation, highest level above
// top, count above capacity).  The ids are checked, and the rows copied, per
// vertex with one wavefront each; an out-of-range id in an earlier record, or
// in an earlier level of the faulting record, is reported before the fault,
// which is the order of the sequential reader.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"

using namespace artea;
using namespace artea::dev;

namespace {

constexpr uint32_t kMagic = 0x48475241u;  // "HGRA"
constexpr uint32_t kVersion = 1u;
constexpr uint64_t kNone = ~0ull;
constexpr uint32_t kDefaultSegment = 4096;  // words per segment
constexpr uint32_t kDefaultTrial = 4;       // records a guessed start must parse
constexpr size_t kChunk = size_t(64) << 20; // bytes per host <-> device staging copy

struct Body {
    const uint32_t* w;
    uint64_t W;  // words
    uint32_t n, top, ul, bl;
};

struct SegResult {
    uint64_t exit;   // first record start at or past the segment's end; kNone after a fault
    uint64_t at;     // start of the faulting record
    uint32_t count;  // records parsed (started in the segment) before a fault
    uint32_t fault;  // 0 or the HgraFault of record `count`
    uint32_t level, a, b, pad;
};

// The structure of the record at word p (ids not checked): its end in *next,
// or the fault the sequential reader meets first.
__device__ uint32_t parse_record(const Body& B, uint64_t p, uint64_t* next, uint32_t* level, uint32_t* a,
                                 uint32_t* b) {
    if (p >= B.W) return kHgraTruncVid;
    const uint32_t h = B.w[p++];
    if (h == kInvalidLevel) {
        *next = p;
        return 0;
    }
    if (h > B.top) {
        *a = h;
        return kHgraHighest;
    }
    for (uint32_t l = h;; --l) {
        if (p >= B.W) {
            *level = l;
            return kHgraTruncLevel;
        }
        const uint32_t c = B.w[p++];
        const uint32_t cap = l == 0 ? B.bl : B.ul;
        if (c > cap) {
            *level = l;
            *a = c;
            *b = cap;
            return kHgraCapacity;
        }
        if (uint64_t(c) > B.W - p) {
            *level = l;
            return kHgraTruncNbrs;
        }
        p += c;
        if (l == 0) break;
    }
    *next = p;
    return 0;
}

// Guessed first record start of every segment (segment 0: word 0).
__global__ void guess_kernel(Body B, uint64_t S, uint32_t G, uint32_t trial, uint64_t* start) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= G) return;
    const uint64_t s = uint64_t(j) * S, e = min(s + S, B.W);
    uint64_t found = j == 0 ? 0 : kNone;
    for (uint64_t p = s; p < e && found == kNone; ++p) {
        uint64_t q = p;
        bool ok = true;
        for (uint32_t t = 0; t < trial && q < B.W; ++t) {
            uint64_t nx = 0;
            uint32_t lv, a, b;
            if (parse_record(B, q, &nx, &lv, &a, &b)) {
                ok = false;
                break;
            }
            q = nx;
        }
        if (ok) found = p;
    }
    start[j] = found;
}

// Segments first..first+count parsed from their starts.
__global__ void parse_kernel(Body B, uint64_t S, uint32_t first, uint32_t count, const uint64_t* start,
                             SegResult* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t j = first + i;
    const uint64_t e = min(uint64_t(j) * S + S, B.W);
    SegResult r{kNone, kNone, 0, 0, 0, 0, 0, 0};
    uint64_t p = start[j];
    if (p != kNone) {
        while (p < e) {
            uint64_t nx = 0;
            const uint32_t f = parse_record(B, p, &nx, &r.level, &r.a, &r.b);
            if (f) {
                r.fault = f;
                r.at = p;
                break;
            }
            ++r.count;
            p = nx;
        }
        if (!r.fault) r.exit = p;
    }
    out[j] = r;
}

// Segment j parsed from `pos`, its verified start, and the segments after it
// from their verified starts (the previous one's exit), until a segment's
// guess equals that exit (its parse stands), a fault, or the end; *upto = the
// first segment left as it was.  Segments in which no record starts get start
// kNone.  One thread.
__global__ void repair_kernel(Body B, uint64_t S, uint32_t G, uint32_t j, uint64_t pos, uint64_t* start,
                              SegResult* res, uint32_t* upto) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t k = j;
    while (true) {
        const uint64_t e = min(uint64_t(k) * S + S, B.W);
        start[k] = pos;
        SegResult r{kNone, kNone, 0, 0, 0, 0, 0, 0};
        uint64_t p = pos;
        while (p < e) {
            uint64_t nx = 0;
            const uint32_t f = parse_record(B, p, &nx, &r.level, &r.a, &r.b);
            if (f) {
                r.fault = f;
                r.at = p;
                break;
            }
            ++r.count;
            p = nx;
        }
        if (!r.fault) r.exit = p;
        res[k] = r;
        ++k;
        if (r.fault) break;
        pos = p;
        while (k < G && pos >= min(uint64_t(k) * S + S, B.W)) start[k++] = kNone;
        if (k >= G || start[k] == pos) break;
    }
    *upto = k;
}

// Record start of every vertex below V, from the verified segment starts.
__global__ void offsets_kernel(Body B, uint64_t S, uint32_t G, const uint64_t* start, const uint64_t* base,
                               uint64_t V, uint64_t* rec) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= G) return;
    uint64_t p = start[j], v = base[j];
    if (p == kNone) return;
    const uint64_t e = min(uint64_t(j) * S + S, B.W);
    while (p < e && v < V) {
        rec[v++] = p;
        uint64_t nx = 0;
        uint32_t lv, a, b;
        if (parse_record(B, p, &nx, &lv, &a, &b)) break;  // not reached below V
        p = nx;
    }
}

// Per vertex (one wavefront each): its highest level and number of upper
// rows, the largest L0 and upper-level counts, and the first row (in vid
// order, then file order) holding an id >= n, as vid << 32 | level step.
__global__ __launch_bounds__(256) void check_kernel(Body B, const uint64_t* rec, uint32_t V, uint32_t* hi,
                                                    uint64_t* upper, uint32_t* maxc, unsigned long long* bad) {
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    uint32_t m0 = 0, mu = 0;
    for (uint32_t v = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v < V; v += waves) {
        uint64_t p = rec[v];
        const uint32_t h = B.w[p++];
        if (lane == 0) {
            hi[v] = h;
            upper[v] = h == kInvalidLevel ? 0 : h;
        }
        if (h == kInvalidLevel) continue;
        for (uint32_t l = h;; --l) {
            const uint32_t c = B.w[p++];
            if (l == 0) m0 = max(m0, c);
            else mu = max(mu, c);
            bool hit = false;
            for (uint32_t k = lane; k < c && !hit; k += 64) hit = B.w[p + k] >= B.n;
            if (ballot(hit)) {
                if (lane == 0) atomicMin(bad, (static_cast<unsigned long long>(v) << 32) | (h - l));
                break;
            }
            p += c;
            if (l == 0) break;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        m0 = max(m0, uint32_t(__shfl_xor(int(m0), off)));
        mu = max(mu, uint32_t(__shfl_xor(int(mu), off)));
    }
    if (lane == 0) {
        atomicMax(maxc + 0, m0);
        atomicMax(maxc + 1, mu);
    }
}

// The first id >= n in level steps [0, steps) of the record at p (steps =
// kNone: every level), in file order; kNone when there is none.
__global__ void first_bad_kernel(Body B, uint64_t p, uint32_t steps, uint64_t* out) {
    const uint32_t lane = lane_id();
    const uint32_t h = B.w[p++];
    uint64_t found = kNone;
    if (h != kInvalidLevel) {
        for (uint32_t s = 0; s <= h && s < steps && found == kNone; ++s) {
            const uint32_t c = B.w[p++];
            for (uint32_t k0 = 0; k0 < c && found == kNone; k0 += 64) {
                const uint32_t id = k0 + lane < c ? B.w[p + k0 + lane] : 0u;
                const uint64_t m = ballot(k0 + lane < c && id >= B.n);
                if (m) found = readlane_u(id, __ffsll((unsigned long long)m) - 1);
            }
            p += c;
        }
    }
    if (lane == 0) *out = found;
}

// ELL rows and vertex info of every vertex, one wavefront each.
__global__ __launch_bounds__(256) void fill_kernel(Body B, const uint64_t* rec, const uint32_t* hi,
                                                   const uint64_t* ub, uint32_t n, uint32_t* ell0, uint32_t stride0,
                                                   uint32_t* ellU, uint32_t strideU, uint2* vinfo) {
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t v = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v < n; v += waves) {
        const uint32_t h = hi[v];
        uint32_t* r0 = ell0 + size_t(v) * stride0;
        if (h == kInvalidLevel) {
            for (uint32_t k = lane; k < stride0; k += 64) r0[k] = kNoVertex;
            if (lane == 0) vinfo[v] = make_uint2(kNoVertex, 0);
            continue;
        }
        if (lane == 0) vinfo[v] = make_uint2(h, uint32_t(ub[v]));
        uint64_t p = rec[v] + 1;
        for (uint32_t l = h;; --l) {
            const uint32_t c = B.w[p++];
            uint32_t* dst = l == 0 ? r0 : ellU + size_t(ub[v] + (h - l)) * strideU;
            const uint32_t stride = l == 0 ? stride0 : strideU;
            for (uint32_t k = lane; k < stride; k += 64) dst[k] = k < c ? B.w[p + k] : kNoVertex;
            p += c;
            if (l == 0) break;
        }
    }
}

// Writer: valid entries of an ELL row (they are contiguous from its start).
__device__ uint32_t row_count(const uint32_t* row, uint32_t stride) {
    uint32_t c = 0;
    for (uint32_t k0 = 0; k0 < stride; k0 += 64) {
        const uint32_t k = k0 + lane_id();
        const uint32_t m = uint32_t(__popcll(ballot(k < stride && row[k] != kNoVertex)));
        c += m;
        if (m < 64) break;
    }
    return c;
}

struct Ell {
    const uint32_t* ell0;
    const uint32_t* ellU;
    const uint2* vinfo;
    uint32_t stride0, strideU;
};

__device__ const uint32_t* ell_row(const Ell& g, uint32_t v, uint32_t h, uint32_t ub, uint32_t l, uint32_t* stride) {
    *stride = l == 0 ? g.stride0 : g.strideU;
    return l == 0 ? g.ell0 + size_t(v) * g.stride0 : g.ellU + size_t(ub + (h - l)) * g.strideU;
}

// Words of every vertex's record; len[n] = 0 closes the scan.
__global__ __launch_bounds__(256) void length_kernel(Ell g, uint32_t n, uint64_t* len) {
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t v = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v <= n; v += waves) {
        uint64_t words = 0;
        if (v < n) {
            const uint2 vi = g.vinfo[v];
            words = 1;
            if (vi.x != kInvalidLevel) {
                for (uint32_t l = vi.x;; --l) {
                    uint32_t stride;
                    const uint32_t* row = ell_row(g, v, vi.x, vi.y, l, &stride);
                    words += 1 + row_count(row, stride);
                    if (l == 0) break;
                }
            }
        }
        if (lane == 0) len[v] = words;
    }
}

// Every vertex's record at its offset: highest level, then per level from
// the highest down its count and ids (file_manager.hpp:88-145).
__global__ __launch_bounds__(256) void record_kernel(Ell g, uint32_t n, const uint64_t* off, uint32_t* out) {
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t v = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v < n; v += waves) {
        const uint2 vi = g.vinfo[v];
        uint64_t p = off[v];
        if (lane == 0) out[p] = vi.x;
        ++p;
        if (vi.x == kInvalidLevel) continue;
        for (uint32_t l = vi.x;; --l) {
            uint32_t stride;
            const uint32_t* row = ell_row(g, v, vi.x, vi.y, l, &stride);
            const uint32_t c = row_count(row, stride);
            if (lane == 0) out[p] = c;
            for (uint32_t k = lane; k < c; k += 64) out[p + 1 + k] = row[k];
            p += 1 + c;
            if (l == 0) break;
        }
    }
}

// Device buffers of one call, freed on every path.
struct Scratch {
    std::vector<void*> d;
    std::vector<void*> h;
    std::vector<hipEvent_t> ev;
    ~Scratch() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (void* p : d) (void)hipFree(p);
        for (void* p : h) (void)hipHostFree(p);
    }
    template <class T>
    hipError_t dev(T** p, size_t count) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) d.push_back(q);
        *p = static_cast<T*>(q);
        return e;
    }
    hipError_t host(void** p, size_t bytes) {
        const hipError_t e = hipHostMalloc(p, bytes);
        if (e == hipSuccess) h.push_back(*p);
        return e;
    }
    hipError_t event(hipEvent_t* e) {
        const hipError_t r = hipEventCreateWithFlags(e, hipEventDisableTiming);
        if (r == hipSuccess) ev.push_back(*e);
        return r;
    }
};

#define HG_TRY(expr)                                                                                   \
    do {                                                                                               \
        const hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                        \
            set_error(std::string("HGRA device path: ") + #expr + ": " + hipGetErrorString(e_));       \
            return ARTEA_EHIP;                                                                         \
        }                                                                                              \
    } while (0)

uint32_t grid_waves(uint64_t items, uint32_t per_block_waves = 4) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint64_t want = (items + per_block_waves - 1) / per_block_waves;
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(want, uint64_t(cus) * 16)));
}

// The body (file bytes 28 .. 28 + 4 W) into d_body, through two pinned staging
// buffers so the file reads overlap the copies.
int upload_body(FILE* f, uint64_t W, uint32_t* d_body, hipStream_t st, Scratch& S, const std::string& path) {
    const uint64_t bytes = W * 4;
    if (bytes == 0) return ARTEA_OK;
    void* pin[2];
    hipEvent_t done[2];
    const size_t cap = size_t(std::min<uint64_t>(bytes, kChunk));
    for (int i = 0; i < 2; ++i) {
        HG_TRY(S.host(&pin[i], cap));
        HG_TRY(S.event(&done[i]));
    }
    bool used[2] = {false, false};
    int i = 0;
    for (uint64_t off = 0; off < bytes; off += cap, i ^= 1) {
        if (used[i]) HG_TRY(hipEventSynchronize(done[i]));
        const size_t len = size_t(std::min<uint64_t>(cap, bytes - off));
        if (fread(pin[i], 1, len, f) != len) {
            set_error("HierarchicalGraphFileManager::restore: read failed in " + path);
            return ARTEA_EIO;
        }
        HG_TRY(hipMemcpyAsync(reinterpret_cast<char*>(d_body) + off, pin[i], len, hipMemcpyHostToDevice, st));
        HG_TRY(hipEventRecord(done[i], st));
        used[i] = true;
    }
    HG_TRY(hipStreamSynchronize(st));
    return ARTEA_OK;
}

double seconds_since(std::chrono::steady_clock::time_point& t) {
    const auto now = std::chrono::steady_clock::now();
    const double s = std::chrono::duration<double>(now - t).count();
    t = now;
    return s;
}

struct FileCloser {
    FILE* f = nullptr;
    ~FileCloser() {
        if (f) fclose(f);
    }
};

}  // namespace

int artea::hgra_load_device(artea_index* idx, const char* path, uint32_t segment_words, uint32_t trial_records) {
    const std::string p = path;
    auto clock = std::chrono::steady_clock::now();
    double phase[3] = {0, 0, 0};
    HgraHeader hh;
    FileCloser fc;
    int rc = hgra_open(path, &fc.f, &hh);
    if (rc) return rc;
    const uint64_t S = segment_words ? segment_words : kDefaultSegment;
    const uint32_t trial = trial_records ? trial_records : kDefaultTrial;
    hipStream_t st = idx->stream;
    Scratch X;

    // n == 0: the empty graph; the body is not read (file_manager.hpp:196-206).
    const uint64_t W = hh.n == 0 ? 0 : hh.body_words;
    uint32_t* d_body = nullptr;
    HG_TRY(X.dev(&d_body, W));
    rc = upload_body(fc.f, W, d_body, st, X, p);
    if (rc) return rc;
    phase[0] = seconds_since(clock);
    const Body B{d_body, W, hh.n, hh.top, hh.ul_max, hh.bl_max};

    // Segmented parse and the walk that verifies it.
    const uint32_t G = uint32_t((W + S - 1) / S);
    uint64_t V = 0;  // records before the first structural fault (or n)
    SegResult fault{};
    fault.fault = 0;
    std::vector<uint64_t> start(G), base(G);
    if (hh.n > 0) {
        uint64_t* d_start = nullptr;
        SegResult* d_res = nullptr;
        uint32_t* d_upto = nullptr;
        HG_TRY(X.dev(&d_start, G));
        HG_TRY(X.dev(&d_res, G));
        HG_TRY(X.dev(&d_upto, 1));
        std::vector<SegResult> res(G);
        if (G) {
            hipLaunchKernelGGL(guess_kernel, dim3((G + 255) / 256), dim3(256), 0, st, B, S, G, trial, d_start);
            hipLaunchKernelGGL(parse_kernel, dim3((G + 255) / 256), dim3(256), 0, st, B, S, 0u, G, d_start, d_res);
            HG_TRY(hipGetLastError());
            HG_TRY(hipMemcpyAsync(start.data(), d_start, G * 8, hipMemcpyDeviceToHost, st));
            HG_TRY(hipMemcpyAsync(res.data(), d_res, G * sizeof(SegResult), hipMemcpyDeviceToHost, st));
            HG_TRY(hipStreamSynchronize(st));
        }
        uint64_t pos = 0;
        bool stopped = false;
        for (uint32_t j = 0; j < G; ++j) {
            base[j] = V;
            if (stopped || V >= hh.n) {
                start[j] = kNone;
                continue;
            }
            const uint64_t seg_end = std::min<uint64_t>(uint64_t(j) * S + S, W);
            if (pos >= seg_end) {  // no record starts in this segment
                start[j] = kNone;
                continue;
            }
            if (start[j] != pos) {  // a wrong guess: parse again from the verified start
                uint32_t upto = 0;
                hipLaunchKernelGGL(repair_kernel, dim3(1), dim3(64), 0, st, B, S, G, j, pos, d_start, d_res, d_upto);
                HG_TRY(hipGetLastError());
                HG_TRY(hipMemcpyAsync(&upto, d_upto, 4, hipMemcpyDeviceToHost, st));
                HG_TRY(hipStreamSynchronize(st));
                HG_TRY(hipMemcpyAsync(start.data() + j, d_start + j, size_t(upto - j) * 8, hipMemcpyDeviceToHost, st));
                HG_TRY(hipMemcpyAsync(res.data() + j, d_res + j, size_t(upto - j) * sizeof(SegResult),
                                      hipMemcpyDeviceToHost, st));
                HG_TRY(hipStreamSynchronize(st));
            }
            const SegResult& r = res[j];
            const uint64_t take = std::min<uint64_t>(r.count, hh.n - V);
            V += take;
            if (r.fault && take == r.count && V < hh.n) {
                fault = r;
                stopped = true;
                continue;
            }
            pos = r.exit;
        }
        if (!fault.fault && V < hh.n) {  // the body ended before record V
            fault.fault = kHgraTruncVid;
            fault.at = W;
        }
    }

    phase[1] = seconds_since(clock);

    // Per-vertex record starts and the checks of the complete records.
    uint64_t* d_rec = nullptr;
    uint32_t* d_hi = nullptr;
    uint64_t* d_up = nullptr;  // upper rows per vertex, then their exclusive sum in d_ub
    uint64_t* d_ub = nullptr;
    uint32_t* d_maxc = nullptr;
    unsigned long long* d_bad = nullptr;
    uint64_t* d_first = nullptr;
    HG_TRY(X.dev(&d_rec, V));
    HG_TRY(X.dev(&d_hi, V));
    HG_TRY(X.dev(&d_up, V + 1));
    HG_TRY(X.dev(&d_ub, V + 1));
    HG_TRY(X.dev(&d_maxc, 2));
    HG_TRY(X.dev(&d_bad, 1));
    HG_TRY(X.dev(&d_first, 1));
    HG_TRY(hipMemsetAsync(d_maxc, 0, 8, st));
    HG_TRY(hipMemsetAsync(d_bad, 0xFF, 8, st));
    if (V > 0) {
        uint64_t* d_base = nullptr;
        uint64_t* d_start = nullptr;
        HG_TRY(X.dev(&d_base, G));
        HG_TRY(X.dev(&d_start, G));
        HG_TRY(hipMemcpyAsync(d_start, start.data(), G * 8, hipMemcpyHostToDevice, st));
        HG_TRY(hipMemcpyAsync(d_base, base.data(), G * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(offsets_kernel, dim3((G + 255) / 256), dim3(256), 0, st, B, S, G, d_start, d_base, V,
                           d_rec);
        hipLaunchKernelGGL(check_kernel, dim3(grid_waves(V)), dim3(256), 0, st, B, d_rec, uint32_t(V), d_hi, d_up,
                           d_maxc, d_bad);
        HG_TRY(hipGetLastError());
    }
    unsigned long long bad = ~0ull;
    uint32_t maxc[2] = {0, 0};
    HG_TRY(hipMemcpyAsync(&bad, d_bad, 8, hipMemcpyDeviceToHost, st));
    HG_TRY(hipMemcpyAsync(maxc, d_maxc, 8, hipMemcpyDeviceToHost, st));
    HG_TRY(hipStreamSynchronize(st));
    if (bad != ~0ull) {  // an out-of-range id in a complete record comes first
        const uint32_t v = uint32_t(bad >> 32), step = uint32_t(bad & 0xFFFFFFFFu);
        uint64_t rec_v = 0, id = kNone;
        HG_TRY(hipMemcpy(&rec_v, d_rec + v, 8, hipMemcpyDeviceToHost));
        hipLaunchKernelGGL(first_bad_kernel, dim3(1), dim3(64), 0, st, B, rec_v, step + 1, d_first);
        HG_TRY(hipMemcpyAsync(&id, d_first, 8, hipMemcpyDeviceToHost, st));
        HG_TRY(hipStreamSynchronize(st));
        return hgra_fault(kHgraNeighbor, v, 0, uint32_t(id), 0, hh.top, p);
    }
    if (fault.fault) {
        // The faulting record's levels above the fault hold only whole rows.
        uint32_t steps = 0;
        if (fault.fault == kHgraTruncLevel || fault.fault == kHgraCapacity || fault.fault == kHgraTruncNbrs) {
            uint32_t h = 0;
            HG_TRY(hipMemcpy(&h, d_body + fault.at, 4, hipMemcpyDeviceToHost));
            steps = h - fault.level;
        }
        if (steps > 0) {
            uint64_t id = kNone;
            hipLaunchKernelGGL(first_bad_kernel, dim3(1), dim3(64), 0, st, B, fault.at, steps, d_first);
            HG_TRY(hipMemcpyAsync(&id, d_first, 8, hipMemcpyDeviceToHost, st));
            HG_TRY(hipStreamSynchronize(st));
            if (id != kNone) return hgra_fault(kHgraNeighbor, uint32_t(V), 0, uint32_t(id), 0, hh.top, p);
        }
        return hgra_fault(fault.fault, uint32_t(V), fault.level, fault.a, fault.b, hh.top, p);
    }
    if (hh.n != idx->n) {
        set_error("artea_index_set_graph_hgra: graph has " + std::to_string(hh.n) + " vertices, index has " +
                  std::to_string(idx->n) + " vectors");
        return ARTEA_EINVAL;
    }

    // Upper-row bases, the ELL arrays and the fill.
    const uint32_t n = hh.n;
    uint64_t rowsU = 0;
    if (n > 0) {
        size_t tmp_bytes = 0;
        HG_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_up, d_ub, int(n + 1), st));
        char* tmp = nullptr;
        HG_TRY(X.dev(&tmp, tmp_bytes));
        HG_TRY(hipMemsetAsync(d_up + n, 0, 8, st));
        HG_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, d_up, d_ub, int(n + 1), st));
        HG_TRY(hipMemcpyAsync(&rowsU, d_ub + n, 8, hipMemcpyDeviceToHost, st));
        HG_TRY(hipStreamSynchronize(st));
    }
    if (rowsU >= 0xFFFFFFFFull) {
        set_error("upper-level row count exceeds 2^32");
        return ARTEA_EINVAL;
    }
    DeviceGraph g;
    g.top = hh.top;
    g.entry = hh.entry;
    g.ul_max = hh.ul_max;
    g.bl_max = hh.bl_max;
    g.stride0 = round_up(std::max(1u, maxc[0]), 16);
    g.strideU = round_up(std::max(1u, maxc[1]), 16);
    g.rowsU = rowsU;
    auto release = [&]() {
        for (void* q : {(void*)g.ell0, (void*)g.ellU, (void*)g.vinfo})
            if (q) (void)hipFree(q);
    };
    hipError_t e = hipMalloc(&g.ell0, std::max<size_t>(size_t(n) * g.stride0, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&g.ellU, std::max<uint64_t>(rowsU, 1) * g.strideU * 4);
    if (e == hipSuccess) e = hipMalloc(&g.vinfo, std::max<size_t>(n, 1) * sizeof(uint2));
    if (e == hipSuccess && n == 0) {
        const uint2 none{kNoVertex, 0};
        e = hipMemcpyAsync(g.vinfo, &none, sizeof none, hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess && rowsU == 0) e = hipMemsetAsync(g.ellU, 0xFF, size_t(g.strideU) * 4, st);
    if (e == hipSuccess && n == 0) e = hipMemsetAsync(g.ell0, 0xFF, 4, st);
    if (e == hipSuccess && n > 0) {
        hipLaunchKernelGGL(fill_kernel, dim3(grid_waves(n)), dim3(256), 0, st, B, d_rec, d_hi, d_ub, n, g.ell0,
                           g.stride0, g.ellU, g.strideU, g.vinfo);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        release();
        set_error(std::string("HGRA device load: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? ARTEA_ENOMEM : ARTEA_EHIP;
    }
    g.attached = true;
    replace_device_graph(idx, g);
    idx->csr = HostCsr{};
    idx->csr_current = false;
    phase[2] = seconds_since(clock);
    std::copy(phase, phase + 3, idx->hgra_load_s);
    return ARTEA_OK;
}

int artea::hgra_store_device(artea_index* idx, const char* path) {
    const std::string p = path;
    auto clock = std::chrono::steady_clock::now();
    double phase[3] = {0, 0, 0};
    const DeviceGraph& dg = idx->dg;
    const uint32_t n = idx->n;
    hipStream_t st = idx->stream;
    Scratch X;
    const Ell g{dg.ell0, dg.ellU, dg.vinfo, dg.stride0, dg.strideU};
    uint64_t* d_len = nullptr;
    uint64_t* d_off = nullptr;
    HG_TRY(X.dev(&d_len, size_t(n) + 1));
    HG_TRY(X.dev(&d_off, size_t(n) + 1));
    uint64_t words = 0;
    uint32_t* d_out = nullptr;
    if (n > 0) {
        hipLaunchKernelGGL(length_kernel, dim3(grid_waves(uint64_t(n) + 1)), dim3(256), 0, st, g, n, d_len);
        HG_TRY(hipGetLastError());
        size_t tmp_bytes = 0;
        HG_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_len, d_off, int(n + 1), st));
        char* tmp = nullptr;
        HG_TRY(X.dev(&tmp, tmp_bytes));
        HG_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, d_len, d_off, int(n + 1), st));
        HG_TRY(hipMemcpyAsync(&words, d_off + n, 8, hipMemcpyDeviceToHost, st));
        HG_TRY(hipStreamSynchronize(st));
        HG_TRY(X.dev(&d_out, words));
        hipLaunchKernelGGL(record_kernel, dim3(grid_waves(n)), dim3(256), 0, st, g, n, d_off, d_out);
        HG_TRY(hipGetLastError());
        HG_TRY(hipStreamSynchronize(st));
    }
    phase[0] = seconds_since(clock);
    FileCloser fc;
    fc.f = fopen(p.c_str(), "wb");
    if (!fc.f) {
        set_error("HierarchicalGraphFileManager::snapshot: failed to open " + p);
        return ARTEA_EIO;
    }
    const uint32_t hdr[7] = {kMagic, kVersion, n, n == 0 ? 0u : dg.top, dg.ul_max, dg.bl_max, dg.entry};
    bool ok = fwrite(hdr, 4, 7, fc.f) == 7;
    const uint64_t bytes = words * 4;
    if (ok && bytes) {
        // Two pinned buffers: the copy of chunk i + 1 runs while chunk i is written.
        const size_t cap = size_t(std::min<uint64_t>(bytes, kChunk));
        void* pin[2];
        hipEvent_t done[2];
        for (int i = 0; i < 2; ++i) {
            HG_TRY(X.host(&pin[i], cap));
            HG_TRY(X.event(&done[i]));
        }
        const uint64_t chunks = (bytes + cap - 1) / cap;
        auto issue = [&](uint64_t c) -> hipError_t {
            const uint64_t off = c * cap;
            const size_t len = size_t(std::min<uint64_t>(cap, bytes - off));
            hipError_t e = hipMemcpyAsync(pin[c & 1], reinterpret_cast<const char*>(d_out) + off, len,
                                          hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipEventRecord(done[c & 1], st);
            return e;
        };
        HG_TRY(issue(0));
        for (uint64_t c = 0; c < chunks && ok; ++c) {
            HG_TRY(hipEventSynchronize(done[c & 1]));
            if (c + 1 < chunks) HG_TRY(issue(c + 1));
            const size_t len = size_t(std::min<uint64_t>(cap, bytes - c * cap));
            ok = fwrite(pin[c & 1], 1, len, fc.f) == len;
        }
        HG_TRY(hipStreamSynchronize(st));
    }
    if (!ok || fflush(fc.f) != 0 || ferror(fc.f)) {
        set_error("HierarchicalGraphFileManager::snapshot: write failure on " + p);
        return ARTEA_EIO;
    }
    phase[1] = seconds_since(clock);
    std::copy(phase, phase + 3, idx->hgra_store_s);
    return ARTEA_OK;
}

int artea::graph_from_device(const artea_index* idx, HostCsr* out) {
    const DeviceGraph& dg = idx->dg;
    const uint32_t n = idx->n;
    std::vector<uint2> vinfo(std::max<uint32_t>(n, 1));
    std::vector<uint32_t> ell0(size_t(n) * dg.stride0), ellU(std::max<uint64_t>(dg.rowsU, 1) * dg.strideU);
    HG_TRY(hipMemcpy(vinfo.data(), dg.vinfo, vinfo.size() * sizeof(uint2), hipMemcpyDeviceToHost));
    if (!ell0.empty()) HG_TRY(hipMemcpy(ell0.data(), dg.ell0, ell0.size() * 4, hipMemcpyDeviceToHost));
    HG_TRY(hipMemcpy(ellU.data(), dg.ellU, ellU.size() * 4, hipMemcpyDeviceToHost));
    HostCsr c;
    c.n = n;
    c.top = dg.top;
    c.ul_max = dg.ul_max;
    c.bl_max = dg.bl_max;
    c.entry = dg.entry;
    c.highest.assign(n, kInvalidLevel);
    c.local.assign(n, kNoVertex);
    c.offsets.assign(size_t(dg.top) + 1, {});
    std::vector<std::vector<uint32_t>> bucket(size_t(dg.top) + 1);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t h = vinfo[v].x;
        if (h == kInvalidLevel) continue;
        c.highest[v] = h;
        c.local[v] = uint32_t(bucket[h].size());
        bucket[h].push_back(v);
    }
    auto count = [](const uint32_t* row, uint32_t stride) {
        uint32_t k = 0;
        while (k < stride && row[k] != kNoVertex) ++k;
        return k;
    };
    for (uint32_t h = 0; h <= dg.top; ++h) {
        auto& o = c.offsets[h];
        o.assign(bucket[h].size() * (h + 1) + 1, 0);
        for (size_t loc = 0; loc < bucket[h].size(); ++loc) {
            const uint32_t v = bucket[h][loc];
            for (uint32_t l = h;; --l) {
                const uint32_t* row = l == 0 ? ell0.data() + size_t(v) * dg.stride0
                                             : ellU.data() + size_t(vinfo[v].y + (h - l)) * dg.strideU;
                const uint32_t k = count(row, l == 0 ? dg.stride0 : dg.strideU);
                o[loc * (h + 1) + (h - l)] = c.nbrs.size();
                c.nbrs.insert(c.nbrs.end(), row, row + k);
                if (l == 0) break;
            }
        }
        o.back() = c.nbrs.size();
    }
    *out = std::move(c);
    return ARTEA_OK;
}

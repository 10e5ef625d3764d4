// HGRA v1 snapshot / restore of the compact hierarchical graph
// (persistence/hierarchical_graph_file_manager.hpp:20-40 wire format,
// :88-145 snapshot, :152-324 restore).  Same validation order and message
// texts as the reference's ARTEA_ERROR calls.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "host.hpp"

namespace artea {

namespace {
constexpr uint32_t kMagic = 0x48475241u;  // "HGRA"
constexpr uint32_t kVersion = 1u;

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

bool rd32(FILE* f, uint32_t* x) { return fread(x, 4, 1, f) == 1; }
}  // namespace

int hgra_open(const char* path, FILE** out, HgraHeader* hh) {
    const std::string p = path ? path : "";
    *out = nullptr;
    File fh;
    fh.f = fopen(p.c_str(), "rb");
    if (!fh.f) {
        set_error("HierarchicalGraphFileManager::restore: failed to open " + p);
        return ARTEA_EIO;
    }
    FILE* f = fh.f;
    // Bytes in the file: per-vertex tables are sized from the header only
    // once the file is long enough to hold them (every vertex takes at least
    // 4 bytes), so a corrupt count cannot request a huge allocation.
    uint64_t file_bytes = 0;
    if (fseek(f, 0, SEEK_END) == 0) {
        const long e = ftell(f);
        file_bytes = e > 0 ? uint64_t(e) : 0;
    }
    if (fseek(f, 0, SEEK_SET) != 0) {
        set_error("HierarchicalGraphFileManager::restore: failed to open " + p);
        return ARTEA_EIO;
    }
    uint32_t hdr[7];
    if (fread(hdr, 4, 7, f) != 7) {
        set_error("HierarchicalGraphFileManager::restore: header read failed: " + p);
        return ARTEA_EIO;
    }
    char buf[256];
    if (hdr[0] != kMagic) {
        snprintf(buf, sizeof buf, "HierarchicalGraphFileManager::restore: bad magic %#x in ", hdr[0]);
        set_error(buf + p);
        return ARTEA_EIO;
    }
    if (hdr[1] != kVersion) {
        snprintf(buf, sizeof buf, "HierarchicalGraphFileManager::restore: unsupported version %u in ", hdr[1]);
        set_error(buf + p);
        return ARTEA_EIO;
    }
    HgraHeader h;
    h.n = hdr[2];
    h.top = hdr[3];
    h.ul_max = hdr[4];
    h.bl_max = hdr[5];
    h.entry = hdr[6];
    if (h.top == kInvalidLevel || (h.entry != kNoVertex && h.entry >= h.n)) {
        set_error("HierarchicalGraphFileManager::restore: invalid header in " + p);
        return ARTEA_EIO;
    }
    // The device layout addresses at most 64 levels (the reference has no
    // such bound; a graph of ln(N / 1000) levels needs 64 only past 2^92
    // vertices).
    if (h.top >= 64) {
        snprintf(buf, sizeof buf,
                 "HierarchicalGraphFileManager::restore: top_level_id %u exceeds the supported maximum of 63 in ",
                 h.top);
        set_error(buf + p);
        return ARTEA_EIO;
    }
    h.body_words = file_bytes > 28 ? (file_bytes - 28) / 4 : 0;
    *hh = h;
    *out = fh.f;
    fh.f = nullptr;
    return ARTEA_OK;
}

int hgra_fault(uint32_t fault, uint32_t vid, uint32_t level, uint32_t a, uint32_t b, uint32_t top,
               const std::string& p) {
    char buf[256];
    switch (fault) {
        case kHgraTruncVid:
            snprintf(buf, sizeof buf, "HierarchicalGraphFileManager::restore: truncated at vid %u in ", vid);
            break;
        case kHgraHighest:
            snprintf(buf, sizeof buf,
                     "HierarchicalGraphFileManager::restore: vid %u has highest_level_id %u > top_level_id %u in ",
                     vid, a, top);
            break;
        case kHgraTruncLevel:
            snprintf(buf, sizeof buf, "HierarchicalGraphFileManager::restore: truncated at vid %u level %u in ", vid,
                     level);
            break;
        case kHgraCapacity:
            snprintf(buf, sizeof buf,
                     "HierarchicalGraphFileManager::restore: vid %u level %u count %u > capacity %u in ", vid, level,
                     a, b);
            break;
        case kHgraTruncNbrs:
            snprintf(buf, sizeof buf,
                     "HierarchicalGraphFileManager::restore: truncated reading nbrs of vid %u level %u in ", vid,
                     level);
            break;
        default:
            snprintf(buf, sizeof buf, "HierarchicalGraphFileManager::restore: invalid neighbor %u in ", a);
            break;
    }
    set_error(buf + p);
    return ARTEA_EIO;
}

namespace {

int hgra_read_impl(const char* path, HostCsr* out) {
    const std::string p = path ? path : "";
    HgraHeader hh;
    File fh;
    int rc = hgra_open(path, &fh.f, &hh);
    if (rc) return rc;
    FILE* f = fh.f;
    HostCsr g;
    g.n = hh.n;
    g.top = hh.top;
    g.ul_max = hh.ul_max;
    g.bl_max = hh.bl_max;
    g.entry = hh.entry;
    const uint64_t body_words = hh.body_words;
    const uint32_t groups = g.top + 1;
    const size_t table = size_t(std::min<uint64_t>(g.n, body_words + 1));
    g.highest.assign(table, kInvalidLevel);
    g.offsets.assign(groups, {});
    // Per vid, levels h..0: counts then ids, kept in file order first.
    std::vector<uint32_t> lens;
    std::vector<uint32_t> ids;
    std::vector<uint64_t> hist(groups, 0);
    lens.reserve(table * 2);
    for (uint32_t v = 0; v < g.n; ++v) {
        uint32_t h;
        if (!rd32(f, &h)) return hgra_fault(kHgraTruncVid, v, 0, 0, 0, g.top, p);
        g.highest[v] = h;  // v < table: vertices 0..v took at least 4(v + 1) bytes of the body
        if (h == kInvalidLevel) continue;
        if (h > g.top) return hgra_fault(kHgraHighest, v, 0, h, 0, g.top, p);
        hist[h]++;
        for (uint32_t l = h;; --l) {
            uint32_t c;
            if (!rd32(f, &c)) return hgra_fault(kHgraTruncLevel, v, l, 0, 0, g.top, p);
            const uint32_t cap = l == 0 ? g.bl_max : g.ul_max;
            if (c > cap) return hgra_fault(kHgraCapacity, v, l, c, cap, g.top, p);
            lens.push_back(c);
            const size_t at = ids.size();
            if (uint64_t(at) + c > body_words)  // more ids than the file holds
                return hgra_fault(kHgraTruncNbrs, v, l, 0, 0, g.top, p);
            ids.resize(at + c);
            if (c && fread(ids.data() + at, 4, c, f) != c) return hgra_fault(kHgraTruncNbrs, v, l, 0, 0, g.top, p);
            for (uint32_t i = 0; i < c; ++i)
                if (ids[at + i] >= g.n) return hgra_fault(kHgraNeighbor, v, l, ids[at + i], 0, g.top, p);
            if (l == 0) break;
        }
    }
    // Every vertex was read: the per-vertex tables can take their full size.
    g.highest.resize(g.n, kInvalidLevel);
    g.local.assign(g.n, kNoVertex);
    // Buckets in ascending vid order; row counts, prefix sum, then ids.
    std::vector<uint64_t> fill(groups, 0);
    for (uint32_t h = 0; h < groups; ++h) g.offsets[h].assign(hist[h] * (h + 1) + 1, 0);
    size_t li = 0;
    for (uint32_t v = 0; v < g.n; ++v) {
        const uint32_t h = g.highest[v];
        if (h == kInvalidLevel) continue;
        const uint64_t loc = fill[h]++;
        g.local[v] = uint32_t(loc);
        for (uint32_t l = h;; --l) {
            g.offsets[h][loc * (h + 1) + (h - l)] = lens[li++];
            if (l == 0) break;
        }
    }
    uint64_t total = 0;
    for (uint32_t h = 0; h < groups; ++h) {
        auto& o = g.offsets[h];
        for (size_t r = 0; r + 1 < o.size(); ++r) {
            const uint64_t len = o[r];
            o[r] = total;
            total += len;
        }
        o.back() = total;
    }
    g.nbrs.resize(total);
    size_t pos = 0;
    for (uint32_t v = 0; v < g.n; ++v) {
        const uint32_t h = g.highest[v];
        if (h == kInvalidLevel) continue;
        for (uint32_t l = h;; --l) {
            const uint64_t r = uint64_t(g.local[v]) * (h + 1) + (h - l);
            const uint64_t c = g.offsets[h][r + 1] - g.offsets[h][r];
            memcpy(g.nbrs.data() + g.offsets[h][r], ids.data() + pos, c * 4);
            pos += c;
            if (l == 0) break;
        }
    }
    *out = std::move(g);
    return ARTEA_OK;
}

}  // namespace

int hgra_read(const char* path, HostCsr* out) {
    try {
        return hgra_read_impl(path, out);
    } catch (const std::bad_alloc&) {
        set_error(std::string("HierarchicalGraphFileManager::restore: out of host memory reading ") +
                  (path ? path : ""));
        return ARTEA_ENOMEM;
    }
}

int hgra_write(const HostCsr& g, const char* path) {
    const std::string p = path ? path : "";
    File fh;
    fh.f = fopen(p.c_str(), "wb");
    if (!fh.f) {
        set_error("HierarchicalGraphFileManager::snapshot: failed to open " + p);
        return ARTEA_EIO;
    }
    const uint32_t hdr[7] = {kMagic, kVersion, g.n, g.n == 0 ? 0u : g.top, g.ul_max, g.bl_max, g.entry};
    bool ok = fwrite(hdr, 4, 7, fh.f) == 7;
    for (uint32_t v = 0; ok && v < g.n; ++v) {
        const uint32_t h = g.highest[v];
        ok = fwrite(&h, 4, 1, fh.f) == 1;
        if (h == kInvalidLevel) continue;
        for (uint32_t l = h; ok; --l) {
            uint32_t c;
            const uint32_t* r = g.row(v, l, &c);
            ok = fwrite(&c, 4, 1, fh.f) == 1 && (c == 0 || fwrite(r, 4, c, fh.f) == c);
            if (l == 0) break;
        }
    }
    if (!ok || ferror(fh.f)) {
        set_error("HierarchicalGraphFileManager::snapshot: write failure on " + p);
        return ARTEA_EIO;
    }
    return ARTEA_OK;
}

int csr_validate(const HostCsr& g) {
    char buf[256];
    if (g.n == 0) return ARTEA_OK;
    if (g.offsets.size() != size_t(g.top) + 1) {
        set_error("HierarchicalGraph: group count does not match top level");
        return ARTEA_EINVAL;
    }
    std::vector<uint64_t> count(g.top + 1, 0);
    for (uint32_t v = 0; v < g.n; ++v) {
        const uint32_t h = g.highest[v];
        if (h == kInvalidLevel) continue;
        if (h > g.top) {
            snprintf(buf, sizeof buf, "vid %u has highest level %u > top level %u", v, h, g.top);
            set_error(buf);
            return ARTEA_EINVAL;
        }
        count[h]++;
    }
    std::vector<std::vector<uint8_t>> seen(g.top + 1);
    for (uint32_t h = 0; h <= g.top; ++h) {
        if (g.offsets[h].size() != count[h] * (h + 1) + 1) {
            snprintf(buf, sizeof buf, "group %u offsets size does not match its %llu vertices", h,
                     (unsigned long long)count[h]);
            set_error(buf);
            return ARTEA_EINVAL;
        }
        for (size_t r = 0; r + 1 < g.offsets[h].size(); ++r) {
            if (g.offsets[h][r + 1] < g.offsets[h][r] || g.offsets[h][r + 1] > g.nbrs.size()) {
                snprintf(buf, sizeof buf, "group %u offsets are not monotone within the neighbour array", h);
                set_error(buf);
                return ARTEA_EINVAL;
            }
        }
        seen[h].assign(count[h], 0);
    }
    for (uint32_t v = 0; v < g.n; ++v) {
        const uint32_t h = g.highest[v];
        if (h == kInvalidLevel) continue;
        if (g.local[v] >= count[h] || seen[h][g.local[v]]) {
            snprintf(buf, sizeof buf, "vid %u local id %u is out of range or repeated in group %u", v, g.local[v], h);
            set_error(buf);
            return ARTEA_EINVAL;
        }
        seen[h][g.local[v]] = 1;
    }
    // Neighbour ids in parallel; the first bad one, in order, is reported.
    bool bad = false;
    const int64_t nn = int64_t(g.nbrs.size());
#pragma omp parallel for schedule(static) reduction(|| : bad)
    for (int64_t i = 0; i < nn; ++i) bad = bad || g.nbrs[size_t(i)] >= g.n;
    if (bad)
        for (uint32_t x : g.nbrs) {
            if (x >= g.n) {
                snprintf(buf, sizeof buf, "invalid neighbor %u", x);
                set_error(buf);
                return ARTEA_EINVAL;
            }
        }
    if (g.entry != kNoVertex && g.entry >= g.n) {
        set_error("entry point vid is out of range");
        return ARTEA_EINVAL;
    }
    return ARTEA_OK;
}

}  // namespace artea

// ------------------------------------------------------------ C-ABI: graph
using artea::HostCsr;

extern "C" {

int artea_graph_from_csr(uint32_t n, uint32_t top, uint32_t ul_max, uint32_t bl_max, uint32_t entry,
                         const uint32_t* vi, const uint64_t* const* go, const uint32_t* nbrs, uint64_t n_nbrs,
                         artea_graph** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(out, "artea_graph_from_csr: null output");
    ARTEA_CHECK_ARG(top < 64, "artea_graph_from_csr: top level must be < 64");
    ARTEA_CHECK_ARG(n == 0 || (vi && go), "artea_graph_from_csr: null CSR arrays");
    ARTEA_CHECK_ARG(n_nbrs == 0 || nbrs, "artea_graph_from_csr: null neighbour array");
    HostCsr c;
    c.n = n;
    c.top = top;
    c.ul_max = ul_max;
    c.bl_max = bl_max;
    c.entry = entry;
    c.highest.resize(n);
    c.local.resize(n);
    std::vector<uint64_t> count(top + 1, 0);
    for (uint32_t v = 0; v < n; ++v) {
        c.highest[v] = vi[2 * size_t(v)];
        c.local[v] = vi[2 * size_t(v) + 1];
        if (c.highest[v] != artea::kInvalidLevel && c.highest[v] <= top) count[c.highest[v]]++;
    }
    c.offsets.resize(top + 1);
    for (uint32_t h = 0; h <= top; ++h) {
        if (n == 0) {
            c.offsets[h].assign(1, 0);
            continue;
        }
        ARTEA_CHECK_ARG(go[h], "artea_graph_from_csr: null group offsets");
        c.offsets[h].assign(go[h], go[h] + count[h] * (h + 1) + 1);
    }
    if (n_nbrs) c.nbrs.assign(nbrs, nbrs + n_nbrs);
    int rc = artea::csr_validate(c);
    if (rc) return rc;
    *out = new artea_graph{std::move(c)};
    return ARTEA_OK;
}

int artea_graph_restore(const char* path, artea_graph** out) {
    artea::clear_error();
    ARTEA_CHECK_ARG(path && out, "artea_graph_restore: null argument");
    HostCsr c;
    int rc = artea::hgra_read(path, &c);
    if (rc) return rc;
    *out = new artea_graph{std::move(c)};
    return ARTEA_OK;
}

int artea_graph_snapshot(const artea_graph* g, const char* path) {
    artea::clear_error();
    ARTEA_CHECK_ARG(g && path, "artea_graph_snapshot: null argument");
    return artea::hgra_write(g->c, path);
}

int artea_graph_info(const artea_graph* g, uint32_t* n, uint32_t* top, uint32_t* ul, uint32_t* bl, uint32_t* entry,
                     uint64_t* n_nbrs) {
    artea::clear_error();
    ARTEA_CHECK_ARG(g, "artea_graph_info: null graph");
    if (n) *n = g->c.n;
    if (top) *top = g->c.n == 0 ? artea::kInvalidLevel : g->c.top;
    if (ul) *ul = g->c.ul_max;
    if (bl) *bl = g->c.bl_max;
    if (entry) *entry = g->c.entry;
    if (n_nbrs) *n_nbrs = g->c.nbrs.size();
    return ARTEA_OK;
}

const uint32_t* artea_graph_vertex_info(const artea_graph* g) {
    if (!g) return nullptr;
    static thread_local std::vector<uint32_t> scratch;  // interleaved view built on demand
    scratch.resize(size_t(g->c.n) * 2);
    for (uint32_t v = 0; v < g->c.n; ++v) {
        scratch[2 * size_t(v)] = g->c.highest[v];
        scratch[2 * size_t(v) + 1] = g->c.local[v];
    }
    return scratch.data();
}

const uint64_t* artea_graph_group_offsets(const artea_graph* g, uint32_t h, uint64_t* count) {
    if (!g || h >= g->c.offsets.size()) {
        if (count) *count = 0;
        return nullptr;
    }
    if (count) *count = g->c.count(h);
    return g->c.offsets[h].data();
}

const uint32_t* artea_graph_nbrs(const artea_graph* g) { return g ? g->c.nbrs.data() : nullptr; }

const uint32_t* artea_graph_level_nbrs(const artea_graph* g, uint32_t vid, uint32_t level, uint32_t* count) {
    if (!g || vid >= g->c.n) {
        if (count) *count = 0;
        return nullptr;
    }
    uint32_t c;
    const uint32_t* r = g->c.row(vid, level, &c);
    if (count) *count = c;
    return r;
}

int artea_graph_destroy(artea_graph* g) {
    delete g;
    return ARTEA_OK;
}

}  // extern "C"

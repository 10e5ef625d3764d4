// Host-only helpers of libartea_hip.so (no HIP headers): error plumbing and
// the host copy of the compact grouped CSR.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/artea_hip.h"

namespace artea {

void set_error(const std::string& msg);
void clear_error();

#define ARTEA_CHECK_ARG(cond, msg)                                                            \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            ::artea::set_error(msg);                                                          \
            return ARTEA_EINVAL;                                                              \
        }                                                                                     \
    } while (0)

constexpr uint32_t kInvalidVid = 0x7FFFFFFFu;
constexpr uint32_t kNoVertex = 0xFFFFFFFFu;  // ELL padding and unassigned sentinels
constexpr uint32_t kInvalidLevel = 0xFFFFFFFFu;

// Host copy of a compact hierarchical graph in the reference's grouped CSR
// (index/compact_structure/hierarchical_graph.hpp:33-201).
struct HostCsr {
    uint32_t n = 0, top = 0, ul_max = 0, bl_max = 0, entry = kNoVertex;
    std::vector<uint32_t> highest, local;        // per vid
    std::vector<std::vector<uint64_t>> offsets;  // per group h: count_h*(h+1)+1
    std::vector<uint32_t> nbrs;
    uint64_t count(uint32_t h) const { return (offsets[h].size() - 1) / (h + 1); }
    // fetch_level_nbrs (:99-108); empty when level > highest.
    const uint32_t* row(uint32_t vid, uint32_t level, uint32_t* cnt) const {
        const uint32_t h = highest[vid];
        if (h == kInvalidLevel || level > h) {
            *cnt = 0;
            return nbrs.data();
        }
        const uint64_t r = uint64_t(local[vid]) * (h + 1) + (h - level);
        *cnt = uint32_t(offsets[h][r + 1] - offsets[h][r]);
        return nbrs.data() + offsets[h][r];
    }
};

// HGRA v1 restore (persistence/hierarchical_graph_file_manager.hpp:152-324),
// shared by the host reader and the device loader.
struct HgraHeader {
    uint32_t n = 0, top = 0, ul_max = 0, bl_max = 0, entry = 0;
    uint64_t body_words = 0;  // whole 4-byte words after the 28-byte header
};
// Opens `path` and reads and checks the header in the reference's order;
// on success *f is the open file, positioned at the body.
int hgra_open(const char* path, std::FILE** f, HgraHeader* h);
// Failures inside the body, in the order the reference meets them for one
// record: its highest level (missing, above top), then per level from the
// highest down its count (missing, above capacity) and its ids (missing,
// out of range).
enum HgraFault : uint32_t {
    kHgraTruncVid = 1,    // vid
    kHgraHighest = 2,     // vid, a = highest, top
    kHgraTruncLevel = 3,  // vid, level
    kHgraCapacity = 4,    // vid, level, a = count, b = capacity
    kHgraTruncNbrs = 5,   // vid, level
    kHgraNeighbor = 6,    // a = the id
};
// Sets the reference's text for a body failure; returns ARTEA_EIO.
int hgra_fault(uint32_t fault, uint32_t vid, uint32_t level, uint32_t a, uint32_t b, uint32_t top,
               const std::string& path);
int hgra_read(const char* path, HostCsr* g);         // restore semantics + messages
int hgra_write(const HostCsr& g, const char* path);  // snapshot semantics
int csr_validate(const HostCsr& g);                 // consistency of a caller-built CSR

inline uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

}  // namespace artea

// The C-ABI's host graph handle.
struct artea_graph {
    artea::HostCsr c;
};

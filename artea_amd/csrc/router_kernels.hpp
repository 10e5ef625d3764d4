// Router kernel sets: the launch arguments and the per-(metric family,
// candidate-list blocks) kernel tables shared by the host code (router.hip)
// and the instantiation unit (router_inst.hip, compiled once per pair).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace artea {
namespace rk {

// Waves (queries) per block.  Two rather than four: a block's LDS is then
// small enough that 14 waves fit a CU's 160 KiB at D = 128 with a 4-block
// candidate list (12 with four-wave blocks), which the one-byte kernels'
// register use (under 128 VGPRs) can take.
constexpr int kWavesPerBlock = 2;
// Lossy visited cache per wave, sized with the candidate list and the row
// kind.  One-byte-row kernels use under 128 VGPRs, so LDS bounds their
// residency: 7.5 KiB with lists of up to 4 blocks, 5.5 KiB up to 8, 4.1 KiB
// above keep a wave near 10-11 KiB (14-16 waves per CU at D = 128; a smaller
// cache there costs fewer recomputed distances than the waves it buys).
// FP32-row kernels are held to 12 waves per CU by their registers, so their
// caches take the LDS up to that: 10, 8.5, 6 and 4.5 KiB.  The cache holds
// vis_bytes(K, U8) 8-bit codes, or half as many 16-bit codes.
__host__ __device__ constexpr uint32_t vis_bytes(uint32_t K, bool u8) {
    return u8 ? (K <= 4 ? 7680u : (K <= 8 ? 5632u : 4224u))
              : (K <= 4 ? 10240u : (K <= 8 ? 8704u : (K <= 13 ? 6144u : 4608u)));
}

// LDS of one wavefront: the (distance, vid) mirror of a K-block candidate
// list (the first 64 entries of each half double as the expansion chunk's
// candidate staging, whose use never overlaps a merge), the visited cache
// and the query row.  At D = 128 and K = 4 that is 10 KiB with one-byte rows,
// so 16 waves fit a CU's 160 KiB.
__host__ __device__ inline size_t wave_lds_bytes(uint32_t D, uint32_t K, bool u8) {
    return size_t(K) * 512 + vis_bytes(K, u8) + (size_t(D) * 4 + 15) / 16 * 16;
}

struct RouterArgs {
    const float* vecs;
    const uint32_t* vecs_u8;                      // one-byte rows (4*C words each), or null
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
    unsigned long long* stats;                    // [0] ndc, [1] expansions, [2] tail overflow, [3] spare,
                                                  // then the overflow list (kStatsBytes in all)
    uint32_t hmask;                               // 2^B - 1, B = max(13, ceil(log2 n))
    uint32_t vis_mode;                            // visited cache codes: 0 8-bit (n <= 2^20), 1 16-bit (n <= 2^27), 2 as 1 with one choice
    uint32_t* work_counter;                       // persistent mode when non-null
};

// Queries whose candidate list dropped an entry tied with its bound (a
// result that may differ from the reference's) are listed after the four
// counters: uint32 count, then up to kOverflowList query indices.
constexpr uint32_t kOverflowList = 255;
constexpr size_t kStatsBytes = 4 * sizeof(unsigned long long) + (1 + kOverflowList) * sizeof(uint32_t);

using KernelFn = void (*)(RouterArgs);

// Persistent-kernel variants (the batched router's autotuner picks one):
// distance rows per pass and the register budget of the compiled kernel.
enum PersistentVariant : int {
    kVarNarrow = 0,  // 16 rows per pass, no register bound
    kVarWide = 1,    // 32 rows per pass, no register bound
    kVarWide3 = 2,   // 32 rows per pass, registers for 3 waves per SIMD
    kVarWide4 = 3,   // 32 rows per pass, registers for 4 waves per SIMD
    kNumVariants = 4
};

struct KernelSet {
    KernelFn plain;
    KernelFn persistent[kNumVariants];
    KernelFn plain_u8;  // with the one-byte path (indexes that keep one-byte rows)
    KernelFn persistent_u8[kNumVariants];
    bool u8_kernels;    // the *_u8 entries carry the one-byte path (and its LDS layout)
};

// Metric family 0 = L2SQR / L2, 1 = DOT, 2 = COSINE; K = 64-entry register
// blocks of the candidate list.
#define ARTEA_RK_LIST(X) \
    X(0, 1) X(0, 2) X(0, 4) X(0, 8) X(0, 13) X(0, 16) \
    X(1, 1) X(1, 2) X(1, 4) X(1, 8) X(1, 13) X(1, 16) \
    X(2, 1) X(2, 2) X(2, 4) X(2, 8) X(2, 13) X(2, 16)
#define ARTEA_RK_DECLARE(MF, K) KernelSet kernels_mf##MF##_k##K(uint32_t chunks);
ARTEA_RK_LIST(ARTEA_RK_DECLARE)
#undef ARTEA_RK_DECLARE

}  // namespace rk
}  // namespace artea

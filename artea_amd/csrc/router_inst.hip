// Instantiates the router kernels of one (metric family, candidate-list
// blocks) pair; build.py compiles this file once per pair with
// -DARTEA_RK_MF=<0|1|2> -DARTEA_RK_K=<1|2|4|8|13|16>.
#include "router_device.hpp"

#if !defined(ARTEA_RK_MF) || !defined(ARTEA_RK_K)
#error "router_inst.hip needs -DARTEA_RK_MF and -DARTEA_RK_K"
#endif

namespace {

template <int CT>
KernelSet make_set() {
    constexpr int K = ARTEA_RK_K, MF = ARTEA_RK_MF;
    KernelSet s;
    s.plain = router_kernel<K, CT, MF, 0>;
    s.persistent[kVarNarrow] = router_persistent_kernel<K, CT, MF, 1, false, 0>;
    s.persistent[kVarWide] = router_persistent_kernel<K, CT, MF, 1, true, 0>;
    s.persistent[kVarWide3] = router_persistent_kernel<K, CT, MF, 3, true, 0>;
    s.persistent[kVarWide4] = router_persistent_kernel<K, CT, MF, 4, true, 0>;
    // One-byte rows: the same query code with the integer distance path;
    // the variants differ in rows per distance pass (64, 32, 16) and, for the
    // last, a register budget of 4 waves per SIMD.  Queries with other
    // values fall back to 16-row FP32 passes.
    s.plain_u8 = router_kernel<K, CT, MF, 4>;
    s.persistent_u8[0] = router_persistent_kernel<K, CT, MF, 1, false, 4>;
    s.persistent_u8[1] = router_persistent_kernel<K, CT, MF, 1, false, 2>;
    s.persistent_u8[2] = router_persistent_kernel<K, CT, MF, 1, false, 1>;
    s.persistent_u8[3] = router_persistent_kernel<K, CT, MF, 4, false, 4>;
    s.u8_kernels = true;
    return s;
}

// Row widths other than 96 / 112 / 128 floats: one generic kernel serves
// every variant slot (FP32 rows only).
KernelSet make_generic() {
    constexpr int K = ARTEA_RK_K, MF = ARTEA_RK_MF;
    KernelSet s;
    s.plain = s.plain_u8 = router_kernel<K, 0, MF, 0>;
    for (int v = 0; v < kNumVariants; ++v)
        s.persistent[v] = s.persistent_u8[v] = router_persistent_kernel<K, 0, MF, 1, false, 0>;
    s.u8_kernels = false;
    return s;
}

}  // namespace

#define ARTEA_RK_CAT(a, b, c, d) a##b##c##d
#define ARTEA_RK_NAME(MF, K) ARTEA_RK_CAT(kernels_mf, MF, _k, K)

namespace artea {
namespace rk {

KernelSet ARTEA_RK_NAME(ARTEA_RK_MF, ARTEA_RK_K)(uint32_t chunks) {
    switch (chunks) {
    case 6: return make_set<6>();
    case 7: return make_set<7>();
    case 8: return make_set<8>();
    default: return make_generic();
    }
}

}  // namespace rk
}  // namespace artea

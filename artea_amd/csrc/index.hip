// Index handle: device vectors plus the device form of an attached compact
// hierarchical graph; error plumbing; the distance-matrix utility.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <memory>
#include <string>

#include "common.hpp"
#include "device.hpp"

namespace artea {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
}  // namespace artea

using namespace artea;

namespace {

void free_graph(DeviceGraph& g) {
    if (g.ell0) (void)hipFree(g.ell0);
    if (g.vinfo) (void)hipFree(g.vinfo);
    if (g.ellU) (void)hipFree(g.ellU);
    g = DeviceGraph{};
}

}  // namespace

void artea::replace_device_graph(artea_index* idx, const DeviceGraph& g) {
    free_graph(idx->dg);
    idx->dg = g;
}

namespace {

// Level-0 ELL rows (n x stride0, the largest array) into g.ell0: built in
// parallel into two pinned staging halves of 64 MB, each uploaded while the
// other is filled.
hipError_t upload_level0(const HostCsr& c, const DeviceGraph& g, hipStream_t s) {
    const uint32_t n = c.n, stride = g.stride0;
    if (n == 0) return hipMemcpy(g.ell0, &kNoVertex, 4, hipMemcpyHostToDevice);
    const uint32_t chunk_rows = std::min(n, uint32_t(std::max<size_t>(1, (size_t(64) << 20) / (size_t(stride) * 4))));
    const size_t chunk_elems = size_t(chunk_rows) * stride;
    uint32_t* stage = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&stage), 2 * chunk_elems * 4, hipHostMallocDefault);
    for (int h = 0; h < 2 && e == hipSuccess; ++h) e = hipEventCreateWithFlags(&ev[h], hipEventDisableTiming);
    const uint32_t chunks = (n + chunk_rows - 1) / chunk_rows;
    for (uint32_t k = 0; k < chunks && e == hipSuccess; ++k) {
        // the half this chunk fills was last uploaded two chunks ago
        if (k >= 2) e = hipEventSynchronize(ev[k & 1u]);
        if (e != hipSuccess) break;
        const uint32_t v0 = k * chunk_rows, v1 = std::min(n, v0 + chunk_rows);
        uint32_t* half = stage + (k & 1u) * chunk_elems;
#pragma omp parallel for schedule(static)
        for (int64_t vv = v0; vv < int64_t(v1); ++vv) {
            const uint32_t v = uint32_t(vv);
            uint32_t cnt;
            const uint32_t* r = c.row(v, 0, &cnt);
            uint32_t* dst = half + size_t(v - v0) * stride;
            memcpy(dst, r, size_t(cnt) * 4);
            std::fill(dst + cnt, dst + stride, kNoVertex);
        }
        e = hipMemcpyAsync(g.ell0 + size_t(v0) * stride, half, size_t(v1 - v0) * stride * 4, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(ev[k & 1u], s);
    }
    const hipError_t f = hipStreamSynchronize(s);
    if (e == hipSuccess) e = f;
    if (stage) (void)hipHostFree(stage);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    return e;
}

// Converts the grouped CSR into the device ELL layout and uploads it.
int upload_graph(artea_index* idx, const HostCsr& c) {
    DeviceGraph g;
    g.top = c.top;
    g.entry = c.entry;
    g.ul_max = c.ul_max;
    g.bl_max = c.bl_max;
    const uint32_t n = c.n;
    uint32_t max0 = 1, maxU = 1;
    uint64_t rowsU = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t h = c.highest[v];
        if (h == kInvalidLevel) continue;
        uint32_t cnt;
        c.row(v, 0, &cnt);
        max0 = std::max(max0, cnt);
        for (uint32_t l = 1; l <= h; ++l) {
            c.row(v, l, &cnt);
            maxU = std::max(maxU, cnt);
        }
        rowsU += h;
    }
    g.stride0 = round_up(max0, 16);
    g.strideU = round_up(maxU, 16);
    g.rowsU = rowsU;
    const size_t n0 = std::max<size_t>(size_t(n) * g.stride0, 1);
    std::vector<uint32_t> ellU(std::max<uint64_t>(rowsU, 1) * g.strideU, kNoVertex);
    std::vector<uint2> vinfo(std::max<uint32_t>(n, 1), uint2{kNoVertex, 0});
    uint64_t ub = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t h = c.highest[v];
        if (h == kInvalidLevel) continue;
        vinfo[v] = uint2{h, uint32_t(ub)};
        for (uint32_t l = h; l >= 1; --l) {
            uint32_t cnt;
            const uint32_t* r = c.row(v, l, &cnt);
            memcpy(ellU.data() + size_t(ub + (h - l)) * g.strideU, r, size_t(cnt) * 4);
        }
        ub += h;
    }
    if (ub >= 0xFFFFFFFFull) {
        set_error("upper-level row count exceeds 2^32");
        return ARTEA_EINVAL;
    }
    auto fail = [&](hipError_t e, const char* what) {
        free_graph(g);
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return ARTEA_ENOMEM;
    };
    hipError_t e;
    if ((e = hipMalloc(&g.ell0, n0 * 4)) != hipSuccess) return fail(e, "hipMalloc ell0");
    if ((e = hipMalloc(&g.ellU, ellU.size() * 4)) != hipSuccess) return fail(e, "hipMalloc ellU");
    if ((e = hipMalloc(&g.vinfo, vinfo.size() * sizeof(uint2))) != hipSuccess) return fail(e, "hipMalloc vinfo");
    if ((e = upload_level0(c, g, idx->stream)) != hipSuccess) return fail(e, "upload ell0");
    if ((e = hipMemcpy(g.ellU, ellU.data(), ellU.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "upload ellU");
    if ((e = hipMemcpy(g.vinfo, vinfo.data(), vinfo.size() * sizeof(uint2), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "upload vinfo");
    g.attached = true;
    free_graph(idx->dg);
    idx->dg = g;
    return ARTEA_OK;
}

int make_index(uint32_t n, uint32_t dim_pad, int metric, artea_index** out) {
    ARTEA_CHECK_ARG(out, "artea_index_create: null output handle");
    ARTEA_CHECK_ARG(dim_pad > 0 && dim_pad % 16 == 0 && dim_pad <= 4096,
                    "artea_index_create: dim_pad must be a positive multiple of 16 (at most 4096)");
    ARTEA_CHECK_ARG(metric >= 0 && metric <= 3, "artea_index_create: metric must be 0..3");
    ARTEA_CHECK_ARG(n < 0x7FFFFFFFu, "artea_index_create: too many vectors");
    auto idx = std::make_unique<artea_index>();
    idx->n = n;
    idx->dim_pad = dim_pad;
    idx->metric = metric;
    ARTEA_HIP_TRY(hipMalloc(&idx->d_vecs, std::max<size_t>(size_t(n) * dim_pad * 4, 256)));
    ARTEA_HIP_TRY(hipStreamCreateWithFlags(&idx->stream, hipStreamNonBlocking));
    *out = idx.release();
    return ARTEA_OK;
}

__global__ void distance_matrix_kernel(int metric, const float* __restrict__ a, uint32_t na,
                                       const float* __restrict__ b, uint32_t nb, uint32_t D, float* __restrict__ out) {
    // One quad per (i, j) pair; 64 pairs per block.
    const uint32_t pair = blockIdx.x * 64 + threadIdx.x / 4;
    const uint32_t sub = threadIdx.x & 3u;
    const bool ok = pair < uint64_t(na) * nb;
    const uint32_t i = ok ? pair / nb : 0, j = ok ? pair % nb : 0;
    const float* x = a + size_t(i) * D;
    const float* y = b + size_t(j) * D;
    const float qn2 = metric == dev::kCosine ? dev::quad_norm2(x, int(D / 16), sub) : 0.f;
    const float d = dev::quad_distance<0>(x, y, int(D / 16), metric, sub, qn2);
    if (ok && sub == 0) out[size_t(i) * nb + j] = d;
}

// 8-bit rows.  When every stored value is an integer in [0, 255] and a row
// has at most 256 of them, every FP32 operation dist_func_t performs on a
// query of the same kind is exact: differences, squares and products are
// integers below 2^16, and every partial sum of at most 256 of them stays
// below 2^24.  Distances computed from one-byte copies of the values (by any
// summation order, in integers) then equal the FP32 chains bit for bit, and
// the router reads a quarter of the bytes (dev::byte_valued: either zero
// counts, NaN does not).

__global__ void byte_check_kernel(const float* __restrict__ v, size_t count, uint32_t* __restrict__ ok) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        if (!dev::byte_valued(v[i])) *ok = 0u;
}

__global__ void byte_pack_kernel(const float* __restrict__ v, size_t count, uint32_t* __restrict__ out) {
    // four values per output word, little-endian byte order
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count / 4; i += size_t(gridDim.x) * blockDim.x) {
        const float4 f = reinterpret_cast<const float4*>(v)[i];
        out[i] = uint32_t(f.x) | (uint32_t(f.y) << 8) | (uint32_t(f.z) << 16) | (uint32_t(f.w) << 24);
    }
}

// Builds (enable) or drops the 8-bit copy; returns whether the index has one.
int set_u8_rows(artea_index* idx, bool enable, int* active) {
    if (!enable || idx->dim_pad > 256 || idx->n == 0) {
        if (idx->d_u8) ARTEA_HIP_TRY(hipFree(idx->d_u8));
        idx->d_u8 = nullptr;
        *active = 0;
        return ARTEA_OK;
    }
    if (!idx->d_u8) {
        const size_t count = size_t(idx->n) * idx->dim_pad;
        uint32_t* d_ok = nullptr;
        ARTEA_HIP_TRY(hipMalloc(&d_ok, sizeof(uint32_t)));
        const uint32_t one = 1u;
        uint32_t ok = 0u;
        hipError_t e = hipMemcpy(d_ok, &one, sizeof(one), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(byte_check_kernel, dim3(2048), dim3(256), 0, nullptr, idx->d_vecs, count, d_ok);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(&ok, d_ok, sizeof(ok), hipMemcpyDeviceToHost);
        (void)hipFree(d_ok);
        ARTEA_HIP_TRY(e);
        if (ok) {
            uint8_t* d = nullptr;
            ARTEA_HIP_TRY(hipMalloc(&d, count));
            hipLaunchKernelGGL(byte_pack_kernel, dim3(2048), dim3(256), 0, nullptr, idx->d_vecs, count,
                               reinterpret_cast<uint32_t*>(d));
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
            if (e != hipSuccess) {
                (void)hipFree(d);  // never attach a copy that may not hold the values
                ARTEA_HIP_TRY(e);
            }
            idx->d_u8 = d;
        }
    }
    *active = idx->d_u8 ? 1 : 0;
    return ARTEA_OK;
}

}  // namespace

int artea::attach_csr(artea_index* idx, HostCsr&& c) {
    int rc = csr_validate(c);
    if (rc) return rc;
    rc = upload_graph(idx, c);
    if (rc) return rc;
    idx->csr = std::move(c);
    idx->csr_current = true;
    return ARTEA_OK;
}

extern "C" {

const char* artea_last_error(void) { return artea::g_last_error.c_str(); }

const char* artea_version(void) { return "artea_hip 0.1 gfx950 (" __VERSION__ ")"; }

int artea_index_create(const float* vecs, uint32_t n, uint32_t dim_pad, int metric, artea_index** out) {
    clear_error();
    ARTEA_CHECK_ARG(vecs || n == 0, "artea_index_create: null vectors");
    artea_index* idx = nullptr;
    int rc = make_index(n, dim_pad, metric, &idx);
    if (rc) return rc;
    if (n) {
        hipError_t e = hipMemcpy(idx->d_vecs, vecs, size_t(n) * dim_pad * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            artea_index_destroy(idx);
            set_error(std::string("artea_index_create: upload failed: ") + hipGetErrorString(e));
            return ARTEA_EHIP;
        }
    }
    {
        int active = 0;
        rc = set_u8_rows(idx, true, &active);
        if (rc) {
            artea_index_destroy(idx);
            return rc;
        }
    }
    *out = idx;
    return ARTEA_OK;
}

int artea_index_create_device(const float* d_vecs, uint32_t n, uint32_t dim_pad, int metric, artea_index** out) {
    clear_error();
    ARTEA_CHECK_ARG(d_vecs || n == 0, "artea_index_create_device: null vectors");
    artea_index* idx = nullptr;
    int rc = make_index(n, dim_pad, metric, &idx);
    if (rc) return rc;
    if (n) {
        hipError_t e = hipMemcpy(idx->d_vecs, d_vecs, size_t(n) * dim_pad * 4, hipMemcpyDeviceToDevice);
        if (e != hipSuccess) {
            artea_index_destroy(idx);
            set_error(std::string("artea_index_create_device: copy failed: ") + hipGetErrorString(e));
            return ARTEA_EHIP;
        }
    }
    {
        int active = 0;
        rc = set_u8_rows(idx, true, &active);
        if (rc) {
            artea_index_destroy(idx);
            return rc;
        }
    }
    *out = idx;
    return ARTEA_OK;
}

int artea_index_set_graph_csr(artea_index* idx, uint32_t top, uint32_t ul_max, uint32_t bl_max, uint32_t entry,
                              const uint32_t* vertex_info, const uint64_t* const* group_offsets,
                              const uint32_t* nbrs, uint64_t n_nbrs) {
    clear_error();
    ARTEA_CHECK_ARG(idx, "artea_index_set_graph_csr: null index");
    ARTEA_CHECK_ARG(top < 64, "artea_index_set_graph_csr: top level must be < 64");
    ARTEA_CHECK_ARG(idx->n == 0 || (vertex_info && group_offsets), "artea_index_set_graph_csr: null CSR arrays");
    ARTEA_CHECK_ARG(n_nbrs == 0 || nbrs, "artea_index_set_graph_csr: null neighbour array");
    HostCsr c;
    c.n = idx->n;
    c.top = top;
    c.ul_max = ul_max;
    c.bl_max = bl_max;
    c.entry = entry;
    c.highest.resize(c.n);
    c.local.resize(c.n);
    std::vector<uint64_t> count(top + 1, 0);
    for (uint32_t v = 0; v < c.n; ++v) {
        c.highest[v] = vertex_info[2 * size_t(v)];
        c.local[v] = vertex_info[2 * size_t(v) + 1];
        if (c.highest[v] != kInvalidLevel && c.highest[v] <= top) count[c.highest[v]]++;
    }
    if (c.n) {
        c.offsets.resize(top + 1);
        for (uint32_t h = 0; h <= top; ++h) {
            ARTEA_CHECK_ARG(group_offsets[h], "artea_index_set_graph_csr: null group offsets");
            c.offsets[h].assign(group_offsets[h], group_offsets[h] + count[h] * (h + 1) + 1);
        }
        c.nbrs.assign(nbrs, nbrs + n_nbrs);
    }
    return attach_csr(idx, std::move(c));
}


int artea_index_set_graph_hgra(artea_index* idx, const char* path) {
    clear_error();
    ARTEA_CHECK_ARG(idx && path, "artea_index_set_graph_hgra: null argument");
    return hgra_load_device(idx, path, 0, 0);
}

int artea_index_set_graph_hgra_segmented(artea_index* idx, const char* path, uint32_t segment_words,
                                         uint32_t trial_records) {
    clear_error();
    ARTEA_CHECK_ARG(idx && path, "artea_index_set_graph_hgra_segmented: null argument");
    return hgra_load_device(idx, path, segment_words, trial_records);
}

int artea_index_set_graph(artea_index* idx, const artea_graph* g) {
    clear_error();
    ARTEA_CHECK_ARG(idx && g, "artea_index_set_graph: null argument");
    if (g->c.n != idx->n) {
        set_error("artea_index_set_graph: graph has " + std::to_string(g->c.n) + " vertices, index has " +
                  std::to_string(idx->n) + " vectors");
        return ARTEA_EINVAL;
    }
    HostCsr c = g->c;
    int rc = upload_graph(idx, c);
    if (rc) return rc;
    idx->csr = std::move(c);
    idx->csr_current = true;
    return ARTEA_OK;
}

int artea_index_get_graph(const artea_index* idx, artea_graph** out) {
    clear_error();
    ARTEA_CHECK_ARG(idx && out, "artea_index_get_graph: null argument");
    ARTEA_CHECK_ARG(idx->dg.attached, "artea_index_get_graph: no graph attached");
    if (idx->csr_current) {
        *out = new artea_graph{idx->csr};
        return ARTEA_OK;
    }
    HostCsr c;
    const int rc = graph_from_device(idx, &c);
    if (rc) return rc;
    *out = new artea_graph{std::move(c)};
    return ARTEA_OK;
}

int artea_index_snapshot_hgra(const artea_index* idx, const char* path) {
    clear_error();
    ARTEA_CHECK_ARG(idx && path, "artea_index_snapshot_hgra: null argument");
    ARTEA_CHECK_ARG(idx->dg.attached, "artea_index_snapshot_hgra: no graph attached");
    return hgra_store_device(const_cast<artea_index*>(idx), path);
}

int artea_index_layout(const artea_index* idx, uint32_t* stride0, uint32_t* stride_upper, uint64_t* rows_upper) {
    clear_error();
    ARTEA_CHECK_ARG(idx, "artea_index_layout: null index");
    ARTEA_CHECK_ARG(idx->dg.attached, "artea_index_layout: no graph attached");
    if (stride0) *stride0 = idx->dg.stride0;
    if (stride_upper) *stride_upper = idx->dg.strideU;
    if (rows_upper) *rows_upper = idx->dg.rowsU;
    return ARTEA_OK;
}

int artea_index_hgra_seconds(const artea_index* idx, double* load3, double* store3) {
    clear_error();
    ARTEA_CHECK_ARG(idx, "artea_index_hgra_seconds: null index");
    for (int i = 0; i < 3; ++i) {
        if (load3) load3[i] = idx->hgra_load_s[i];
        if (store3) store3[i] = idx->hgra_store_s[i];
    }
    return ARTEA_OK;
}

int artea_index_info(const artea_index* idx, uint32_t* n, uint32_t* dim_pad, uint32_t* top, uint32_t* entry,
                     uint32_t* ul_max, uint32_t* bl_max) {
    clear_error();
    ARTEA_CHECK_ARG(idx, "artea_index_info: null index");
    if (n) *n = idx->n;
    if (dim_pad) *dim_pad = idx->dim_pad;
    if (top) *top = idx->dg.attached ? (idx->n == 0 ? kInvalidLevel : idx->dg.top) : kInvalidLevel;
    if (entry) *entry = idx->dg.attached ? idx->dg.entry : kNoVertex;
    if (ul_max) *ul_max = idx->dg.ul_max;
    if (bl_max) *bl_max = idx->dg.bl_max;
    return ARTEA_OK;
}

const float* artea_index_device_vectors(const artea_index* idx) { return idx ? idx->d_vecs : nullptr; }

int artea_index_set_u8_rows(artea_index* idx, int enable, int* active) {
    clear_error();
    ARTEA_CHECK_ARG(idx, "artea_index_set_u8_rows: null index");
    int on = 0;
    const int rc = set_u8_rows(idx, enable != 0, &on);
    if (active) *active = on;
    return rc;
}

int artea_index_u8_rows(const artea_index* idx) { return idx && idx->d_u8 ? 1 : 0; }

int artea_index_destroy(artea_index* idx) {
    if (!idx) return ARTEA_OK;
    free_graph(idx->dg);
    if (idx->d_vecs) (void)hipFree(idx->d_vecs);
    if (idx->d_u8) (void)hipFree(idx->d_u8);
    if (idx->stream) (void)hipStreamDestroy(idx->stream);
    delete idx;
    return ARTEA_OK;
}

int artea_distance_matrix_device(int metric, const float* d_a, uint32_t na, const float* d_b, uint32_t nb,
                                 uint32_t dim_pad, float* d_out, void* stream) {
    clear_error();
    ARTEA_CHECK_ARG(metric >= 0 && metric <= 3, "artea_distance_matrix_device: metric must be 0..3");
    ARTEA_CHECK_ARG(dim_pad > 0 && dim_pad % 16 == 0, "artea_distance_matrix_device: dim_pad % 16 != 0");
    const uint64_t pairs = uint64_t(na) * nb;
    if (pairs == 0) return ARTEA_OK;
    ARTEA_CHECK_ARG(pairs < (1ull << 32) * 64 / 64 && pairs < 0xFFFFFFFFull,
                    "artea_distance_matrix_device: too many pairs for one call");
    ARTEA_CHECK_ARG(d_a && d_b && d_out, "artea_distance_matrix_device: null buffer");
    const uint32_t blocks = uint32_t((pairs + 63) / 64);
    hipLaunchKernelGGL(distance_matrix_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), metric,
                       d_a, na, d_b, nb, dim_pad, d_out);
    ARTEA_HIP_TRY(hipGetLastError());
    return ARTEA_OK;
}

}  // extern "C"

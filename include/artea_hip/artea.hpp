// artea_hip/artea.hpp -- C++ drop-in surface over the C-ABI (artea_hip.h).
//
// Re-exposes the reference's type names and template signatures for the
// search path (include/artea/cpu/framework/type_context/default_context.hpp
// of the reference): dist_func_t<Metric, Dim>, vector_array_t,
// compact::hierarchical_graph_t, hierarchical_graph_file_manager_t,
// hierarchical_graph_router_t<Metric, Dim> and the CandidateEntry results.
// Status codes of the C-ABI become the reference's exception types:
// std::runtime_error (ARTEA_ERROR), std::overflow_error (thresholds).
// Header-only; link with libartea_hip.so.
#pragma once

#include <cfloat>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../artea_hip.h"

namespace artea {
namespace hip {

enum class DistanceMetricsT : uint8_t { EUCLIDEAN_SQR = 0, DOT = 1, COSINE = 2, EUCLIDEAN = 3 };

using vertex_id_t = uint32_t;
using vertex_num_t = uint32_t;
using layer_id_t = uint32_t;
using distance_t = float;

inline void throw_on(int rc) {
    if (rc == ARTEA_OK) return;
    const std::string msg = artea_last_error();
    if (rc == ARTEA_EOVERFLOW) throw std::overflow_error(msg);
    throw std::runtime_error(msg);
}

// dist_func_t<Metric, Dim> (utils/simd_distance.hpp:64-79): stateless,
// Dim = padded dimension.  The device kernels compute it; this tag carries
// the compile-time axes, as the reference's template does.
template <DistanceMetricsT Metric, uint32_t Dim>
struct dist_func_t {
    static_assert(Dim % 16 == 0, "Dim must be the SIMD-padded dimension (a multiple of 16)");
    static constexpr DistanceMetricsT metric = Metric;
    static constexpr uint32_t dim = Dim;
};

// CandidateEntry (router/data_structures/candidate_entry.hpp:42-201).
class CandidateEntry {
  public:
    static constexpr vertex_id_t MASK_ID = 0x7FFFFFFFu;
    CandidateEntry() : _vid(ARTEA_INVALID_VID), _dist(0.0f) {}
    CandidateEntry(vertex_id_t vid, distance_t d) : _vid(vid & MASK_ID), _dist(d) {}
    static CandidateEntry make_invalid_entry() { return {ARTEA_INVALID_VID, FLT_MAX}; }
    vertex_id_t get_vid() const { return _vid & MASK_ID; }
    distance_t get_distance() const { return _dist; }
    bool is_invalid() const { return (_vid & MASK_ID) == MASK_ID; }
    bool operator<(const CandidateEntry& o) const {
        return _dist != o._dist ? _dist < o._dist : get_vid() < o.get_vid();
    }
    bool operator==(const CandidateEntry& o) const { return _dist == o._dist && get_vid() == o.get_vid(); }

  private:
    vertex_id_t _vid;
    distance_t _dist;
};
using knn_results_t = std::vector<CandidateEntry>;

// vector_array_t (containers/vector_array.hpp:26-506): row-major rows of
// the padded dimension, zero tail.  Host storage.
class vector_array_t {
  public:
    vector_array_t(vertex_num_t n, uint32_t dim) : _n(n), _dim(dim), _data(size_t(n) * dim, 0.0f) {}
    float* get(vertex_id_t i) { return _data.data() + size_t(i) * _dim; }
    const float* get(vertex_id_t i) const { return _data.data() + size_t(i) * _dim; }
    float* get_all() { return _data.data(); }
    const float* get_all() const { return _data.data(); }
    vertex_num_t get_num_vecs() const { return _n; }
    uint32_t get_vec_dim() const { return _dim; }

  private:
    vertex_num_t _n;
    uint32_t _dim;
    std::vector<float> _data;
};

namespace compact {
// compact::HierarchicalGraph (index/compact_structure/hierarchical_graph.hpp).
class hierarchical_graph_t {
  public:
    static constexpr bool is_compacted = true;
    static constexpr layer_id_t invalid_level_id = ARTEA_INVALID_LEVEL;

    explicit hierarchical_graph_t(artea_graph* g) : _g(g, &artea_graph_destroy) { _refresh(); }
    // Builder path: grouped CSR arrays (copied).
    hierarchical_graph_t(vertex_num_t n, layer_id_t top, vertex_num_t ul_max, vertex_num_t bl_max,
                         vertex_id_t entry, const std::vector<uint32_t>& vertex_info,
                         const std::vector<std::vector<uint64_t>>& group_offsets,
                         const std::vector<vertex_id_t>& nbrs)
        : _g(nullptr, &artea_graph_destroy) {
        std::vector<const uint64_t*> ptrs;
        for (auto& o : group_offsets) ptrs.push_back(o.data());
        artea_graph* g = nullptr;
        throw_on(artea_graph_from_csr(n, top, ul_max, bl_max, entry, vertex_info.data(), ptrs.data(), nbrs.data(),
                                      nbrs.size(), &g));
        _g.reset(g);
        _refresh();
    }

    vertex_num_t get_num_vertices() const { return _n; }
    layer_id_t top_occupied_level_id() const { return _top; }
    vertex_id_t entry_point_vid() const { return _entry; }
    vertex_num_t ul_max_nbr_size() const { return _ul; }
    vertex_num_t bl_max_nbr_size() const { return _bl; }
    std::span<const vertex_id_t> fetch_level_nbrs(vertex_id_t vid, layer_id_t level) const {
        uint32_t c = 0;
        const uint32_t* p = artea_graph_level_nbrs(_g.get(), vid, level, &c);
        return {p, c};
    }
    const artea_graph* handle() const { return _g.get(); }

  private:
    void _refresh() {
        uint64_t nn = 0;
        throw_on(artea_graph_info(_g.get(), &_n, &_top, &_ul, &_bl, &_entry, &nn));
    }
    std::unique_ptr<artea_graph, int (*)(artea_graph*)> _g;
    vertex_num_t _n = 0, _ul = 0, _bl = 0;
    layer_id_t _top = 0;
    vertex_id_t _entry = 0;
};
}  // namespace compact

// HierarchicalGraphFileManager (persistence/hierarchical_graph_file_manager.hpp).
struct hierarchical_graph_file_manager_t {
    static void snapshot(const compact::hierarchical_graph_t& g, const std::string& path) {
        throw_on(artea_graph_snapshot(g.handle(), path.c_str()));
    }
    static compact::hierarchical_graph_t restore(const std::string& path) {
        artea_graph* g = nullptr;
        throw_on(artea_graph_restore(path.c_str(), &g));
        return compact::hierarchical_graph_t(g);
    }
};

// hierarchical_graph_router_t<Metric, Dim>
// (router/hierarchical_graph_router.hpp:93-430).  Borrows base_vecs as the
// reference does; the device copy of the vectors is made once at
// construction, the graph is attached on first use (and whenever a
// different graph object is passed).
template <DistanceMetricsT Metric, uint32_t Dim>
class hierarchical_graph_router_t {
  public:
    hierarchical_graph_router_t(const vector_array_t& base_vecs, const dist_func_t<Metric, Dim>&, uint32_t topk,
                                vertex_num_t candidate_queue_size = 16)
        : _base(base_vecs), _topk(topk), _queue(candidate_queue_size), _idx(nullptr, &artea_index_destroy),
          _router(nullptr, &artea_router_destroy) {
        if (base_vecs.get_vec_dim() != Dim) throw std::runtime_error("base_vecs dimension differs from Dim");
        if (candidate_queue_size < topk)
            throw std::runtime_error("candidate_queue_size (" + std::to_string(candidate_queue_size) +
                                     ") must be >= topk (" + std::to_string(topk) + ")");
        artea_index* idx = nullptr;
        throw_on(artea_index_create(base_vecs.get_all(), base_vecs.get_num_vecs(), Dim, int(Metric), &idx));
        _idx.reset(idx);
        artea_router* r = nullptr;
        throw_on(artea_router_create(idx, topk, candidate_queue_size, &r));
        _router.reset(r);
    }

    void initialize() {}

    // batch_query<RandomSeeding, enable_early_stop> (:293-319).  The
    // explicit-threshold overload takes the per-level thresholds in query
    // distance units (what _early_stop_thresholds computes, :388-408).
    template <bool RandomSeeding = false, bool enable_early_stop = false>
    knn_results_t batch_query(const vector_array_t& queries, const compact::hierarchical_graph_t& g,
                              const std::vector<float>& thresholds = {}) const {
        static_assert(!RandomSeeding, "RandomSeeding draws from MKL streams; the device router seeds from the "
                                      "graph's entry point");
        return _run(queries, g, enable_early_stop ? uint32_t(ARTEA_EARLY_STOP) : 0u, thresholds);
    }
    template <bool RandomSeeding = false>
    knn_results_t batch_query_l0_only(const vector_array_t& queries, const compact::hierarchical_graph_t& g) const {
        static_assert(!RandomSeeding, "RandomSeeding is not available on the device router");
        return _run(queries, g, ARTEA_L0_ONLY, {});
    }
    // query / query_l0_only (:164-192, :258-286): results trimmed to the
    // found entries, as the reference returns min(topk, |T|) entries.
    template <bool RandomSeeding = false, bool enable_early_stop = false>
    knn_results_t query(const float* q, const compact::hierarchical_graph_t& g,
                        const std::vector<float>& thresholds = {}) const {
        vector_array_t one(1, Dim);
        std::copy(q, q + Dim, one.get(0));
        auto r = batch_query<RandomSeeding, enable_early_stop>(one, g, thresholds);
        while (!r.empty() && r.back().is_invalid()) r.pop_back();
        return r;
    }
    template <bool RandomSeeding = false>
    knn_results_t query_l0_only(const float* q, const compact::hierarchical_graph_t& g) const {
        vector_array_t one(1, Dim);
        std::copy(q, q + Dim, one.get(0));
        auto r = batch_query_l0_only<RandomSeeding>(one, g);
        while (!r.empty() && r.back().is_invalid()) r.pop_back();
        return r;
    }

  private:
    knn_results_t _run(const vector_array_t& queries, const compact::hierarchical_graph_t& g, uint32_t flags,
                       const std::vector<float>& thr) const {
        if (_attached != g.handle()) {
            throw_on(artea_index_set_graph(_idx.get(), g.handle()));
            _attached = g.handle();
        }
        const uint32_t nq = queries.get_num_vecs();
        std::vector<uint32_t> v(size_t(nq) * _topk);
        std::vector<float> d(size_t(nq) * _topk);
        throw_on(artea_router_batch_query(_router.get(), queries.get_all(), nq, flags,
                                          thr.empty() ? nullptr : thr.data(), v.data(), d.data()));
        knn_results_t out(v.size());
        for (size_t i = 0; i < v.size(); ++i) out[i] = CandidateEntry(v[i], d[i]);
        return out;
    }

    const vector_array_t& _base;
    uint32_t _topk;
    vertex_num_t _queue;
    std::unique_ptr<artea_index, int (*)(artea_index*)> _idx;
    std::unique_ptr<artea_router, int (*)(artea_router*)> _router;
    mutable const artea_graph* _attached = nullptr;
};

}  // namespace hip
}  // namespace artea

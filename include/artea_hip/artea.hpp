// artea_hip/artea.hpp -- C++ drop-in surface over the C-ABI (artea_hip.h).
//
// Re-exposes the reference's type names and template signatures for the
// search and refine paths (include/artea/cpu/framework/type_context/
// default_context.hpp of the reference): dist_func_t<Metric, Dim>,
// vector_array_t, compact::hierarchical_graph_t,
// hierarchical_graph_file_manager_t, hierarchical_graph_router_t<Metric, Dim>
// and the CandidateEntry results; the configs the router's early-stop
// thresholds come from (rgraph_config_t, pruning_config_t); and the propagate
// engine with its single-layer graph, Neighbor rows and updaters
// (propagate_engine_t<Metric, Dim>, dynamic::single_layer_graph_t).
// Status codes of the C-ABI become the reference's exception types:
// std::runtime_error (ARTEA_ERROR), std::overflow_error (thresholds).
// Header-only; link with libartea_hip.so.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>
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

// PruningConfig of the artea graph (index/artea_graph/configs.hpp:66-120):
// scale (alpha), shift (tau) and the L0 distance scale (rho).
struct pruning_config_t {
    pruning_config_t(float scale_coeffs, float shifted_coeffs, float l0_min_distance = 1.0f)
        : _scale(scale_coeffs), _shift(shifted_coeffs), _l0(l0_min_distance) {}
    pruning_config_t& scale_coeffs(float v) { _scale = v; return *this; }
    pruning_config_t& shifted_coeffs(float v) { _shift = v; return *this; }
    pruning_config_t& l0_min_distance(float v) { _l0 = v; return *this; }
    float scale_coeffs() const { return _scale; }
    float shifted_coeffs() const { return _shift; }
    float l0_min_distance() const { return _l0; }

  private:
    float _scale, _shift, _l0;
};

// PropagateConfig (index/conv_graph/configs.hpp:19-97).
struct propagate_config_t {
    propagate_config_t(uint32_t num_build_loops, uint32_t num_triu_iters, float prefill_ratio = 1.0f,
                       uint32_t num_routing_loops = 0, uint32_t routing_topk = 64, uint32_t routing_queue_size = 96)
        : _loops(num_build_loops), _triu(num_triu_iters), _prefill(prefill_ratio), _routing(num_routing_loops),
          _rtopk(routing_topk), _rqueue(routing_queue_size) {}
    uint32_t num_build_loops() const { return _loops; }
    uint32_t num_triu_iters() const { return _triu; }
    float prefill_ratio() const { return _prefill; }
    uint32_t num_routing_loops() const { return _routing; }
    uint32_t routing_topk() const { return _rtopk; }
    uint32_t routing_queue_size() const { return _rqueue; }
    uint32_t resolve_routing_topk(uint32_t) const { return _rtopk > 0 ? _rtopk : 64u; }
    uint32_t resolve_routing_queue_size(uint32_t) const { return _rqueue > 0 ? _rqueue : 96u; }

  private:
    uint32_t _loops, _triu;
    float _prefill;
    uint32_t _routing, _rtopk, _rqueue;
};

// RGraphConfig (index/stacked_rgraph/configs.hpp:65-285): the r-net's radii
// R_0 = rho, R_h = rho (1 + tau_k) beta^(h-1), and the routing bounds the
// early-stop thresholds and ARC cutoffs derive from.  The arithmetic follows
// the reference step by step (radii rounded to float, sums in double).
struct rgraph_config_t {
    rgraph_config_t(float rnet_beta, float tau_k, float tau, float l0_min_distance, uint32_t search_nn_qs,
                    uint32_t ul_select_nbrs_qs = 100, uint32_t bl_select_nbrs_qs = 100, uint32_t ul_max_nbr_size = 32,
                    uint32_t bl_max_nbr_size = 64)
        : _beta(rnet_beta), _tau_k(tau_k), _tau(tau), _l0(l0_min_distance), _search(search_nn_qs),
          _ul_sel(ul_select_nbrs_qs), _bl_sel(bl_select_nbrs_qs), _ul(ul_max_nbr_size), _bl(bl_max_nbr_size) {
        // validation order and messages of configs.hpp:86-114
        if (!std::isfinite(rnet_beta) || rnet_beta <= 1.0f)
            throw std::runtime_error("rnet_beta (" + std::to_string(rnet_beta) + ") must be finite and > 1");
        if (!std::isfinite(tau_k) || tau_k < 0.0f)
            throw std::runtime_error("tau_k (" + std::to_string(tau_k) + ") must be finite and >= 0");
        if (!std::isfinite(tau) || tau < 0.0f)
            throw std::runtime_error("shifted_coeffs / tau (" + std::to_string(tau) + ") must be finite and >= 0");
        if (!std::isfinite(l0_min_distance) || l0_min_distance <= 0.0f)
            throw std::runtime_error("l0_min_distance (" + std::to_string(l0_min_distance) +
                                     ") must be finite and > 0");
        const double l1 = double(l0_min_distance) * (1.0 + double(tau_k));
        if (!std::isfinite(l1) || l1 > double(std::numeric_limits<float>::max()))
            throw std::runtime_error("tau_k (" + std::to_string(tau_k) + ") produces an unrepresentable L1 radius");
        if (search_nn_qs < 1) throw std::runtime_error("search_nn_qs (0) must be >= 1");
        if (ul_select_nbrs_qs < 1) throw std::runtime_error("ul_select_nbrs_qs (0) must be >= 1");
        if (bl_select_nbrs_qs < 1) throw std::runtime_error("bl_select_nbrs_qs (0) must be >= 1");
    }
    float rnet_beta() const { return _beta; }
    float tau_k() const { return _tau_k; }
    float tau() const { return _tau; }
    float l0_min_distance() const { return _l0; }
    uint32_t search_nn_qs() const { return _search; }
    uint32_t ul_select_nbrs_qs() const { return _ul_sel; }
    uint32_t bl_select_nbrs_qs() const { return _bl_sel; }
    uint32_t ul_max_nbr_size() const { return _ul; }
    uint32_t bl_max_nbr_size() const { return _bl; }

    // radius_at (:130-138).
    float radius_at(layer_id_t h) const {
        if (h == 0) return _l0;
        return float(double(_l0) * (1.0 + double(_tau_k)) * std::pow(double(_beta), double(h - 1)));
    }
    // aspect_ratio_constraint (:165-179): Delta_h.
    float aspect_ratio_constraint(layer_id_t h, const pruning_config_t& p) const {
        const Bounds b = _routing_bounds(h, p);
        const double next_radius = radius_at(h + 1);
        if (!std::isfinite(next_radius) || next_radius <= 0.0)
            throw std::overflow_error("ARC requires finite positive layer radii");
        const double psi = b.distance_bound / b.radius;
        const double arc = (b.gamma + 1.0) * psi + b.gamma * (next_radius / b.radius);
        if (!std::isfinite(arc) || arc > double(std::numeric_limits<float>::max()))
            throw std::overflow_error("ARC threshold is not representable");
        return float(arc);
    }
    // early_stop_threshold (:208-221): T_h = gamma (tau rho + sum_{i<=h} R_i).
    float early_stop_threshold(layer_id_t h, const pruning_config_t& p) const {
        if (h == 0) throw std::out_of_range("Early stopping applies only to upper layers (h >= 1)");
        const Bounds b = _routing_bounds(h, p);
        const double t = b.gamma * b.distance_bound;
        if (!std::isfinite(t) || t > double(std::numeric_limits<float>::max()))
            throw std::overflow_error("Early-stop threshold is not representable");
        return float(t);
    }
    // compute_max_allowed_level_id (:241-249).
    static layer_id_t compute_max_allowed_level_id(vertex_num_t total) {
        if (total == 0) return 1;
        const double ratio = double(total) / 1000.0;
        if (ratio <= 1.0) return 1;
        const layer_id_t c = layer_id_t(std::ceil(std::log(ratio)));
        return c > 1 ? c : 1;
    }

  private:
    struct Bounds {
        double gamma, radius, distance_bound;
    };
    // _routing_bounds (:260-284), theta = 1.
    Bounds _routing_bounds(layer_id_t h, const pruning_config_t& p) const {
        const double alpha = p.scale_coeffs(), tau = p.shifted_coeffs(), rho = p.l0_min_distance();
        if (!std::isfinite(alpha) || alpha <= 1.0 || !std::isfinite(tau) || tau < 0.0 || !std::isfinite(rho) ||
            rho <= 0.0)
            throw std::invalid_argument("Routing thresholds require finite alpha > 1, tau >= 0 and rho > 0");
        if (h == std::numeric_limits<layer_id_t>::max())
            throw std::out_of_range("Routing thresholds require a valid layer ID");
        const double radius = radius_at(h);
        if (!std::isfinite(radius) || radius <= 0.0)
            throw std::overflow_error("Routing thresholds require finite positive layer radii");
        double sum = radius;
        for (layer_id_t i = 0; i < h; ++i) sum += double(radius_at(i));
        const double gamma = (alpha + 1.0) / (alpha - 1.0) + 1.0;
        return {gamma, radius, tau * rho + sum};
    }
    float _beta, _tau_k, _tau, _l0;
    uint32_t _search, _ul_sel, _bl_sel, _ul, _bl;
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
    // The config overload (:305-319): with enable_early_stop the per-level
    // thresholds come from rgraph_config.early_stop_threshold(h, pruning),
    // squared for EUCLIDEAN_SQR, as _early_stop_thresholds computes them
    // (:388-408); std::overflow_error when one is not representable.
    template <bool RandomSeeding = false, bool enable_early_stop = false>
    knn_results_t batch_query(const vector_array_t& queries, const compact::hierarchical_graph_t& g,
                              const rgraph_config_t& rgraph_config, const pruning_config_t& pruning_config) const {
        static_assert(!RandomSeeding, "RandomSeeding draws from MKL streams; the device router seeds from the "
                                      "graph's entry point");
        if constexpr (enable_early_stop) {
            return _run(queries, g, ARTEA_EARLY_STOP, early_stop_thresholds(g, rgraph_config, pruning_config));
        } else {
            return _run(queries, g, 0u, {});
        }
    }
    // _early_stop_thresholds (:388-408): index = level, entry 0 unused; empty
    // for a graph of one level.
    static std::vector<float> early_stop_thresholds(const compact::hierarchical_graph_t& g,
                                                    const rgraph_config_t& rgraph_config,
                                                    const pruning_config_t& pruning_config) {
        const layer_id_t top = g.top_occupied_level_id();
        if (top == compact::hierarchical_graph_t::invalid_level_id || top == 0) return {};
        std::vector<float> thr(size_t(top) + 1, 0.0f);
        for (layer_id_t h = 1; h <= top; ++h) {
            double t = rgraph_config.early_stop_threshold(h, pruning_config);
            if constexpr (Metric == DistanceMetricsT::EUCLIDEAN_SQR) t *= t;
            if (!std::isfinite(t) || t <= 0 || t > double(std::numeric_limits<float>::max()) || float(t) == 0.0f)
                throw std::overflow_error("Early-stop threshold is not representable in query distance units");
            thr[h] = float(t);
        }
        return thr;
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

// ------------------------------------------------------------- refine path

// Neighbor (index/neighbor.hpp:43-258): 8 bytes, vid:31 | new:1, then the
// FP32 distance -- the same bytes the device rows hold.
struct Neighbor {
    static constexpr vertex_id_t MASK_STATUS_NEW = 0x80000000u;
    static constexpr vertex_id_t MASK_ID = 0x7FFFFFFFu;
    vertex_id_t nbr_vid_and_status;
    distance_t distance;

    constexpr Neighbor() : nbr_vid_and_status(0xFFFFFFFFu), distance(0.0f) {}
    constexpr Neighbor(vertex_id_t vid, distance_t d) : nbr_vid_and_status(vid & MASK_ID), distance(d) {}
    constexpr Neighbor(vertex_id_t vid, distance_t d, bool is_new)
        : nbr_vid_and_status(is_new ? ((vid & MASK_ID) | MASK_STATUS_NEW) : (vid & MASK_ID)), distance(d) {}
    static constexpr Neighbor make_invalid_nbr() { return Neighbor(0xFFFFFFFFu, FLT_MAX); }
    static Neighbor make_new_nbr(vertex_id_t vid, distance_t d) { return Neighbor(vid, d, true); }
    static Neighbor make_old_nbr(vertex_id_t vid, distance_t d) { return Neighbor(vid, d, false); }
    vertex_id_t get_vid() const { return nbr_vid_and_status & MASK_ID; }
    bool is_new() const { return (nbr_vid_and_status & MASK_STATUS_NEW) != 0; }
    bool is_old() const { return !is_new(); }
    void mark_as_new() { nbr_vid_and_status |= MASK_STATUS_NEW; }
    void mark_as_old() { nbr_vid_and_status &= MASK_ID; }
    void set_status(bool is_new_) { is_new_ ? mark_as_new() : mark_as_old(); }
    bool is_invalid() const { return get_vid() == MASK_ID; }
    distance_t get_distance() const { return distance; }
    void set_distance(distance_t d) { distance = d; }
};
static_assert(sizeof(Neighbor) == 8, "Neighbor must stay 8 bytes (the device row layout)");
using nbr_t = Neighbor;
using nbr_arr_t = std::vector<nbr_t>;

// SingleLayerConfig: the layer's per-vertex neighbour capacity.
struct sl_config_t {
    explicit sl_config_t(vertex_num_t max_nbr_size = 0) : _max(max_nbr_size) {}
    vertex_num_t max_nbr_size() const { return _max; }
    sl_config_t& max_nbr_size(vertex_num_t v) {
        _max = v;
        return *this;
    }

  private:
    vertex_num_t _max;
};

namespace dynamic {
// dynamic::SingleLayerGraph (index/dynamic_structure/single_layer_graph.hpp:
// 29-227): one host row of Neighbor entries per local vertex; dense (row =
// global vid) or sparse (local_to_global / global_to_local maps).  Rows hold
// global vids.  The propagate engine moves the rows to the device for a pass
// and back.
class single_layer_graph_t {
  public:
    single_layer_graph_t(const vector_array_t& vecs, sl_config_t cfg)
        : _n(vecs.get_num_vecs()), _cfg(cfg), _vecs(vecs), _rows(_n) {}
    single_layer_graph_t(const vector_array_t& global_vecs, sl_config_t cfg, std::vector<vertex_id_t> local_to_global,
                         std::vector<vertex_id_t> global_to_local)
        : _n(vertex_num_t(local_to_global.size())), _cfg(cfg), _vecs(global_vecs), _rows(_n),
          _l2g(std::move(local_to_global)), _g2l(std::move(global_to_local)) {}
    single_layer_graph_t(const single_layer_graph_t&) = delete;
    single_layer_graph_t& operator=(const single_layer_graph_t&) = delete;

    vertex_num_t get_num_vertices() const { return _n; }
    bool is_identity_mapped() const { return _g2l.empty(); }
    const std::vector<vertex_id_t>& local_to_global() const { return _l2g; }
    const std::vector<vertex_id_t>& global_to_local() const { return _g2l; }
    vertex_id_t local_id_of(vertex_id_t storage_vid) const { return _g2l.empty() ? storage_vid : _g2l[storage_vid]; }
    vertex_id_t get_storage_vid(vertex_num_t local) const { return _l2g.empty() ? local : _l2g[local]; }
    const sl_config_t& layer_config() const { return _cfg; }
    sl_config_t& layer_config() { return _cfg; }
    std::vector<nbr_arr_t>& get_nbrs_arr() { return _rows; }
    const std::vector<nbr_arr_t>& get_nbrs_arr() const { return _rows; }
    nbr_arr_t& fetch_nbrs(vertex_id_t storage_vid) { return _rows[local_id_of(storage_vid)]; }
    const nbr_arr_t& fetch_nbrs(vertex_id_t storage_vid) const { return _rows[local_id_of(storage_vid)]; }
    const vector_array_t& get_vecs_data() const { return _vecs; }

  private:
    vertex_num_t _n;
    sl_config_t _cfg;
    const vector_array_t& _vecs;
    std::vector<nbr_arr_t> _rows;
    std::vector<vertex_id_t> _l2g, _g2l;
};
}  // namespace dynamic

// The updaters (refiner/updaters/*.hpp) as descriptors: the device runs
// them; make_updater fills in what the engine supplies, as the reference's
// factory does (propagate_engine.hpp:281-309).
// TriangleUpdater(scale_coeffs, shifted_coeffs) (triangle_updater.hpp:114-237).
template <DistanceMetricsT Metric, uint32_t Dim>
struct triangle_updater_t {
    float scale_coeffs = 1.0f;
    float shifted_coeffs = 0.0f;
};
// ReverseUpdater (reverse_updater.hpp:72-100).
template <DistanceMetricsT Metric, uint32_t Dim>
struct reverse_updater_t {};
// RandomUpdater(rand_gen_size, start_vid, end_vid) (random_updater.hpp:67-142):
// ids drawn from [start_vid, end_vid) of the layer's rows.
template <DistanceMetricsT Metric, uint32_t Dim>
struct random_updater_t {
    vertex_num_t rand_gen_size = 0;
    vertex_id_t start_vid = 0;
    vertex_id_t end_vid = 0;
};
// TruncateUpdater / arc_updater_t (truncate_updater.hpp:58-76).
template <DistanceMetricsT Metric, uint32_t Dim>
struct truncate_updater_t {};
template <DistanceMetricsT Metric, uint32_t Dim>
using arc_updater_t = truncate_updater_t<Metric, Dim>;
// PruningUpdater(scale_coeffs, shifted_coeffs) (pruning_updater.hpp:130-177):
// the scaled rule, pairs of two old entries skipped; the shift is not used
// by that rule.  Rows are pruned in place, no logs.
template <DistanceMetricsT Metric, uint32_t Dim>
struct pruning_updater_t {
    float scale_coeffs = 1.0f;
    float shifted_coeffs = 0.0f;
};
// RoutingUpdater(topk, candidate_queue_size) (routing_updater.hpp:57-114).
// It only logs into its pivot's own row, so the device runs it fused with
// the merge and the pruning that refine_layer always applies next:
// propagate_engine_t::routing_loops takes both updaters.
template <DistanceMetricsT Metric, uint32_t Dim>
struct routing_updater_t {
    vertex_num_t topk = 64;
    vertex_num_t candidate_queue_size = 96;
};

// PropagateEngine (refiner/propagate_engine.hpp:38-325) on the GPU.  Every
// pass (next, next_range, each iteration of run) runs the updater on every
// row (or the window) and merges its logs in the same call, as next() does in
// the reference; the rows move from the bound graph to the device before the
// call and back after it (run keeps them there across its iterations).
// Rows must be sorted by (distance, vid) and hold at most max_nbr_size
// entries, as fill_single_layer_graph_from_layer leaves them.  Each merge
// keeps at most max_nbr_size entries per row (the arc updater that
// refine_layer always runs next).  The random updater draws from a seeded
// counter-based stream: set_seed picks the seed, every random pass takes the
// next stream.  Not offered: propagate without a merge and the log table.
template <DistanceMetricsT Metric, uint32_t Dim>
class propagate_engine_t {
  public:
    explicit propagate_engine_t(const dist_func_t<Metric, Dim>&) {}
    propagate_engine_t(const propagate_engine_t&) = delete;
    propagate_engine_t& operator=(const propagate_engine_t&) = delete;
    ~propagate_engine_t() {
        if (_r) artea_refine_destroy(_r);
        if (_d_vecs) artea_device_free(_d_vecs);
    }

    // Bind the graph; the vectors are copied to the device once per array.
    void set_graph(dynamic::single_layer_graph_t& g) {
        const vector_array_t& v = g.get_vecs_data();
        if (v.get_vec_dim() != Dim) throw std::runtime_error("single_layer_graph: vector dimension differs from Dim");
        if (_r) {
            artea_refine_destroy(_r);
            _r = nullptr;
        }
        if (_vec_src != &v || _vec_n != v.get_num_vecs()) {
            if (_d_vecs) artea_device_free(_d_vecs);
            _d_vecs = nullptr;
            const size_t bytes = std::max<size_t>(size_t(v.get_num_vecs()) * Dim * sizeof(float), 16);
            throw_on(artea_device_alloc(bytes, &_d_vecs));
            if (v.get_num_vecs())
                throw_on(artea_memcpy_htod(_d_vecs, v.get_all(), size_t(v.get_num_vecs()) * Dim * sizeof(float),
                                           nullptr));
            throw_on(artea_stream_synchronize(nullptr));
            _vec_src = &v;
            _vec_n = v.get_num_vecs();
        }
        const auto& l2g = g.local_to_global();
        throw_on(artea_refine_create(static_cast<const float*>(_d_vecs), v.get_num_vecs(), Dim, int(Metric),
                                     g.layer_config().max_nbr_size(), l2g.empty() ? nullptr : l2g.data(),
                                     g.get_num_vertices(), _seed, &_r));
        throw_on(artea_refine_set_join(_r, _join));
        _g = &g;
    }

    template <typename UpdaterT, typename... Args>
    UpdaterT make_updater(Args&&... args) {
        return UpdaterT{std::forward<Args>(args)...};
    }

    template <typename UpdaterT>
    propagate_engine_t& next(UpdaterT& u) {
        _upload();
        _apply(u, 0, _g->get_num_vertices());
        _download();
        return *this;
    }
    template <typename UpdaterT>
    propagate_engine_t& next_range(UpdaterT& u, vertex_num_t local_start, vertex_num_t local_end) {
        _upload();
        _apply(u, local_start, local_end);
        _download();
        return *this;
    }
    template <typename UpdaterT>
    propagate_engine_t& run(uint32_t num_iters, UpdaterT& u) {
        _upload();
        for (uint32_t i = 0; i < num_iters; ++i) _apply(u, 0, _g->get_num_vertices());
        _download();
        return *this;
    }
    // refine_layer's routing loops (index/artea_graph/index_factory.hpp
    // :400-404), which the reference writes as
    //   for (l = 0; l < loops; ++l) {
    //       next(routing_updater).next(pruning_updater);
    //       if (l != loops - 1) next(reverse_updater).next(arc_updater);
    //   }
    // Rows longer than max_nbr_size after the last loop keep their first
    // max_nbr_size entries (the writeback's truncation).
    propagate_engine_t& routing_loops(uint32_t loops, const routing_updater_t<Metric, Dim>& routing,
                                      const pruning_updater_t<Metric, Dim>& pruning) {
        _upload();
        throw_on(artea_refine_routing_loops(_r, loops, routing.topk, routing.candidate_queue_size,
                                            pruning.scale_coeffs));
        _download();
        return *this;
    }
    // Every pass merged its logs already; nothing is pending.
    size_t merge_logs() { return 0; }
    // Logs merged by the passes since the previous call.
    size_t get_merged_logs_count() {
        uint64_t logs = 0, exact = 0, approx = 0;
        throw_on(artea_refine_last_stats(_bound(), &logs, &exact, &approx));
        return size_t(logs);
    }

    // Device extras (no reference counterpart): the triangle join's pair
    // distances (ARTEA_JOIN_FP32 / ARTEA_JOIN_BF16, identical rows) and the
    // random stream's seed (takes effect at the next set_graph).
    void set_join(int precision) {
        _join = precision;
        if (_r) throw_on(artea_refine_set_join(_r, precision));
    }
    void set_seed(uint64_t seed) { _seed = seed; }

  private:
    artea_refine* _bound() {
        if (!_r || !_g) throw std::runtime_error("propagate_engine: set_graph was not called");
        return _r;
    }
    void _upload() {
        _bound();
        const auto& rows = _g->get_nbrs_arr();
        std::vector<uint32_t> counts(rows.size());
        size_t total = 0;
        for (size_t i = 0; i < rows.size(); ++i) total += (counts[i] = uint32_t(rows[i].size()));
        std::vector<Neighbor> flat;
        flat.reserve(std::max<size_t>(total, 1));
        for (const auto& r : rows) flat.insert(flat.end(), r.begin(), r.end());
        if (flat.empty()) flat.emplace_back();
        throw_on(artea_refine_upload_rows(_r, counts.data(), reinterpret_cast<const uint32_t*>(flat.data())));
    }
    void _download() {
        auto& rows = _g->get_nbrs_arr();
        std::vector<uint32_t> counts(rows.size());
        std::vector<Neighbor> flat(std::max<size_t>(size_t(rows.size()) * _g->layer_config().max_nbr_size(), 1));
        throw_on(artea_refine_download_rows(_r, counts.data(), reinterpret_cast<uint32_t*>(flat.data())));
        size_t pos = 0;
        for (size_t i = 0; i < rows.size(); ++i) {
            rows[i].assign(flat.begin() + pos, flat.begin() + pos + counts[i]);
            pos += counts[i];
        }
    }
    void _apply(const triangle_updater_t<Metric, Dim>& u, vertex_num_t s, vertex_num_t e) {
        throw_on(artea_refine_triangle_range(_r, u.scale_coeffs, u.shifted_coeffs, s, e));
    }
    void _apply(const reverse_updater_t<Metric, Dim>&, vertex_num_t s, vertex_num_t e) {
        throw_on(artea_refine_reverse_range(_r, s, e));
    }
    void _apply(const random_updater_t<Metric, Dim>& u, vertex_num_t s, vertex_num_t e) {
        throw_on(artea_refine_prefill(_r, u.rand_gen_size, u.start_vid, u.end_vid, s, e, _random_stream++));
    }
    void _apply(const truncate_updater_t<Metric, Dim>&, vertex_num_t, vertex_num_t) {
        throw_on(artea_refine_truncate(_r));
    }
    void _apply(const pruning_updater_t<Metric, Dim>& u, vertex_num_t s, vertex_num_t e) {
        throw_on(artea_refine_prune_range(_r, u.scale_coeffs, s, e));
    }

    dynamic::single_layer_graph_t* _g = nullptr;
    artea_refine* _r = nullptr;
    void* _d_vecs = nullptr;
    const vector_array_t* _vec_src = nullptr;
    vertex_num_t _vec_n = 0;
    uint64_t _seed = 0, _random_stream = 0;
    int _join = ARTEA_JOIN_FP32;
};

}  // namespace hip
}  // namespace artea

// C++ drop-in surface test of the refine path (artea_hip/artea.hpp), written
// like the reference's GoogleTest drivers: the triangle KATs of
// unit_tests/test_propagate_engine.cpp:235-318, its integrated random +
// reverse scenario (:318-400), the reverse-between-rounds regression of
// unit_tests/test_graph_build_regressions.cpp:97-121, row windows
// (next_range), and the config form of batch_query with the early-stop
// thresholds of index/stacked_rgraph/configs.hpp:208-284, all through
// propagate_engine_t / hierarchical_graph_router_t on the GPU.
// Prints "PASS <name>" per case; exit status = number of failed checks.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <string>
#include <vector>

#include "artea_hip/artea.hpp"

using namespace artea::hip;

static int g_failures = 0;
#define EXPECT_TRUE(c)                                                             \
    do {                                                                           \
        if (!(c)) {                                                                \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);               \
            ++g_failures;                                                          \
        }                                                                          \
    } while (0)
#define EXPECT_EQ(a, b) EXPECT_TRUE((a) == (b))

constexpr auto L2 = DistanceMetricsT::EUCLIDEAN;
constexpr uint32_t DIM = 16;
using engine_t = propagate_engine_t<L2, DIM>;
using triangle_t = triangle_updater_t<L2, DIM>;
using reverse_t = reverse_updater_t<L2, DIM>;
using random_t = random_updater_t<L2, DIM>;
using arc_t = arc_updater_t<L2, DIM>;

// The 3 x 4 grid of the reference's fixture: v = row * 4 + col at (col, row).
static vector_array_t grid() {
    vector_array_t v(12, DIM);
    for (uint32_t i = 0; i < 12; ++i) {
        v.get(i)[0] = float(i % 4);
        v.get(i)[1] = float(i / 4);
    }
    return v;
}

static float dist(const vector_array_t& v, uint32_t a, uint32_t b) {
    // dist_func_t's L2 on these exact small values: sqrt of the exact sum
    const float dx = v.get(a)[0] - v.get(b)[0], dy = v.get(a)[1] - v.get(b)[1];
    return std::sqrt(dx * dx + dy * dy);
}

static bool has_edge(const nbr_arr_t& row, uint32_t vid, float* d = nullptr) {
    for (const auto& e : row)
        if (e.get_vid() == vid) {
            if (d) *d = e.get_distance();
            return true;
        }
    return false;
}

static void set_row(dynamic::single_layer_graph_t& g, uint32_t u, const vector_array_t& v,
                    std::vector<uint32_t> ids) {
    std::sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) {
        const float da = dist(v, u, a), db = dist(v, u, b);
        return da != db ? da < db : a < b;
    });
    auto& row = g.fetch_nbrs(u);
    row.clear();
    for (uint32_t id : ids) row.emplace_back(id, dist(v, u, id), true);
}

static void triangle_prunes_using_farthest_retained_first() {
    // :235-259: pivot 0, retained 1 and 8, candidate 10; 8 -> 10 recommended.
    auto v = grid();
    dist_func_t<L2, DIM> df;
    dynamic::single_layer_graph_t g(v, sl_config_t(8));
    set_row(g, 0, v, {1, 8, 10});
    engine_t engine(df);
    engine.set_graph(g);
    auto u = engine.make_updater<triangle_t>(1.0f, 0.0f);
    engine.next(u);
    const auto& row = g.fetch_nbrs(0);
    EXPECT_EQ(row.size(), 2u);
    EXPECT_TRUE(row.size() == 2 && row[0].get_vid() == 1u && row[1].get_vid() == 8u);
    EXPECT_TRUE(row.size() == 2 && row[0].is_old() && row[1].is_old());
    // the log (row 8, vid 10, 2.0) was merged into row 8 as a new entry
    const auto& r8 = g.fetch_nbrs(8);
    EXPECT_EQ(r8.size(), 1u);
    EXPECT_TRUE(r8.size() == 1 && r8[0].get_vid() == 10u && r8[0].get_distance() == 2.0f && r8[0].is_new());
    EXPECT_TRUE(g.fetch_nbrs(1).empty());
    std::printf("PASS triangle_prunes_using_farthest_retained_first\n");
}

static void triangle_selects_closest_recommendation_without_pruning() {
    // :261-283: candidate 11, scale 2: recommendation to 1 at sqrt(8).
    auto v = grid();
    dist_func_t<L2, DIM> df;
    dynamic::single_layer_graph_t g(v, sl_config_t(8));
    set_row(g, 0, v, {1, 8, 11});
    engine_t engine(df);
    engine.set_graph(g);
    auto u = engine.make_updater<triangle_t>(2.0f, 0.0f);
    engine.next(u);
    const auto& row = g.fetch_nbrs(0);
    EXPECT_TRUE(row.size() == 3 && row[0].get_vid() == 1u && row[1].get_vid() == 8u && row[2].get_vid() == 11u);
    const auto& r1 = g.fetch_nbrs(1);
    EXPECT_TRUE(r1.size() == 1 && r1[0].get_vid() == 11u && r1[0].get_distance() == std::sqrt(8.0f));
    EXPECT_TRUE(g.fetch_nbrs(8).empty());
    std::printf("PASS triangle_selects_closest_recommendation_without_pruning\n");
}

static void triangle_uses_alpha_dependent_shift() {
    // :285-316: pivot 0, retained 2, candidate 3 at distance 1 from 2.
    struct Case {
        float alpha, shift;
        bool retain;
    };
    auto v = grid();
    dist_func_t<L2, DIM> df;
    for (const Case c : {Case{2.0f, 0.0f, false}, Case{2.0f, 0.2f, false}, Case{2.0f, 0.4f, true},
                         Case{2.0f, 1.0f / 3.0f, true}, Case{1.0f, 1.25f, true}, Case{1.0f, 1.5f, true},
                         Case{1.0f, 2.0f, true}}) {
        dynamic::single_layer_graph_t g(v, sl_config_t(8));
        set_row(g, 0, v, {2, 3});
        engine_t engine(df);
        engine.set_graph(g);
        auto u = engine.make_updater<triangle_t>(c.alpha, c.shift);
        engine.next(u);
        const auto& row = g.fetch_nbrs(0);
        EXPECT_TRUE(has_edge(row, 2));
        EXPECT_EQ(has_edge(row, 3), c.retain);
        EXPECT_EQ(row.size(), c.retain ? 2u : 1u);
        const auto& r2 = g.fetch_nbrs(2);
        EXPECT_TRUE(r2.size() == 1 && r2[0].get_vid() == 3u && r2[0].get_distance() == 1.0f);
    }
    std::printf("PASS triangle_uses_alpha_dependent_shift\n");
}

static void integrated_random_and_reverse() {
    // :318-400: random edges are asymmetric; one reverse pass makes every
    // edge bidirectional with equal distances.
    // The device merge keeps max_nbr_size entries per row (the arc updater
    // fused); 16 holds every row here (5 draws out, at most 11 in), as the
    // reference's uncapped merge does.
    auto v = grid();
    dist_func_t<L2, DIM> df;
    dynamic::single_layer_graph_t g(v, sl_config_t(16));
    engine_t engine(df);
    engine.set_seed(7);
    engine.set_graph(g);
    auto rnd = engine.make_updater<random_t>(5u, 0u, 12u);
    engine.run(1, rnd);
    size_t edges = 0, missing = 0;
    for (uint32_t u = 0; u < 12; ++u) {
        edges += g.fetch_nbrs(u).size();
        for (const auto& e : g.fetch_nbrs(u)) missing += has_edge(g.fetch_nbrs(e.get_vid()), u) ? 0 : 1;
        for (const auto& e : g.fetch_nbrs(u)) EXPECT_TRUE(e.get_vid() != u && e.get_vid() < 12u);
    }
    EXPECT_TRUE(edges > 0 && missing > 0);
    auto rev = engine.make_updater<reverse_t>();
    engine.run(1, rev);
    size_t after = 0, missing_after = 0, mismatch = 0;
    for (uint32_t u = 0; u < 12; ++u) {
        after += g.fetch_nbrs(u).size();
        for (const auto& e : g.fetch_nbrs(u)) {
            float back = 0.f;
            if (!has_edge(g.fetch_nbrs(e.get_vid()), u, &back)) {
                ++missing_after;
            } else if (back != e.get_distance()) {
                ++mismatch;
            }
        }
        EXPECT_TRUE(g.fetch_nbrs(u).size() <= 16u);
    }
    EXPECT_TRUE(after > edges);
    EXPECT_EQ(missing_after, 0u);
    EXPECT_EQ(mismatch, 0u);
    std::printf("PASS integrated_random_and_reverse\n");
}

static void reverse_edges_between_build_rounds() {
    // test_graph_build_regressions.cpp:97-121: one edge 0 -> 1, refine_layer
    // with build_loops 0, 1, 2 and no triangle passes: the reverse edge
    // appears only when a reverse pass runs between two rounds.
    constexpr uint32_t count = 32;
    vector_array_t v(count, DIM);
    v.get(1)[0] = 1.0f;
    dist_func_t<L2, DIM> df;
    for (uint32_t loops : {0u, 1u, 2u}) {
        dynamic::single_layer_graph_t g(v, sl_config_t(8));
        g.fetch_nbrs(0).push_back(nbr_t(1, 1.0f));
        engine_t engine(df);
        engine.set_graph(g);
        auto rev = engine.make_updater<reverse_t>();
        auto arc = engine.make_updater<arc_t>();
        for (uint32_t b = 0; b < loops; ++b) {  // index_factory.hpp:385-393 with triu_iters = 0
            if (b + 1 < loops) engine.next(rev);
            engine.next(arc);
        }
        EXPECT_EQ(g.fetch_nbrs(0)[0].get_vid(), 1u);
        EXPECT_EQ(g.fetch_nbrs(1).size(), loops == 2 ? 1u : 0u);
        if (loops == 2) EXPECT_EQ(g.fetch_nbrs(1)[0].get_vid(), 0u);
    }
    std::printf("PASS reverse_edges_between_build_rounds\n");
}

static void next_range_limits_the_updater_to_the_window() {
    // Rows outside the window are not pruned; logs still merge anywhere.
    auto v = grid();
    dist_func_t<L2, DIM> df;
    dynamic::single_layer_graph_t g(v, sl_config_t(8));
    set_row(g, 0, v, {1, 8, 10});
    set_row(g, 5, v, {4, 6, 9, 1});
    engine_t engine(df);
    engine.set_graph(g);
    auto u = engine.make_updater<triangle_t>(1.0f, 0.0f);
    engine.next_range(u, 1, 12);  // row 0 outside
    EXPECT_EQ(g.fetch_nbrs(0).size(), 3u);
    EXPECT_TRUE(g.fetch_nbrs(0)[2].is_new());
    engine.next_range(u, 0, 1);  // now row 0
    EXPECT_EQ(g.fetch_nbrs(0).size(), 2u);
    EXPECT_TRUE(has_edge(g.fetch_nbrs(8), 10));
    // reverse over rows [0, 1): only row 0's edges are reversed
    dynamic::single_layer_graph_t h(v, sl_config_t(8));
    h.fetch_nbrs(0).push_back(nbr_t(3, dist(v, 0, 3)));
    h.fetch_nbrs(6).push_back(nbr_t(7, dist(v, 6, 7)));
    engine.set_graph(h);
    auto rev = engine.make_updater<reverse_t>();
    engine.next_range(rev, 0, 1);
    EXPECT_TRUE(has_edge(h.fetch_nbrs(3), 0));
    EXPECT_TRUE(!has_edge(h.fetch_nbrs(7), 6));
    std::printf("PASS next_range_limits_the_updater_to_the_window\n");
}

static void sparse_layer_rows_hold_global_ids() {
    // A sparse layer over vids {0, 2, 3, 8, 10}: local rows, global ids.
    auto v = grid();
    dist_func_t<L2, DIM> df;
    std::vector<uint32_t> l2g = {0, 2, 3, 8, 10}, g2l(12, 0xFFFFFFFFu);
    for (uint32_t i = 0; i < l2g.size(); ++i) g2l[l2g[i]] = i;
    dynamic::single_layer_graph_t g(v, sl_config_t(4), l2g, g2l);
    set_row(g, 0, v, {2, 3});  // storage vid 0 = local row 0
    engine_t engine(df);
    engine.set_graph(g);
    auto u = engine.make_updater<triangle_t>(2.0f, 0.0f);
    engine.next(u);
    EXPECT_EQ(g.fetch_nbrs(0).size(), 1u);   // 3 pruned by 2
    EXPECT_TRUE(has_edge(g.fetch_nbrs(2), 3));  // recommendation into local row of vid 2
    EXPECT_EQ(g.get_nbrs_arr()[1].size(), 1u);
    std::printf("PASS sparse_layer_rows_hold_global_ids\n");
}

static void rgraph_thresholds_closed_forms_and_errors() {
    // gamma = (alpha + 1) / (alpha - 1) + 1 = 3 at alpha = 3; tau rho = 1.5;
    // R_0 = 1, R_h = 1.5 * 2^(h-1) at tau_k = 0.5, beta = 2.
    rgraph_config_t rc(2.0f, 0.5f, 1.5f, 1.0f, 64);
    pruning_config_t pc(3.0f, 1.5f, 1.0f);
    EXPECT_EQ(rc.radius_at(0), 1.0f);
    EXPECT_EQ(rc.radius_at(1), 1.5f);
    EXPECT_EQ(rc.radius_at(3), 6.0f);
    EXPECT_EQ(rc.early_stop_threshold(1, pc), 12.0f);  // 3 (1.5 + 1 + 1.5)
    EXPECT_EQ(rc.early_stop_threshold(2, pc), 21.0f);  // 3 (1.5 + 1 + 1.5 + 3)
    EXPECT_EQ(rc.early_stop_threshold(3, pc), 39.0f);  // 3 (1.5 + 1 + 1.5 + 3 + 6)
    // ARC Delta_1 = (gamma + 1) psi + gamma R_2 / R_1, psi = 4 / 1.5
    EXPECT_EQ(rc.aspect_ratio_constraint(1, pc), float(4.0 * (4.0 / 1.5) + 3.0 * 2.0));
    EXPECT_EQ(rgraph_config_t::compute_max_allowed_level_id(1000000), 7u);
    EXPECT_EQ(rgraph_config_t::compute_max_allowed_level_id(100000), 5u);
    EXPECT_EQ(rgraph_config_t::compute_max_allowed_level_id(900), 1u);
    int errors = 0;
    try { (void)rc.early_stop_threshold(0, pc); } catch (const std::out_of_range&) { ++errors; }
    try { (void)rc.early_stop_threshold(1, pruning_config_t(1.0f, 0.0f)); } catch (const std::invalid_argument&) { ++errors; }
    try { (void)rgraph_config_t(2.0f, 0.f, 0.f, 1.f, 8).early_stop_threshold(200, pc); } catch (const std::overflow_error&) { ++errors; }
    try { rgraph_config_t bad(1.0f, 0.f, 0.f, 1.f, 8); } catch (const std::runtime_error& e) { ++errors; EXPECT_TRUE(std::string(e.what()).find("rnet_beta") != std::string::npos); }
    EXPECT_EQ(errors, 4);
    std::printf("PASS rgraph_thresholds_closed_forms_and_errors\n");
}

static void batch_query_with_configs_equals_explicit_thresholds() {
    // A path graph 0 - 1 - ... - 39 on two levels; the config overload's
    // thresholds equal _early_stop_thresholds' values (squared for L2SQR)
    // and give the same results as passing them explicitly.
    constexpr uint32_t n = 40, dim = 128;
    vector_array_t vecs(n, dim);
    for (uint32_t i = 0; i < n; ++i) vecs.get(i)[0] = float(i);
    std::vector<uint32_t> vi(2 * n);
    uint32_t c0 = 0, c1 = 0;
    std::vector<uint32_t> highest(n);
    for (uint32_t i = 0; i < n; ++i) {
        highest[i] = (i % 4 == 0) ? 1u : 0u;
        vi[2 * i] = highest[i];
        vi[2 * i + 1] = highest[i] ? c1++ : c0++;
    }
    std::vector<std::vector<uint64_t>> go(2);
    std::vector<uint32_t> nbrs;
    go[0].push_back(0);
    for (uint32_t i = 0; i < n; ++i) {
        if (highest[i]) continue;
        if (i > 0) nbrs.push_back(i - 1);
        if (i + 1 < n) nbrs.push_back(i + 1);
        go[0].push_back(nbrs.size());
    }
    go[1].push_back(nbrs.size());
    for (uint32_t i = 0; i < n; ++i) {
        if (!highest[i]) continue;
        if (i >= 4) nbrs.push_back(i - 4);  // level 1
        if (i + 4 < n) nbrs.push_back(i + 4);
        go[1].push_back(nbrs.size());
        if (i > 0) nbrs.push_back(i - 1);  // level 0
        if (i + 1 < n) nbrs.push_back(i + 1);
        go[1].push_back(nbrs.size());
    }
    compact::hierarchical_graph_t g(n, 1, 4, 4, 0, vi, go, nbrs);
    rgraph_config_t rc(2.0f, 0.5f, 1.5f, 1.0f, 64);
    pruning_config_t pc(3.0f, 1.5f, 1.0f);
    using router_t = hierarchical_graph_router_t<DistanceMetricsT::EUCLIDEAN_SQR, dim>;
    const auto thr = router_t::early_stop_thresholds(g, rc, pc);
    EXPECT_EQ(thr.size(), 2u);
    EXPECT_EQ(thr[1], 144.0f);  // 12^2
    dist_func_t<DistanceMetricsT::EUCLIDEAN_SQR, dim> df;
    router_t router(vecs, df, 3, 8);
    vector_array_t qs(3, dim);
    qs.get(0)[0] = 30.2f;
    qs.get(1)[0] = 17.0f;
    qs.get(2)[0] = 3.5f;
    const auto a = router.batch_query<false, true>(qs, g, rc, pc);
    const auto b = router.batch_query<false, true>(qs, g, thr);
    const auto c = router.batch_query<false, false>(qs, g, rc, pc);
    const auto d = router.batch_query(qs, g);
    EXPECT_TRUE(a == b);
    EXPECT_TRUE(c == d);
    EXPECT_EQ(a[0].get_vid(), 30u);
    EXPECT_EQ(d[3].get_vid(), 17u);
    std::printf("PASS batch_query_with_configs_equals_explicit_thresholds\n");
}

static void pruning_updater_reference_kats() {
    // unit_tests/test_exact_artea_arc_candidates.cpp:382-399: data {0, 1,
    // 1.25}, scale 2, shift 1.5 (unused by the refiner's scaled rule).
    vector_array_t v(3, DIM);
    v.get(1)[0] = 1.0f;
    v.get(2)[0] = 1.25f;
    dist_func_t<L2, DIM> df;
    dynamic::single_layer_graph_t g(v, sl_config_t(4));
    engine_t engine(df);
    engine.set_graph(g);
    auto prune = engine.make_updater<pruning_updater_t<L2, DIM>>(2.0f, 1.5f);
    g.fetch_nbrs(0) = {Neighbor(1, 1.0f, false), Neighbor(2, 1.25f, false)};
    engine.next(prune);
    EXPECT_EQ(g.fetch_nbrs(0).size(), 2u);  // two old entries: the pair is skipped
    g.fetch_nbrs(0) = {Neighbor(1, 1.0f, true), Neighbor(2, 1.25f, true)};
    engine.next(prune);
    const auto& row = g.fetch_nbrs(0);
    EXPECT_TRUE(row.size() == 1 && row[0].get_vid() == 1u && row[0].is_old());
    g.fetch_nbrs(0).clear();
    engine.next(prune);
    EXPECT_TRUE(g.fetch_nbrs(0).empty());
    std::printf("PASS pruning_updater_reference_kats\n");
}

static void routing_loop_grid_closed_form() {
    // Rows of the chain 0-1-...-11 over the grid; one routing loop with
    // topk = queue = 12 merges every other vertex into each row, and the
    // scale-1 pruning leaves exactly the grid's axis neighbours.
    auto v = grid();
    dist_func_t<L2, DIM> df;
    dynamic::single_layer_graph_t g(v, sl_config_t(16));
    for (uint32_t u = 0; u < 12; ++u) {
        std::vector<uint32_t> ids;
        if (u > 0) ids.push_back(u - 1);
        if (u < 11) ids.push_back(u + 1);
        set_row(g, u, v, ids);
        for (auto& e : g.fetch_nbrs(u)) e.mark_as_old();
    }
    engine_t engine(df);
    engine.set_graph(g);
    auto routing = engine.make_updater<routing_updater_t<L2, DIM>>(12u, 12u);
    auto prune = engine.make_updater<pruning_updater_t<L2, DIM>>(1.0f, 0.0f);
    engine.routing_loops(1, routing, prune);
    for (uint32_t u = 0; u < 12; ++u) {
        const uint32_t r = u / 4, c = u % 4;
        std::vector<uint32_t> want;
        if (r > 0) want.push_back(u - 4);
        if (c > 0) want.push_back(u - 1);
        if (c < 3) want.push_back(u + 1);
        if (r < 2) want.push_back(u + 4);
        std::vector<uint32_t> got;
        for (const auto& e : g.fetch_nbrs(u)) got.push_back(e.get_vid());
        std::sort(got.begin(), got.end());
        EXPECT_TRUE(got == want);
    }
    std::printf("PASS routing_loop_grid_closed_form\n");
}

int main() {
    try {
        rgraph_thresholds_closed_forms_and_errors();
        triangle_prunes_using_farthest_retained_first();
        triangle_selects_closest_recommendation_without_pruning();
        triangle_uses_alpha_dependent_shift();
        integrated_random_and_reverse();
        reverse_edges_between_build_rounds();
        next_range_limits_the_updater_to_the_window();
        sparse_layer_rows_hold_global_ids();
        batch_query_with_configs_equals_explicit_thresholds();
        pruning_updater_reference_kats();
        routing_loop_grid_closed_form();
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        ++g_failures;
    }
    return g_failures;
}

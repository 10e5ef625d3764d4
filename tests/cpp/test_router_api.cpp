// C++ drop-in surface test (artea_hip/artea.hpp), written like the
// reference's GoogleTest drivers: the same scenarios as
// unit_tests/test_compact_csr_query.cpp:20-87 and
// unit_tests/test_router_early_stop.cpp:64-212, through
// hierarchical_graph_router_t / hierarchical_graph_file_manager_t on the GPU.
// Prints "PASS <name>" per case; exit status = number of failed checks.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
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

constexpr auto L2SQR = DistanceMetricsT::EUCLIDEAN_SQR;

// Grouped CSR for highest[v] and rows[(v, level)], buckets in ascending vid.
static compact::hierarchical_graph_t make_graph(const std::vector<uint32_t>& highest,
                                                const std::vector<std::vector<std::vector<uint32_t>>>& rows,
                                                uint32_t ul, uint32_t bl, uint32_t entry) {
    const uint32_t n = highest.size();
    uint32_t top = 0;
    for (auto h : highest) top = std::max(top, h);
    std::vector<uint32_t> vi(2 * n), count(top + 1, 0);
    for (uint32_t v = 0; v < n; ++v) {
        vi[2 * v] = highest[v];
        vi[2 * v + 1] = count[highest[v]]++;
    }
    std::vector<std::vector<uint64_t>> go(top + 1);
    std::vector<uint32_t> nbrs;
    for (uint32_t h = 0; h <= top; ++h) {
        go[h].assign(size_t(count[h]) * (h + 1) + 1, 0);
        for (uint32_t v = 0; v < n; ++v) {
            if (highest[v] != h) continue;
            for (uint32_t l = h;; --l) {
                go[h][size_t(vi[2 * v + 1]) * (h + 1) + (h - l)] = nbrs.size();
                nbrs.insert(nbrs.end(), rows[v][l].begin(), rows[v][l].end());
                if (l == 0) break;
            }
        }
        go[h].back() = nbrs.size();
    }
    return compact::hierarchical_graph_t(n, top, ul, bl, entry, vi, go, nbrs);
}

static void greedy_top10_matches_exact_reference_and_restored_batch() {
    constexpr uint32_t dim = 128, n = 80, topk = 10;
    vector_array_t vectors(n, dim);
    std::vector<uint32_t> highest(n);
    std::vector<std::vector<std::vector<uint32_t>>> rows(n);
    for (uint32_t v = 0; v < n; ++v) {
        vectors.get(v)[0] = float(v);
        highest[v] = v < 40 ? 1 : 0;
        rows[v].resize(highest[v] + 1);
        rows[v][0].resize(n);
        std::iota(rows[v][0].begin(), rows[v][0].end(), 0);
        if (highest[v] == 1) {
            rows[v][1].resize(40);
            std::iota(rows[v][1].begin(), rows[v][1].end(), 0);
        }
    }
    auto graph = make_graph(highest, rows, n, n, 0);
    const auto dir = std::filesystem::temp_directory_path() / "artea_hip_cpp_test";
    std::filesystem::create_directories(dir);
    const auto snap = (dir / "greedy.graph").string();
    hierarchical_graph_file_manager_t::snapshot(graph, snap);
    auto restored = hierarchical_graph_file_manager_t::restore(snap);
    dist_func_t<L2SQR, dim> distance;
    hierarchical_graph_router_t<L2SQR, dim> router(vectors, distance, topk, n);
    router.initialize();
    vector_array_t queries(5, dim);
    const float coords[5] = {-2.0f, 12.25f, 39.5f, 60.75f, 90.0f};
    for (uint32_t i = 0; i < 5; ++i) queries.get(i)[0] = coords[i];
    for (uint32_t i = 0; i < 5; ++i) {
        auto d = [&](uint32_t u) { float x = coords[i] - float(u); return x * x; };
        std::vector<uint32_t> expected(n);
        std::iota(expected.begin(), expected.end(), 0);
        std::sort(expected.begin(), expected.end(),
                  [&](uint32_t a, uint32_t b) { return d(a) < d(b) || (d(a) == d(b) && a < b); });
        auto original = router.query(queries.get(i), graph);
        auto from_file = router.query(queries.get(i), restored);
        EXPECT_EQ(original.size(), topk);
        EXPECT_EQ(from_file.size(), topk);
        for (uint32_t r = 0; r < topk; ++r) {
            EXPECT_EQ(original[r].get_vid(), expected[r]);
            EXPECT_EQ(original[r].get_distance(), d(expected[r]));
            EXPECT_EQ(from_file[r], original[r]);
        }
    }
    auto batch = router.batch_query(queries, restored);
    EXPECT_EQ(batch.size(), size_t(5) * topk);
    for (uint32_t i = 0; i < 5; ++i) {
        auto single = router.query(queries.get(i), graph);
        for (uint32_t r = 0; r < topk; ++r) EXPECT_EQ(batch[i * topk + r], single[r]);
    }
    std::printf("PASS greedy_top10_matches_exact_reference_and_restored_batch\n");
}

static void early_stop_and_errors() {
    constexpr uint32_t dim = 16;
    vector_array_t data(4, dim);
    const float xs[4] = {12, 8, 4, 0};
    for (uint32_t v = 0; v < 4; ++v) data.get(v)[0] = xs[v];
    std::vector<uint32_t> highest = {2, 1, 1, 1};
    std::vector<std::vector<std::vector<uint32_t>>> rows(4);
    for (uint32_t v = 0; v < 4; ++v) {
        rows[v].resize(highest[v] + 1);
        for (uint32_t o = 0; o < 4; ++o)
            if (o != v) rows[v][0].push_back(o);
        if (v < 3) rows[v][1] = {v + 1};
    }
    auto g = make_graph(highest, rows, 4, 4, 0);
    dist_func_t<L2SQR, dim> dist;
    hierarchical_graph_router_t<L2SQR, dim> router(data, dist, 1, 4);
    EXPECT_EQ(router.query(data.get(3), g).front().get_vid(), 3u);
    auto early = router.query<false, true>(data.get(3), g, {0.0f, 64.0f, 256.0f});
    EXPECT_EQ(early.front().get_vid(), 3u);
    EXPECT_EQ(early.front().get_distance(), 0.0f);
    auto batch = router.batch_query(data, g);
    for (uint32_t v = 0; v < 4; ++v) EXPECT_EQ(batch[v].get_vid(), v);
    bool threw = false;
    try {
        router.query<false, true>(data.get(3), g, {0.0f, 0.0f, 256.0f});
    } catch (const std::overflow_error&) {
        threw = true;
    }
    EXPECT_TRUE(threw);
    threw = false;
    try {
        hierarchical_graph_router_t<L2SQR, dim> bad(data, dist, 5, 4);
    } catch (const std::runtime_error& e) {
        threw = std::string(e.what()).find("must be >= topk") != std::string::npos;
    }
    EXPECT_TRUE(threw);
    threw = false;
    try {
        hierarchical_graph_file_manager_t::restore("/nonexistent/x.graph");
    } catch (const std::runtime_error& e) {
        threw = std::string(e.what()).find("failed to open") != std::string::npos;
    }
    EXPECT_TRUE(threw);
    std::printf("PASS early_stop_and_errors\n");
}

int main() {
    greedy_top10_matches_exact_reference_and_restored_batch();
    early_stop_and_errors();
    std::printf("%s (%d failures)\n", g_failures ? "FAILED" : "OK", g_failures);
    return g_failures;
}

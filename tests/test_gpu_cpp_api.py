"""Runs the C++ drop-in-surface tests (tests/cpp/test_router_api.cpp and
test_propagate_api.cpp, built by __graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "test_router_api")


@pytest.mark.gpu
def test_cpp_router_api():
    assert os.path.exists(BIN), "C++ test binary not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS greedy_top10_matches_exact_reference_and_restored_batch" in r.stdout
    assert "PASS early_stop_and_errors" in r.stdout


@pytest.mark.gpu
def test_cpp_propagate_api():
    """propagate_engine_t and the config form of batch_query through the
    C++ drop-in header (tests/cpp/test_propagate_api.cpp)."""
    binp = os.path.join(ROOT, "tests", "cpp", "build", "test_propagate_api")
    assert os.path.exists(binp), "C++ test binary not built (run __graft_entry__.build())"
    r = subprocess.run([binp], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    for case in ("rgraph_thresholds_closed_forms_and_errors", "triangle_prunes_using_farthest_retained_first",
                 "triangle_selects_closest_recommendation_without_pruning", "triangle_uses_alpha_dependent_shift",
                 "integrated_random_and_reverse", "reverse_edges_between_build_rounds",
                 "next_range_limits_the_updater_to_the_window", "sparse_layer_rows_hold_global_ids",
                 "batch_query_with_configs_equals_explicit_thresholds", "pruning_updater_reference_kats",
                 "routing_loop_grid_closed_form"):
        assert "PASS " + case in r.stdout, case


def test_cpp_api_header_compiles(tmp_path):
    # The C++ surface compiles against the C-ABI header with the host compiler.
    src = tmp_path / "t.cpp"
    src.write_text('#include "artea_hip/artea.hpp"\n'
                   'int main() { artea::hip::dist_func_t<artea::hip::DistanceMetricsT::EUCLIDEAN_SQR, 128> d;'
                   ' (void)d; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

"""bench.py's command line without a GPU: a request for more than one GPU
exits with a message (every box here has one MI355X and the path uses no
collectives), and the workloads and options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode != 0 and r.stdout == ""
    assert "one GPU" in r.stderr or "one MI355X" in r.stderr


def test_help_lists_the_workloads_and_timing_options():
    r = run("--help")
    assert r.returncode == 0
    for word in ("search", "cosine", "build", "c0", "--steps", "--warmup", "--streams", "--serve-lanes",
                 "--serve-budget"):
        assert word in r.stdout

"""bench.py's command line, without a GPU: the one-GPU refusal and the
workload switch."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_graph_choices_listed():
    r = run_bench("--help")
    assert r.returncode == 0
    assert "{rmat,er,grid}" in r.stdout


def test_unknown_graph_rejected():
    r = run_bench("--graph", "kron")
    assert r.returncode == 2
    assert "invalid choice" in r.stderr

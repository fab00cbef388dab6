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


def test_traversal_bytes_hand_computed():
    """The executed-byte model on a four-level record sequence, n = 128
    (W = 16 bytes per bitmap), computed by hand from its docstring: init
    4*128 + 5*16 = 592; L0 scan-free, no sweep: 16 + (16 + 24 + 12) + 36*3 =
    176; L1 bottom-up: 16 + 4*16 + 12*100 + 16*20 + 8*50 + 4*40 = 2160; L2
    single-scan after bottom-up, bitmap sweep: 16 + (16 + 800) + (640 + 1600
    + 24) = 3096; L3 scan-free after single-scan, sweep from the levels: 16 +
    (512 + 32 + 120) + (96 + 96 + 8) + 72 = 952.  With parents: + 4*128 +
    4*(3 + 40 + 6 + 2)."""
    sys.path.insert(0, ROOT)
    import bench
    recs = [dict(mode=0, n_f=1, m_f=3, bu_cand=0, bu_scan_v=0, bu_scan_e=0),
            dict(mode=2, n_f=3, m_f=10, bu_cand=100, bu_scan_v=20, bu_scan_e=50),
            dict(mode=1, n_f=40, m_f=200, bu_cand=0, bu_scan_v=0, bu_scan_e=0),
            dict(mode=0, n_f=6, m_f=12, bu_cand=0, bu_scan_v=0, bu_scan_e=0)]
    assert bench.traversal_bytes(recs, 128, 52) == 592 + 176 + 2160 + 3096 + 952
    assert bench.traversal_bytes(recs, 128, 52, parents=True) == 6976 + 512 + 4 * 51
    assert bench.traversal_bytes([], 128, 1) == 592

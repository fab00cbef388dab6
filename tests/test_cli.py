"""The ./bfs driver (xbfs_amd/bfs): the reference's 11-argument command line
(bfs_gpu_update.cu:33-48), its host-side report lines and its three output
files.  The argument and loader paths need no GPU; the traversal run is
marked gpu."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from xbfs_amd import Graph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BFS = os.path.join(ROOT, "xbfs_amd", "bfs")


def run(args, cwd, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([BFS] + [str(a) for a in args], cwd=cwd, capture_output=True,
                          text=True, env=e, timeout=300)


def graph_files(tmp_path, g):
    pre = str(tmp_path / "g")
    g.write(pre)
    return pre + "_beg_pos.bin", pre + "_csr.bin", pre + "_weight.bin"


def test_wrong_argument_count(tmp_path):
    for args in ([], ["a"] * 10, ["a"] * 12):
        p = run(args, tmp_path)
        assert p.returncode == 255  # main returns -1
        assert p.stdout.splitlines()[-1] == "Wrong input"


def test_missing_graph_file(tmp_path):
    p = run(["nope_beg.bin", "nope_csr.bin", "w", 256, 256, 0, "r.txt", 0.1, 0, "d", 1],
            tmp_path)
    assert p.returncode == 1
    assert "graph load failed" in p.stdout


def test_host_report_matches_reference_statistics(tmp_path):
    """The degree histogram (bfs_gpu_update.cu:76-89), workload_gap1's line
    (workload_gap.cuh:27-38) and hub_num (bfs_adaptiveFQ_wb_async.cuh:953-958)
    are printed before any GPU work; an out-of-range source stops there."""
    g = Graph.rmat(10, 8, seed=33)
    beg, csr, w = graph_files(tmp_path, g)
    p = run([beg, csr, w, 256, 256, 0, "r.txt", 0.1, g.n + 5, "rmat10", 1], tmp_path)
    assert p.returncode == 1 and "out of range" in p.stdout
    out = p.stdout
    deg = g.degrees()
    small = int(((deg >= 0) & (deg <= 32)).sum())
    medium = int(((deg > 32) & (deg <= 256)).sum())
    large = int((deg > 256).sum())
    assert f"small degree counts: {small} medium degree counts: {medium} " \
           f"large degree counts: {large}" in out
    assert f"Graph load (success): {g.n} verts, {g.m} edges" in out
    assert "Weight file cannot open" in out
    m = re.search(r"vertexID = (\d+) maximum degree of graph is (\d+)", out)
    assert int(m.group(2)) == int(deg.max())
    assert f"hub_num = {int((deg > 1000).sum())}" in out
    ref = O.ref_loader(beg, csr)
    if ref is not None:  # the reference's own loader and workload_gap1
        n, mm, mx, vid, _, _ = ref
        assert (n, mm, mx, vid) == (g.n, g.m, int(m.group(2)), int(m.group(1)))


@pytest.mark.gpu
@pytest.mark.parametrize("env", [{}, {"XBFS_PERSISTENT": "0"}, {"XBFS_POLICY": "learned"},
                                 {"XBFS_POLICY": "bu"}])
def test_traversal_run_and_outputs(tmp_path, env):
    """RMAT scale 16, edge factor 16: exit status 0, zero mismatches, levels
    equal to the oracle's, the per-level tag line equal to the reference
    heuristic's restatement (heuristic policy), and one line appended to
    each output file per run."""
    g = Graph.rmat(16, 16, seed=1016)
    beg, csr, w = graph_files(tmp_path, g)
    src = int(np.flatnonzero(g.degrees() > 0)[7])
    lvf = str(tmp_path / "levels.bin")
    e = dict(env, XBFS_LEVELS_OUT=lvf)
    for rep in range(2):
        p = run([beg, csr, w, 256, 256, 0, "res.txt", 0.1, src, "rmat16", 1], tmp_path, e)
        print("[case]", env, rep, p.stdout[-400:], flush=True)
        assert p.returncode == 0, p.stdout + p.stderr
        mm = re.search(r"Current level = (\d+) (\d+) padding=1", p.stdout)
        assert mm and int(mm.group(2)) == 0
        assert "Now we finish!" in p.stdout
    want, _, _ = O.bfs_queue(g.beg_pos, g.adj, src)
    got = np.fromfile(lvf, dtype=np.int32)
    assert (got == want).all()
    tags_line = open(tmp_path / "Hybrid_iteration.csv").read()
    assert tags_line.count("rmat16: ") == 2
    if not env:
        got_tags = [int(t) for t in tags_line.split("rmat16: ")[1].split()]
        ref, _, _ = O.reference_tags(g.beg_pos, g.adj, src, 0.1)
        ref = [int(t) for t in ref]
        k = min(len(got_tags), len(ref))
        assert got_tags[:k] == ref[:k] and len(got_tags) >= len(ref) - 1
    times = open(tmp_path / "time_Baseline+adaptive+FQ+wb+async_P6000_Camera_more_synch.csv"
                 ).read().splitlines()
    assert len(times) == 2
    for line in times:
        name, td, bu, tot = line.split(";")
        assert name == "rmat16" and abs(float(td) + float(bu) - float(tot)) < 1e-6 * max(1, float(tot))
    assert open(tmp_path / "res.txt").read() == "\n\n"


@pytest.mark.gpu
def test_per_level_records_csv(tmp_path):
    """XBFS_RECORDS_OUT: one CSV row per level with the frontier statistics;
    n_f and m_f of each level equal the oracle's vertex and degree counts of
    that level (the reference's count[L], bfs_adaptiveFQ_wb_async.cuh:1042-1046)."""
    import csv
    g = Graph.rmat(14, 16, seed=1014)
    beg, csr, w = graph_files(tmp_path, g)
    src = int(np.flatnonzero(g.degrees() > 0)[3])
    out = tmp_path / "records.csv"
    for pers in ("1", "0"):
        p = run([beg, csr, w, 256, 256, 0, "res.txt", 0.1, src, "rmat14", 1], tmp_path,
                {"XBFS_RECORDS_OUT": str(out), "XBFS_PERSISTENT": pers})
        assert p.returncode == 0, p.stdout + p.stderr
    rows = list(csv.DictReader(open(out)))
    assert list(rows[0].keys()) == ["dataset", "src", "level", "mode", "tag", "n_f", "m_f", "n_u",
                                    "m_u", "ratio", "t_s", "bu_cand", "bu_scan_v", "bu_scan_e"]
    want, _, _ = O.bfs_queue(g.beg_pos, g.adj, src)
    cnt, nv, _ = O.level_counts(g.beg_pos, want)
    deg = g.degrees()
    # bottom-up candidates at level L: the unvisited vertices of the source's
    # component when it is the largest-degree vertex's, else every unvisited
    # non-isolated vertex
    hub_comp = O.bfs_queue(g.beg_pos, g.adj, int(np.argmax(deg)))[0] >= 0
    runs = [r for r in rows if r["level"] == "0"]
    assert len(runs) == 2
    n_bu = 0
    for r in rows:
        L = int(r["level"])
        assert r["dataset"] == "rmat14" and int(r["src"]) == src
        assert int(r["n_f"]) == nv[L] and int(r["m_f"]) == cnt[L]
        assert int(r["tag"]) == {0: 0, 1: 1, 2: -1}[int(r["mode"])]
        assert abs(float(r["ratio"]) - cnt[L] / g.m) < 1e-6
        assert float(r["t_s"]) > 0
        cand, sv, se = int(r["bu_cand"]), int(r["bu_scan_v"]), int(r["bu_scan_e"])
        if int(r["mode"]) == 2:
            n_bu += 1
            unv = (want > L) | ((want == -1) & (deg > 0) & ~hub_comp[src])
            assert cand == unv.sum()
            assert 0 <= sv <= cand and 0 <= se <= deg[unv].sum()
        else:
            assert cand == sv == se == 0
    assert n_bu > 0

"""Host-side graph code of libxbfs: generators, binary CSR I/O (the
reference's graph.hpp rules), text edge lists, degree statistics, and the
C-ABI symbol table.  CPU only."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from tests import graphs as G
from xbfs_amd import Graph, XbfsError, _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def csr_invariants(g):
    beg, adj = g.beg_pos, g.adj
    assert beg[0] == 0 and beg[-1] == g.m and (np.diff(beg) >= 0).all()
    assert adj.min(initial=0) >= 0 and adj.max(initial=0) < g.n
    u = np.repeat(np.arange(g.n), np.diff(beg))
    assert not (u == adj).any(), "self loop"
    key = u.astype(np.int64) * g.n + adj
    assert (np.diff(key) > 0).all(), "rows not strictly sorted (duplicate or unsorted)"
    rkey = adj.astype(np.int64) * g.n + u
    assert np.array_equal(np.sort(rkey), key), "not symmetric"


@pytest.mark.parametrize("scale,ef,seed", [(8, 4, 1), (12, 16, 1012), (14, 8, 5)])
def test_rmat_properties_and_determinism(scale, ef, seed):
    g = Graph.rmat(scale, ef, seed=seed)
    assert g.n == 1 << scale
    csr_invariants(g)
    g2 = Graph.rmat(scale, ef, seed=seed)
    assert np.array_equal(g.adj, g2.adj) and np.array_equal(g.beg_pos, g2.beg_pos)
    g3 = Graph.rmat(scale, ef, seed=seed + 1)
    assert not np.array_equal(g.beg_pos, g3.beg_pos)
    # RMAT skew: the maximum degree is far above the mean
    assert g.degrees().max() > 4 * g.m / g.n


def test_rmat_relabel_is_a_permutation():
    a = Graph.rmat(11, 8, seed=9, relabel=False)
    b = Graph.rmat(11, 8, seed=9, relabel=True)
    assert a.m == b.m
    assert np.array_equal(np.sort(a.degrees()), np.sort(b.degrees()))


def test_er_and_grid():
    g = Graph.er(5000, 20000, seed=7)
    csr_invariants(g)
    assert 39000 <= g.m <= 40000  # 2m minus a few self-loops / duplicates
    g = Graph.grid(5, 7)
    csr_invariants(g)
    beg, adj = G.grid(5, 7)
    assert np.array_equal(g.beg_pos, beg) and np.array_equal(g.adj, adj)
    assert g.m == 2 * (5 * 6 + 4 * 7)


def test_write_load_roundtrip(tmp_path):
    g = Graph.rmat(10, 8, seed=3)
    pre = str(tmp_path / "g")
    g.write(pre)
    h = Graph.load(pre + "_beg_pos.bin", pre + "_csr.bin", pre + "_weight.bin")
    assert np.array_equal(g.beg_pos, h.beg_pos) and np.array_equal(g.adj, h.adj)
    assert os.path.getsize(pre + "_beg_pos.bin") == 8 * (g.n + 1)
    assert os.path.getsize(pre + "_csr.bin") == 4 * g.m


def test_load_errors(tmp_path):
    g = Graph.grid(4, 4)
    pre = str(tmp_path / "g")
    g.write(pre)
    with pytest.raises(XbfsError) as e:
        Graph.load(pre + "_missing.bin", pre + "_csr.bin")
    assert e.value.code == 4
    # truncated adjacency file (graph.hpp:61 asserts ret==edge_count)
    with open(pre + "_csr.bin", "r+b") as f:
        f.truncate(4 * g.m - 4)
    with pytest.raises(XbfsError) as e:
        Graph.load(pre + "_beg_pos.bin", pre + "_csr.bin")
    assert e.value.code == 4
    # beg_pos[n] == 0 (graph.hpp:37 asserts > 0)
    np.zeros(5, np.int64).tofile(pre + "_z_beg_pos.bin")
    np.zeros(0, np.int32).tofile(pre + "_z_csr.bin")
    with pytest.raises(XbfsError) as e:
        Graph.load(pre + "_z_beg_pos.bin", pre + "_z_csr.bin")
    assert e.value.code == 5
    # neighbour id out of range
    g.write(pre)
    bad = g.adj.copy(); bad[3] = g.n
    bad.tofile(pre + "_csr.bin")
    with pytest.raises(XbfsError) as e:
        Graph.load(pre + "_beg_pos.bin", pre + "_csr.bin")
    assert e.value.code == 5


def test_loader_matches_reference_loader(tmp_path):
    """oracle/_ref/ref_loader is the reference's own graph.hpp + workload_gap1
    compiled from the reference tree; skipped where that tree is absent."""
    if not os.path.exists(O.REF_LOADER):
        pytest.skip("reference loader not built (reference tree absent)")
    import zlib  # noqa: F401
    g = Graph.rmat(9, 8, seed=77)
    pre = str(tmp_path / "g")
    g.write(pre)
    n, m, mx, vid, hb, hc = O.ref_loader(pre + "_beg_pos.bin", pre + "_csr.bin")
    assert (n, m) == (g.n, g.m)
    ds = g.degree_stats()
    assert mx == ds["max_degree"]
    want_vid = ds["max_vertex"] if ds["max_vertex"] != 0 else int(g.beg_pos[1] - g.beg_pos[0])
    assert vid == want_vid

    def fnv(b):
        h = 1469598103934665603
        for c in b:
            h ^= c
            h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        return h
    h = Graph.load(pre + "_beg_pos.bin", pre + "_csr.bin")
    assert fnv(h.beg_pos.tobytes()) == hb and fnv(h.adj.tobytes()) == hc


def test_edge_text(tmp_path):
    p = tmp_path / "e.txt"
    p.write_text("# comment\n0 1\n1 2\n2 2\n1 0\n5 3 7\n")
    g = Graph.from_edge_text(str(p), reverse=True)
    assert g.n == 6
    csr_invariants(g)
    assert list(g.adj[g.beg_pos[1]:g.beg_pos[2]]) == [0, 2]
    g1 = Graph.from_edge_text(str(p), reverse=False)
    assert list(g1.adj[g1.beg_pos[5]:g1.beg_pos[6]]) == [3] and g1.beg_pos[4] == g1.beg_pos[3]


def test_degree_stats():
    beg, adj = G.disjoint(G.star(300), G.path(3), G.csr_from_edges(2, []))
    g = Graph(beg, adj)
    d = g.degree_stats()
    assert d["max_degree"] == 299 and d["max_vertex"] == 0
    assert d["large"] == 1 and d["isolated"] == 2 and d["small"] == 304
    assert d["hub_num"] == 0


def test_header_symbols_exported():
    """Every function include/xbfs.h declares is exported by libxbfs.so and
    has a ctypes signature in the binding."""
    hdr = open(os.path.join(ROOT, "include", "xbfs.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(xbfs_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 15
    L = _capi.lib()
    for nm in sorted(names):
        assert hasattr(L, nm), f"{nm} not exported"
        assert nm in _capi.SIGNATURES, f"{nm} has no binding"


def test_concat_is_disjoint_union():
    """xbfs_graph_concat: graph i's vertex v is vert_off[i] + v; each slice
    of a multi-source BFS over the union from one seed per graph equals the
    graph's own single-source BFS."""
    gs = [Graph.grid(3, 5), Graph.rmat(7, 4, seed=11), Graph(np.zeros(2, np.int64),
                                                             np.zeros(0, np.int32)),
          Graph.rmat(6, 8, seed=12)]
    u, off = Graph.concat(gs)
    assert off.tolist() == np.cumsum([0] + [g.n for g in gs]).tolist()
    assert u.n == sum(g.n for g in gs) and u.m == sum(g.m for g in gs)
    csr_invariants(u)
    want_beg, want_adj = G.disjoint(*[(g.beg_pos, g.adj) for g in gs])
    assert (u.beg_pos == want_beg).all() and (u.adj == want_adj).all()
    srcs = [0, int(np.argmax(gs[1].degrees())), 0, 5]
    lv, _, _ = O.bfs_multi(u.beg_pos, u.adj, [off[i] + s for i, s in enumerate(srcs)])
    for i, g in enumerate(gs):
        want, _, _ = O.bfs_queue(g.beg_pos, g.adj, srcs[i])
        assert (lv[off[i]:off[i + 1]] == want).all()


def test_concat_rejects_bad_input():
    bad = Graph(np.array([0, 2], np.int64), np.zeros(1, np.int32))  # beg[n] != m
    with pytest.raises(XbfsError):
        Graph.concat([Graph.grid(2, 2), bad])
    with pytest.raises(XbfsError):
        Graph.concat([])

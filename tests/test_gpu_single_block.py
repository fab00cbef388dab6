"""GPU parity of the single-block levels of the persistent kernel: small
scan-free top-down levels (at most 2048 frontier edges) run on one workgroup,
level after level, while the other blocks park.
Levels, parents, reach, and the per-level records (mode, frontier vertices and
edges) must be the same with them (opts.single_block = 0, the default) and
without them (-1), and equal the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import graphs as G
from tests.test_gpu_parity import POLICIES, _check_one, say, seeded_sources
from xbfs_amd import Context, Graph, POLICY_FORCE_TD_SF, POLICY_HEURISTIC

pytestmark = pytest.mark.gpu


def both_ways(ctx, g, src, **kw):
    """Runs src with and without single-block levels; both must equal the
    oracle, and their level records must agree field by field."""
    lv1, recs1, st1 = _check_one(ctx, g, src, persistent=1, single_block=0, **kw)
    lv0, recs0, st0 = _check_one(ctx, g, src, persistent=1, single_block=-1, **kw)
    assert (lv1 == lv0).all()
    assert st1["n_levels"] == st0["n_levels"]
    for a, b in zip(recs1, recs0):
        for k in ("level", "mode", "tag", "n_f", "m_f", "n_u", "m_u"):
            assert a[k] == b[k], (src, kw, k, a, b)
    return recs1


@pytest.mark.parametrize("name", ["path", "grid", "tree", "star", "disjoint", "loops_dups"])
def test_closed_forms_with_and_without(name):
    from tests.test_gpu_parity import CLOSED
    g = Graph(*CLOSED[name]())
    ctx = Context(g)
    for src in (0, g.n // 2, g.n - 1):
        for pol in POLICIES:
            say("single block", name, src, pol)
            both_ways(ctx, g, src, **pol)
    ctx.close()


def test_rmat_and_grids_with_and_without():
    for g in (Graph.rmat(12, 16, seed=1012), Graph.rmat(14, 8, seed=1414), Graph.grid(64, 64),
              Graph.grid(300, 7), Graph.er(20000, 80000, seed=7)):
        ctx = Context(g)
        for s in seeded_sources(g, 3):
            for pol in POLICIES:
                say("single block", g.n, s, pol)
                both_ways(ctx, g, s, **pol)
        ctx.close()


def test_edge_cap():
    """Stars whose centre has 2047 leaves (n_f = 1, m_f = 2047: the largest
    single-block level, 2047 discoveries in one list) and 2048 and 5000
    leaves (at and above the 2048-edge cap).  From a leaf, the centre's level
    is the large one."""
    for k in (2047, 2048, 5000):
        g = Graph(*G.star(k + 1))
        ctx = Context(g)
        for src in (0, 1, k):
            for pol in (dict(policy=POLICY_HEURISTIC), dict(policy=POLICY_FORCE_TD_SF)):
                say("star", k, src, pol)
                both_ways(ctx, g, src, **pol)
        ctx.close()


def test_degree0_frontier_vertices_directed(monkeypatch):
    """A directed graph (top-down only: forced scan-free) whose frontiers hold
    vertices without out-edges.  Level 0 runs on one block and finds the
    out-degree-0 vertex 1, so level 1 returns to the grid; level 2 holds the
    out-degree-0 vertex 3, which the binned queue leaves out, so the
    single-block run that starts there covers one level and its frontier ring
    is cleared whole afterwards.  A long path follows, so every ring is used
    again.  A cycle of 10 vertices that no path from 0 reaches keeps the
    number of vertices with out-edges above the number reached (the end
    condition "every vertex with edges has a level" assumes a symmetric
    graph, where no reached vertex lacks edges).  The context skips the
    create-time component traversal, whose policy may run bottom-up, which
    reads rows as in-edges and so needs a symmetric graph."""
    monkeypatch.setenv("XBFS_COMPONENT", "off")
    n = 50
    pairs = [(0, 1), (0, 2), (2, 3), (2, 4)] + [(i, i + 1) for i in range(4, 39)]
    pairs += [(i, i + 1) for i in range(40, 49)] + [(49, 40)]
    beg, adj = G.csr_from_edges(n, pairs, symmetric=False)
    g = Graph(beg, adj)
    ctx = Context(g)
    want, _, _ = O.bfs_queue(g.beg_pos, g.adj, 0)
    for sb in (0, -1):
        for eng in (1, 0):
            say("directed", sb, eng)
            lv, par, recs, st = ctx.bfs(0, policy=POLICY_FORCE_TD_SF, persistent=eng,
                                        single_block=sb, want_parent=True, rec_cap=64)
            assert (lv == want).all(), (lv, want)
            for v in np.flatnonzero(lv > 0):
                p = par[v]
                assert lv[p] == lv[v] - 1
                assert v in adj[beg[p]:beg[p + 1]]
            assert st["v_cc"] == (want >= 0).sum()
    ctx.close()


@pytest.mark.parametrize("grid_dim", [1, 2, 33])
def test_grid_counts_with_and_without(grid_dim):
    """Grids of one block (the single-block levels need no parking, and
    leaving them is decided by the mode alone), two, and 33 (one block beyond
    the small-frontier count), on a lattice and on an RMAT graph whose
    heuristic schedule turns bottom-up after single-block levels."""
    for g in (Graph.grid(48, 80), Graph.rmat(11, 16, seed=31)):
        ctx = Context(g)
        for s in seeded_sources(g, 3):
            for pol in POLICIES[:4]:
                say("single block grid", grid_dim, g.n, s, pol)
                both_ways(ctx, g, s, grid_dim=grid_dim, **pol)
        ctx.close()

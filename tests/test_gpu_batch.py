"""GPU parity of the batched modes, through the C ABI:
  - multi-source BFS (xbfs_bfs_multi) = the oracle's multi-source BFS;
  - multigraph batch (xbfs_graph_concat + xbfs_bfs_multigraph): every graph's
    slice = the oracle's single-source BFS on that graph alone;
  - 64-source batch (xbfs_bfs_batch64): row s = the oracle's single-source
    BFS from srcs[s], bit for bit;
  - the autotuner returns a valid setting that still gives exact levels."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests import graphs as G
from xbfs_amd import (MODE_BU, MODE_TD_SF, POLICY_FORCE_BU, POLICY_FORCE_TD_SF,
                      POLICY_FORCE_TD_SS, POLICY_HEURISTIC, POLICY_SCHEDULE, Context, Graph,
                      default_opts)

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def seeded_sources(g, k, seed=42):
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(g.degrees() > 0)
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


def G_(beg_adj):
    beg, adj = beg_adj
    return Graph(beg, adj)


CLOSED = {
    "path": lambda: G_(G.path(300)),
    "star": lambda: G_(G.star(257)),
    "grid": lambda: G_(G.grid(33, 47)),
    "tree": lambda: G_(G.binary_tree(2047)),
    "disjoint": lambda: G_(G.disjoint(G.path(40), G.star(100), G.csr_from_edges(7, []),
                                      G.grid(9, 9))),
}
POLICIES = [dict(policy=POLICY_HEURISTIC), dict(policy=POLICY_FORCE_TD_SF),
            dict(policy=POLICY_FORCE_TD_SS), dict(policy=POLICY_FORCE_BU),
            dict(policy=POLICY_FORCE_TD_SF, bin_thread_max=2, bin_wave_max=8, bin_block_max=32)]


# ---- multi-source ------------------------------------------------------------------
def check_multi(ctx, g, srcs, **kw):
    lv, par, recs, st = ctx.bfs_multi(srcs, want_parent=True, rec_cap=64, **kw)
    want, _, nl = O.bfs_multi(g.beg_pos, g.adj, srcs)
    bad = np.flatnonzero(lv != want)
    assert bad.size == 0, f"srcs={list(srcs)[:8]} {kw}: {bad.size} mismatches, first {bad[:5]} " \
                          f"got {lv[bad[:5]]} want {want[bad[:5]]}"
    assert O.validate_multi(g.beg_pos, g.adj, srcs, lv, par) == 0
    reached = want >= 0
    assert st["v_cc"] == reached.sum()
    assert st["m_cc"] == g.degrees()[reached].sum()
    assert st["max_level"] == want.max()
    return st


@pytest.mark.parametrize("name", sorted(CLOSED))
@pytest.mark.parametrize("pi", range(len(POLICIES)))
def test_multi_closed_forms(name, pi):
    g = CLOSED[name]()
    ctx = Context(g)
    rng = np.random.default_rng(pi)
    srcs = rng.choice(g.n, size=5, replace=False).astype(np.int32)
    srcs = np.concatenate([srcs, srcs[:2]])  # duplicates
    for eng in (1, 0):
        say(name, POLICIES[pi], "persistent", eng)
        check_multi(ctx, g, srcs, persistent=eng, **POLICIES[pi])
    ctx.close()


def test_multi_isolated_and_single_seed():
    g = CLOSED["disjoint"]()
    ctx = Context(g)
    iso = [v for v in range(g.n) if g.degrees()[v] == 0]
    for srcs in ([iso[0]], [iso[0], iso[1]], [iso[0], 3], [3], [3, 3, 3]):
        for eng in (1, 0):
            say("seeds", srcs, "persistent", eng)
            check_multi(ctx, g, np.array(srcs, np.int32), persistent=eng)
    ctx.close()


@pytest.mark.parametrize("scale,ef,k", [(12, 8, 2), (14, 16, 16), (14, 4, 64)])
def test_multi_rmat(scale, ef, k):
    g = Graph.rmat(scale, ef, seed=500 + scale + ef)
    ctx = Context(g)
    srcs = seeded_sources(g, k, seed=k)
    for kw in POLICIES[:4]:
        for eng in (1, 0):
            say(f"rmat{scale} ef{ef} k{k}", kw, "persistent", eng)
            check_multi(ctx, g, srcs, persistent=eng, **kw)
    ctx.close()


def test_multi_rejects_bad_seeds():
    g = Graph.rmat(8, 4, seed=3)
    ctx = Context(g)
    with pytest.raises(Exception):
        ctx.bfs_multi([0, g.n])
    with pytest.raises(Exception):
        ctx.bfs_multi([])
    ctx.close()


# ---- multigraph batch -----------------------------------------------------------------
def uneven_graphs(count, scales, efs, seed0):
    gs, srcs = [], []
    for i in range(count):
        rng = np.random.default_rng(seed0 + i)
        scale = int(rng.integers(scales[0], scales[1] + 1))
        ef = int(rng.integers(efs[0], efs[1] + 1))
        g = Graph.rmat(scale, ef, seed=seed0 + i)
        gs.append(g)
        srcs.append(int(seeded_sources(g, 1, seed=seed0 + i)[0]))
    return gs, srcs


def check_multigraph(gs, srcs, **kw):
    u, off = Graph.concat(gs)
    ctx = Context(u)
    lv, st = ctx.bfs_multigraph(off, srcs, **kw)
    m_cc = 0
    for i, g in enumerate(gs):
        want, _, _ = O.bfs_queue(g.beg_pos, g.adj, srcs[i])
        got = lv[off[i]:off[i + 1]]
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, f"graph {i} (n={g.n}) src {srcs[i]} {kw}: {bad.size} mismatches"
        m_cc += int(g.degrees()[want >= 0].sum())
    assert st["m_cc"] == m_cc
    ctx.close()
    return st


@pytest.mark.parametrize("eng", [1, 0])
def test_multigraph_uneven(eng):
    gs, srcs = uneven_graphs(32, (6, 12), (2, 16), 20000)
    gs.append(Graph.grid(20, 30))  # a long-diameter member
    srcs.append(0)
    for kw in POLICIES[:4]:
        say("multigraph 33", kw, "persistent", eng)
        check_multigraph(gs, srcs, persistent=eng, **kw)


def test_multigraph_256_uneven_rmat():
    """256 RMAT graphs, scale U{10..16}, edge factor U{4..32}, seeds 10000+i,
    one seeded source each, in one persistent launch."""
    gs, srcs = uneven_graphs(256, (10, 16), (4, 32), 10000)
    say("multigraph 256, total n", sum(g.n for g in gs))
    st = check_multigraph(gs, srcs)
    assert st["n_levels"] >= 2


def test_multigraph_rejects_bad_offsets():
    gs, srcs = uneven_graphs(3, (6, 7), (4, 4), 30000)
    u, off = Graph.concat(gs)
    ctx = Context(u)
    with pytest.raises(Exception):
        ctx.bfs_multigraph(off, [0, 0, gs[2].n])  # local source outside graph 2
    bad = off.copy()
    bad[-1] -= 1
    with pytest.raises(Exception):
        ctx.bfs_multigraph(bad, srcs)
    ctx.close()


# ---- 64-source batch ---------------------------------------------------------------
def check_batch(ctx, g, srcs, **kw):
    lv, st = ctx.bfs_batch64(srcs, **kw)
    assert lv.shape == (len(srcs), g.n)
    v_cc = m_cc = 0
    deg = g.degrees()
    for s, src in enumerate(srcs):
        want, _, _ = O.bfs_queue(g.beg_pos, g.adj, int(src))
        bad = np.flatnonzero(lv[s] != want)
        assert bad.size == 0, f"row {s} src {src} {kw}: {bad.size} mismatches, first " \
                              f"{bad[:5]} got {lv[s][bad[:5]]} want {want[bad[:5]]}"
        v_cc += int((want >= 0).sum())
        m_cc += int(deg[want >= 0].sum())
    assert st["v_cc"] == v_cc and st["m_cc"] == m_cc
    return st


BATCH_MODES = [dict(policy=POLICY_HEURISTIC), dict(policy=POLICY_FORCE_TD_SF),
               dict(policy=POLICY_FORCE_BU), dict(policy=POLICY_HEURISTIC, alpha=0.001),
               dict(policy=POLICY_SCHEDULE, schedule=[MODE_TD_SF, MODE_BU] * 8)]


@pytest.mark.parametrize("name", sorted(CLOSED))
def test_batch64_closed_forms(name):
    """path and grid outlast the 64-slot history: their long-running sources
    finish on the single-source engine."""
    g = CLOSED[name]()
    ctx = Context(g)
    rng = np.random.default_rng(7)
    srcs = rng.choice(g.n, size=min(64, g.n), replace=False).astype(np.int32)
    for kw in BATCH_MODES[:3]:
        say("batch64", name, kw)
        check_batch(ctx, g, srcs, **kw)
    ctx.close()


@pytest.mark.parametrize("scale,ef", [(12, 4), (14, 16)])
def test_batch64_rmat(scale, ef):
    g = Graph.rmat(scale, ef, seed=700 + scale)
    ctx = Context(g)
    srcs = seeded_sources(g, 64, seed=scale)
    for kw in BATCH_MODES:
        say(f"batch64 rmat{scale} ef{ef}", kw)
        st = check_batch(ctx, g, srcs, **kw)
        assert st["n_levels"] >= 2
    ctx.close()


def test_batch64_source_counts_duplicates_isolated():
    g = G_(G.disjoint(G.path(40), G.star(100), G.csr_from_edges(7, []), G.grid(9, 9)))
    ctx = Context(g)
    iso = int(np.flatnonzero(g.degrees() == 0)[0])
    for srcs in ([5], [iso], [5, 5, iso, 41, 41], list(range(0, 189, 3))[:63],
                 list(range(64))):
        say("batch64 srcs", srcs[:6], len(srcs))
        check_batch(ctx, g, np.array(srcs, np.int32))
    with pytest.raises(Exception):
        ctx.bfs_batch64(np.arange(65, dtype=np.int32))
    with pytest.raises(Exception):
        ctx.bfs_batch64(np.array([g.n], np.int32))
    ctx.close()


def test_batch64_rmat22_full_size():
    """RMAT scale 22, edge factor 16, 64 seeded sources: every row equals the
    oracle's single-source BFS from its source, bit for bit."""
    g = Graph.rmat(22, 16, seed=1022)
    ctx = Context(g)
    srcs = seeded_sources(g, 64)
    say("batch64 rmat22 n", g.n, "m", g.m)
    lv, st = ctx.bfs_batch64(srcs)
    say("batch64 rmat22: oracle rows")
    want = O.bfs_queue_rows(g.beg_pos, g.adj, srcs)
    deg = g.degrees()
    m_cc = 0
    for s in range(len(srcs)):
        bad = np.flatnonzero(lv[s] != want[s])
        assert bad.size == 0, f"row {s} src {srcs[s]}: {bad.size} levels differ from the oracle"
        m_cc += int(deg[want[s] >= 0].sum())
    assert st["m_cc"] == m_cc
    ctx.close()


# ---- autotuner ---------------------------------------------------------------------
def test_autotune_single_and_multi():
    g = Graph.rmat(13, 16, seed=77)
    ctx = Context(g)
    srcs = seeded_sources(g, 3)
    o, trials = ctx.autotune(srcs, reps=1)
    assert len(trials) > 10
    assert 0 < o.bin_thread_max <= o.bin_wave_max <= o.bin_block_max
    assert o.grid_dim >= 1
    assert min(t["t_s"] for t in trials) > 0
    for s in srcs:
        lv, _, _, _ = ctx.bfs(int(s), opts=o)
        assert (lv == O.bfs_queue(g.beg_pos, g.adj, int(s))[0]).all()
    ctx.close()
    gs, ls = uneven_graphs(8, (6, 10), (4, 16), 40000)
    u, off = Graph.concat(gs)
    ctx = Context(u)
    seeds = (off[:-1] + np.array(ls)).astype(np.int32)
    o, trials = ctx.autotune(seeds, multi=True, reps=1)
    assert len(trials) > 10
    lv, _ = ctx.bfs_multigraph(off, ls, opts=o)
    for i, gi in enumerate(gs):
        assert (lv[off[i]:off[i + 1]] == O.bfs_queue(gi.beg_pos, gi.adj, ls[i])[0]).all()
    ctx.close()

"""Pins the CPU oracle: closed forms, brute force, planted mistakes, golden
fixtures, hand-derived heuristic tag sequences, scan and MLP gradients."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import graphs as G

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rmat10_golden.npz")


def both(beg, adj, src):
    a, par, k1 = O.bfs_queue(beg, adj, src, parents=True)
    b, k2 = O.bfs_scan(beg, adj, src)
    assert (a == b).all(), "queue BFS and reference scan BFS disagree"
    assert k1 == k2
    assert O.validate(beg, adj, src, a, par) == 0
    return a


@pytest.mark.parametrize("n,src", [(1, 0), (2, 1), (17, 0), (17, 8), (64, 63)])
def test_path(n, src):
    if n == 1:
        beg, adj = np.zeros(2, np.int64), np.zeros(0, np.int32)
    else:
        beg, adj = G.path(n)
    lv = both(beg, adj, src)
    assert (lv == np.abs(np.arange(n) - src)).all()


@pytest.mark.parametrize("r,c,src", [(5, 7, 0), (8, 8, 27), (1, 9, 4)])
def test_grid_manhattan(r, c, src):
    beg, adj = G.grid(r, c)
    lv = both(beg, adj, src)
    i, j = np.divmod(np.arange(r * c), c)
    si, sj = divmod(src, c)
    assert (lv == np.abs(i - si) + np.abs(j - sj)).all()


def test_star_complete_tree():
    beg, adj = G.star(33)
    assert (both(beg, adj, 0) == np.r_[0, np.ones(32)]).all()
    assert (both(beg, adj, 5) == np.r_[1, np.full(32, 2)].astype(np.int32) - np.eye(33, dtype=np.int32)[5] * 2).all()
    beg, adj = G.complete(12)
    lv = both(beg, adj, 3)
    assert (lv == (np.arange(12) != 3)).all()
    beg, adj = G.binary_tree(127)
    lv = both(beg, adj, 0)
    assert (lv == np.floor(np.log2(np.arange(127) + 1)).astype(np.int32)).all()


def test_disconnected_and_isolated_source():
    beg, adj = G.disjoint(G.path(5), G.star(4), G.csr_from_edges(3, []))
    lv = both(beg, adj, 1)
    assert (lv[:5] == [1, 0, 1, 2, 3]).all() and (lv[5:] == -1).all()
    lv = both(beg, adj, 10)  # degree-0 vertex
    assert lv[10] == 0 and (np.delete(lv, 10) == -1).all()


@pytest.mark.parametrize("seed", range(6))
def test_brute_force(seed):
    n = 48
    beg, adj = G.random_graph(n, 0.06, seed)
    for src in (0, seed + 3, n - 1):
        lv = both(beg, adj, src)
        assert (lv == G.floyd_warshall_levels(beg, adj, src)).all()


def test_validate_catches_planted_mistakes():
    beg, adj = G.grid(6, 6)
    lv, par, _ = O.bfs_queue(beg, adj, 0, parents=True)
    assert O.validate(beg, adj, 0, lv, par) == 0
    bad = lv.copy(); bad[14] += 1
    assert O.validate(beg, adj, 0, bad, par) > 0
    bad = lv.copy(); bad[35] = -1
    assert O.validate(beg, adj, 0, bad, None) > 0
    bad = lv.copy(); bad[0] = 1
    assert O.validate(beg, adj, 0, bad, None) > 0
    bad = lv + 1; bad[0] = 0  # shifted levels: no predecessor one level up
    assert O.validate(beg, adj, 0, bad, None) > 0
    p = par.copy(); p[20] = 0  # parent not adjacent
    assert O.validate(beg, adj, 0, lv, p) > 0
    p = par.copy(); p[20] = 20
    assert O.validate(beg, adj, 0, lv, p) > 0
    # a level that is off by two on an edge
    beg2, adj2 = G.path(4)
    assert O.validate(beg2, adj2, 0, np.array([0, 1, 3, 4], np.int32), None) > 0


def test_golden_fixture():
    d = np.load(GOLDEN)
    beg, adj = d["beg_pos"], d["adj"]
    for s, want in zip(d["sources"], d["levels"]):
        lv = both(beg, adj, int(s))
        assert (lv == want).all()
    for key, alpha in (("tags_a0p1", 0.1), ("tags_a0p01", 0.01)):
        for s, want in zip(d["sources"], d[key]):
            t, _, _ = O.reference_tags(beg, adj, int(s), alpha)
            assert (t == want[: len(t)]).all() and (want[len(t):] == -9).all()


def test_generator_matches_golden():
    from xbfs_amd import Graph
    d = np.load(GOLDEN)
    g = Graph.rmat(10, 16, seed=1010)
    assert (g.beg_pos == d["beg_pos"]).all() and (g.adj == d["adj"]).all()


def test_level_counts_closed_form():
    beg, adj = G.star(11)
    lv = both(beg, adj, 0)
    cnt, nv, nl = O.level_counts(beg, lv)
    assert nl == 2 and list(cnt) == [10, 10] and list(nv) == [1, 10]
    beg, adj = G.path(6)
    lv = both(beg, adj, 0)
    cnt, nv, nl = O.level_counts(beg, lv)
    assert list(cnt) == [1, 2, 2, 2, 2, 1]


def test_heuristic_hand_derived():
    E = 1_000_000
    # plain: all below alpha -> scan-free everywhere
    t, s, tda = O.heuristic_tags([5, 50, 500], E, 0.1)
    assert list(t) == [0, 0, 0] and list(s) == [0, 0, 0] and tda == 10000
    # BU switch then back: level 2 crosses alpha (hub_before=2), level 4 is the
    # first TD after BU (hub_after=4, single-scan), level 5 is "last phase"
    t, s, tda = O.heuristic_tags([5, 500, 200_000, 150_000, 900, 40], E, 0.1)
    assert list(t) == [0, 0, -1, -1, 1, 0], t
    assert list(s) == [0, 0, -1, -1, 1, 2], s
    # level-0 hub source: count[0] >= INF -> TD_adaptive=1, level 1 single-scan
    t, s, tda = O.heuristic_tags([20_000, 30_000], E, 0.1)
    assert tda == 1 and list(t) == [0, 1]
    # jump rule: count[L-1]>=INF, count[L]>=10*INF and ratio>=8 -> TD_adaptive=L
    t, s, tda = O.heuristic_tags([10, 12_000, 120_000, 90_000], E, 0.5)
    assert tda == 2 and list(t) == [0, 0, 1, 1]
    # just below the jump rule (count[2] = 96,000 < 10*INF): nothing fires
    t, s, tda = O.heuristic_tags([10, 12_000, 96_000, 90_000], E, 0.5)
    assert tda == 10000 and list(t) == [0, 0, 0, 0]
    # 8x rule not met at L=3 (ratio 5) but the >=100*INF rule fires there
    t, s, tda = O.heuristic_tags([20, 9_000, 200_000, 1_000_000], 10 * E, 0.5)
    assert tda == 3 and list(t) == [0, 0, 0, 1]
    # small-previous rule: count[L-1]<INF, count>=10*INF and ratio>=1000
    t, s, tda = O.heuristic_tags([10, 100_000], E, 0.5)
    assert tda == 1 and list(t) == [0, 1]
    # ratio exactly alpha goes bottom-up (ratio >= alpha)
    t, s, _ = O.heuristic_tags([1, 100_000], E, 0.1)
    assert list(t) == [0, -1]
    # hub source with count[0] >= INF: bottom-up at 0 sets TD_adaptive=1, so
    # level 1 (first TD after BU) is single-scan
    t, s, tda = O.heuristic_tags([500_000, 50], E, 0.1)
    assert tda == 1 and list(t) == [-1, 1] and list(s) == [-1, 1]
    # bottom-up at level 0 below INF keeps hub_before = 0 (tag[-1] is inert):
    # level 1 takes the scan-free branch right after a bottom-up level
    t, s, tda = O.heuristic_tags([500, 50], 1000, 0.1)
    assert tda == 10000 and list(t) == [-1, 0] and list(s) == [-1, 0]


def test_exclusive_scan():
    rng = np.random.default_rng(3)
    for n in (0, 1, 7, 1000, 65537):
        x = rng.integers(-5, 1000, n, dtype=np.int32)
        want = np.concatenate([[0], np.cumsum(x, dtype=np.int64)[:-1]]) if n else np.zeros(0)
        assert (O.exclusive_scan(x) == want).all()


def test_omp_td_matches_queue():
    from xbfs_amd import Graph
    g = Graph.rmat(12, 16, seed=1012)
    for s in (0, 17, 1000):
        if g.beg_pos[s + 1] == g.beg_pos[s]:
            continue
        a, _, _ = O.bfs_queue(g.beg_pos, g.adj, s)
        b, t = O.omp_td_bfs(g.beg_pos, g.adj, s, 4)
        assert (a == b).all() and t > 0


def test_queue_rows_equal_single_source_rows():
    """oracle_bfs_queue_rows (rows in parallel threads) = oracle_bfs_queue per
    source, including a repeated source and a degree-0 source."""
    from xbfs_amd import Graph
    g = Graph.rmat(12, 8, seed=1212)
    iso = int(np.flatnonzero(g.degrees() == 0)[0])
    srcs = np.array([0, 5, 5, 100, iso, 4095], np.int32)
    rows = O.bfs_queue_rows(g.beg_pos, g.adj, srcs)
    assert rows.shape == (srcs.size, g.n)
    for i, s in enumerate(srcs):
        want, _, _ = O.bfs_queue(g.beg_pos, g.adj, int(s))
        assert (rows[i] == want).all()
    assert (rows[4] >= 0).sum() == 1


def test_mlp_grad_finite_difference():
    rng = np.random.default_rng(0)
    P = (rng.standard_normal(O.MLP_NPARAM) * 0.3).astype(np.float32)
    x = rng.standard_normal((9, 8)).astype(np.float32)
    t = rng.standard_normal((9, 3)).astype(np.float32)
    mask = (rng.random((9, 3)) < 0.7).astype(np.float32)
    loss, g = O.mlp_loss_grad(P, x, t, mask)
    p64 = P.astype(np.float64)
    assert abs(loss - O.mlp_loss(p64, x, t, mask)) < 1e-12
    h = 1e-6
    idx = np.r_[np.arange(0, 256, 7), np.arange(256, 1443, 11), [1440, 1441, 1442]]
    for i in idx:
        e = np.zeros_like(p64); e[i] = h
        fd = (O.mlp_loss(p64 + e, x, t, mask) - O.mlp_loss(p64 - e, x, t, mask)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # forward of a hand-built net: identity-ish weights give a known output
    P2 = np.zeros(O.MLP_NPARAM, np.float32)
    P2[0] = 1.0                       # W1[0][0]
    P2[256 + 32 + 0] = 2.0            # W2[0][0]
    P2[256 + 32 + 1024 + 32 + 0] = 3.0  # W3[0][0]
    P2[-1] = 0.5                      # b3[2]
    y = O.mlp_forward(P2, np.array([[1.5] + [0] * 7, [-1.0] + [0] * 7], np.float32))
    assert np.allclose(y, [[9.0, 0, 0.5], [0, 0, 0.5]])


# ---- multi-source oracle (batched modes) ---------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_multi_source_is_min_of_single_sources(seed):
    """Distance to the nearest seed = min over seeds of the single-source
    distance (brute force over the seeds; -1 = unreached)."""
    rng = np.random.default_rng(seed)
    beg, adj = G.disjoint(G.random_graph(40, 0.06, seed), G.random_graph(25, 0.1, seed + 50),
                          G.path(7))
    n = beg.shape[0] - 1
    k = int(rng.integers(1, 9))
    srcs = rng.integers(0, n, size=k).astype(np.int32)
    srcs = np.concatenate([srcs, srcs[:1]])  # a duplicate
    lv, par, _ = O.bfs_multi(beg, adj, srcs, parents=True)
    singles = np.stack([O.bfs_queue(beg, adj, int(s))[0] for s in srcs])
    big = np.where(singles < 0, 10 ** 9, singles).min(axis=0)
    want = np.where(big >= 10 ** 9, -1, big)
    assert (lv == want).all()
    assert O.validate_multi(beg, adj, srcs, lv, par) == 0


def test_multi_source_path_closed_form():
    beg, adj = G.path(50)
    for a, b in ((0, 49), (10, 11), (25, 25), (3, 40)):
        lv, _, k = O.bfs_multi(beg, adj, [a, b])
        i = np.arange(50)
        assert (lv == np.minimum(np.abs(i - a), np.abs(i - b))).all()
        assert k == lv.max() + 1


def test_multi_source_isolated_seed():
    beg, adj = G.disjoint(G.path(5), (np.zeros(2, np.int64), np.zeros(0, np.int32)))
    lv, par, _ = O.bfs_multi(beg, adj, [5, 2], parents=True)
    assert lv.tolist() == [2, 1, 0, 1, 2, 0]
    assert par[5] == 5 and par[2] == 2
    assert O.validate_multi(beg, adj, [5, 2], lv, par) == 0


def test_validate_multi_catches_planted_mistakes():
    beg, adj = G.grid(6, 7)
    srcs = np.array([0, 41], np.int32)
    lv, par, _ = O.bfs_multi(beg, adj, srcs, parents=True)
    assert O.validate_multi(beg, adj, srcs, lv, par) == 0
    bad = lv.copy(); bad[20] += 1
    assert O.validate_multi(beg, adj, srcs, bad) > 0          # wrong level
    bad = lv.copy(); bad[41] = 1
    assert O.validate_multi(beg, adj, srcs, bad) > 0          # seed not at 0
    bad = lv.copy(); bad[20] = 0
    assert O.validate_multi(beg, adj, srcs, bad) > 0          # non-seed at 0
    bad = lv.copy(); bad[7] = -1
    assert O.validate_multi(beg, adj, srcs, bad) > 0          # missing vertex
    assert O.validate_multi(beg, adj, srcs[:1], lv) > 0       # levels of another seed set
    p = par.copy(); p[20] = 20
    assert O.validate_multi(beg, adj, srcs, lv, p) > 0        # self parent
    p = par.copy(); p[20] = 0
    assert O.validate_multi(beg, adj, srcs, lv, p) > 0        # non-adjacent parent

"""GPU parity: levels from libxbfs (HIP path, through the C ABI) must equal the
oracle's levels bit for bit; parents must be tree edges one level up; the
heuristic policy must emit the reference's tag sequence."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests import graphs as G
from xbfs_amd import (Context, Graph, XbfsError, MODE_BU, MODE_TD_SF, MODE_TD_SS, POLICY_FORCE_BU,
                      POLICY_FORCE_TD_SF, POLICY_FORCE_TD_SS, POLICY_HEURISTIC,
                      POLICY_SCHEDULE, default_opts)

pytestmark = pytest.mark.gpu

SF, SS, BU = MODE_TD_SF, MODE_TD_SS, MODE_BU


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def seeded_sources(g, k, seed=42):
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(g.degrees() > 0)
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


def run_check(ctx, g, src, engines=(1, 0), **kw):
    """Checks both drivers (persistent kernel and host-driven loop); returns
    the persistent run's results."""
    out = None
    for eng in engines:
        r = _check_one(ctx, g, src, persistent=eng, **kw)
        if out is None:
            out = r
    return out


def _check_one(ctx, g, src, **kw):
    lv, par, recs, st = ctx.bfs(int(src), want_parent=True, rec_cap=64, **kw)
    want, _, nl = O.bfs_queue(g.beg_pos, g.adj, int(src))
    bad = np.flatnonzero(lv != want)
    assert bad.size == 0, f"src={src} {kw}: {bad.size} mismatches, first {bad[:5]} " \
                          f"got {lv[bad[:5]]} want {want[bad[:5]]}"
    assert O.validate(g.beg_pos, g.adj, int(src), lv, par) == 0
    reached = want >= 0
    assert st["v_cc"] == reached.sum()
    assert st["m_cc"] == g.degrees()[reached].sum()
    assert st["max_level"] == want.max()
    return lv, recs, st


def assert_tags_prefix(got, want, st):
    """The reference's tag sequence (restated for levels 0..maxl-1) and ours
    must agree on every level both execute.  Either may stop early: the
    traversal ends as soon as every non-isolated vertex has a level, which the
    bottom-up +2 rule can reach one level before maxl (the reference's own
    loop ends on the same condition, bfs_adaptiveFQ_wb_async.cuh:1039,1247)."""
    k = min(len(got), len(want))
    assert (got[:k] == want[:k]).all(), (got, want)
    assert len(got) >= len(want) - 1, (got, want)
    assert st["n_levels"] == len(got)


CLOSED = {
    "path": lambda: G.path(300),
    "star": lambda: G.star(257),
    "grid": lambda: G.grid(33, 47),
    "tree": lambda: G.binary_tree(2047),
    "complete": lambda: G.complete(70),
    "disjoint": lambda: G.disjoint(G.path(40), G.star(100), G.csr_from_edges(7, []),
                                   G.grid(9, 9)),
    "loops_dups": G.loops_and_duplicates,
}
POLICIES = [dict(policy=POLICY_HEURISTIC), dict(policy=POLICY_FORCE_TD_SF),
            dict(policy=POLICY_FORCE_TD_SS), dict(policy=POLICY_FORCE_BU),
            dict(policy=POLICY_HEURISTIC, alpha=0.01),
            dict(policy=POLICY_FORCE_TD_SF, bin_thread_max=2, bin_wave_max=8, bin_block_max=32)]


@pytest.mark.parametrize("name", sorted(CLOSED))
def test_closed_forms_all_policies(name):
    beg, adj = CLOSED[name]()
    g = Graph(beg, adj)
    ctx = Context(g)
    for src in (0, g.n // 2, g.n - 1):
        for pol in POLICIES:
            say(name, src, pol)
            run_check(ctx, g, src, **pol)
    ctx.close()


def test_csr_row_order(monkeypatch):
    """XBFS_ROW_ORDER=csr keeps the uploaded adjacency order (no degree sort,
    bottom-up scans from each row's first entry): same levels, valid parents."""
    monkeypatch.setenv("XBFS_ROW_ORDER", "csr")
    for name in ("loops_dups", "star", "disjoint"):
        g = Graph(*CLOSED[name]())
        ctx = Context(g)
        for src in (0, g.n // 2, g.n - 1):
            for pol in POLICIES[:4]:
                say("csr order", name, src, pol)
                run_check(ctx, g, src, **pol)
        ctx.close()
    g = Graph.rmat(14, 16, seed=1414)
    ctx = Context(g)
    for s in seeded_sources(g, 4):
        for pol in POLICIES[:4]:
            say("csr order rmat14", s, pol)
            run_check(ctx, g, s, **pol)
    ctx.close()


def test_level_cap_then_full_run():
    """opts.max_levels, forced from the host, ends a 299-level traversal with
    XBFS_ERR_LIMIT under both drivers; the same context then runs complete
    traversals correctly (the failure leaves nothing behind)."""
    g = Graph(*G.path(300))
    ctx = Context(g)
    for eng in (1, 0):
        for cap in (1, 10, 298):
            say("level cap", eng, cap)
            with pytest.raises(XbfsError) as ei:
                ctx.bfs(0, persistent=eng, max_levels=cap)
            assert ei.value.code == 7  # XBFS_ERR_LIMIT
            run_check(ctx, g, 0, engines=(eng,))
    run_check(ctx, g, 0, max_levels=299)  # exactly enough levels
    ctx.close()


def test_barrier_wait_limit_then_full_run():
    """A 1 us bound on each grid-barrier wait of the persistent kernel (the
    kernel's own bounded wait, tightened from the host) ends the traversal
    with XBFS_ERR_DEVICE: blocks that time out set the error word, every other
    block leaves at its next barrier, so the kernel still exits normally.  The
    same context then runs complete traversals correctly."""
    g = Graph(*G.grid(64, 64))
    ctx = Context(g)
    for pol in (dict(policy=POLICY_HEURISTIC), dict(policy=POLICY_FORCE_BU)):
        say("wait limit", pol)
        with pytest.raises(XbfsError) as ei:
            ctx.bfs(0, persistent=1, wait_limit_us=1, **pol)
        assert ei.value.code == 8  # XBFS_ERR_DEVICE
        run_check(ctx, g, 0, engines=(1,), **pol)
    ctx.close()


def test_launch_kinds(monkeypatch):
    """The persistent kernel is launched plainly with a residency check at its
    first barrier, cooperatively under XBFS_LAUNCH=coop, and with a check
    bound of one clock tick (XBFS_RESIDENCY_US=0) the first traversal's
    blocks cannot all arrive in time: it is relaunched cooperatively, and the
    context stays cooperative.  Levels and parents equal the oracle's in every
    case."""
    g = Graph.rmat(12, 16, seed=1012)
    srcs = seeded_sources(g, 3)
    for env, want_fb in ((None, 0), ("coop", 0), ("tick", 1)):
        monkeypatch.delenv("XBFS_LAUNCH", raising=False)
        monkeypatch.delenv("XBFS_RESIDENCY_US", raising=False)
        if env == "coop":
            monkeypatch.setenv("XBFS_LAUNCH", "coop")
        elif env == "tick":
            monkeypatch.setenv("XBFS_RESIDENCY_US", "0")
        ctx = Context(g)
        for s in srcs:
            for pol in POLICIES[:4]:
                say("launch", env, s, pol)
                _check_one(ctx, g, s, persistent=1, **pol)
        assert ctx.launch_fallbacks() == want_fb, (env, ctx.launch_fallbacks())
        ctx.close()


def test_nontemporal_bottom_up_loads(monkeypatch):
    """Bottom-up phases read the hub records and row starts with
    non-temporal loads on graphs whose hub table exceeds 64 MB and with plain
    loads below; XBFS_NT_BU=1 / 0 forces either on a context.  A graph of
    either size gives the oracle's levels and parents both ways."""
    g = Graph.rmat(16, 16, seed=1016)
    srcs = seeded_sources(g, 3)
    for env in ("1", "0"):
        monkeypatch.setenv("XBFS_NT_BU", env)
        ctx = Context(g)
        for s in srcs:
            for pol in (dict(policy=POLICY_HEURISTIC), dict(policy=POLICY_FORCE_BU)):
                say("nt_bu", env, s, pol)
                _check_one(ctx, g, s, persistent=1, **pol)
        ctx.close()


def test_hub_component_termination(monkeypatch):
    """A traversal from the largest-degree vertex's component ends as soon as
    every member has its level, one level (or, after +2 discoveries, two)
    sooner than waiting for an empty frontier, and its bottom-up candidates
    leave out the other components.  Sources outside the component run as
    before.  Levels, parents and reach equal the oracle's either way, and
    equal a context built with XBFS_COMPONENT=off."""
    beg, adj = G.disjoint(G.star(100), G.path(40), G.grid(9, 9), G.csr_from_edges(2, [(0, 1)]),
                          G.csr_from_edges(5, []))
    g = Graph(beg, adj)
    ctx = Context(g)
    monkeypatch.setenv("XBFS_COMPONENT", "off")
    plain = Context(g)
    for src in (0, 1, 50, 99, 100, 139, 140, 200, 221, 222):
        for pol in POLICIES:
            say("hub component", src, pol)
            for eng in (1, 0):
                lv, _, st = _check_one(ctx, g, src, persistent=eng, **pol)
                lv0, _, st0 = _check_one(plain, g, src, persistent=eng, **pol)
                assert (lv == lv0).all()
                if src < 100:
                    assert st["n_levels"] < st0["n_levels"], (st, st0)
                else:
                    assert st["n_levels"] == st0["n_levels"], (st, st0)
    # multi-source traversals keep the empty-frontier end (seeds may lie in
    # several components)
    for srcs in ([0], [5, 120], [120, 5, 222]):
        lv, _, _, st = ctx.bfs_multi(np.array(srcs, np.int32))
        want = O.bfs_multi(g.beg_pos, g.adj, np.array(srcs, np.int32))[0]
        assert (lv == want).all()
    ctx.close()
    plain.close()


def test_degree_zero_source():
    beg, adj = G.disjoint(G.path(10), G.csr_from_edges(3, []))
    g = Graph(beg, adj)
    ctx = Context(g)
    say("isolated source")
    lv, recs, st = run_check(ctx, g, 11)
    assert st["n_levels"] == 0 and st["v_cc"] == 1


@pytest.fixture(scope="module")
def rmat10():
    d = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__),
                                           "golden", "rmat10_golden.npz"))
    g = Graph(d["beg_pos"], d["adj"])
    return g, Context(g), d


def test_golden_rmat10(rmat10):
    g, ctx, d = rmat10
    for s, want in zip(d["sources"], d["levels"]):
        for pol in POLICIES:
            say("rmat10", s, pol)
            lv, _, _ = run_check(ctx, g, s, **pol)
            assert (lv == want).all()


def test_heuristic_tags_match_reference(rmat10):
    g, ctx, d = rmat10
    for key, alpha in (("tags_a0p1", 0.1), ("tags_a0p01", 0.01)):
        for s, want in zip(d["sources"], d[key]):
            say("tags", s, alpha)
            _, recs, st = run_check(ctx, g, s, alpha=alpha)
            want = want[want != -9]
            got = np.array([r["tag"] for r in recs], np.int32)
            assert_tags_prefix(got, want, st)


SCHEDULES = [
    [SF, BU, SF], [SF, SF, BU, SF, SF], [BU, BU, SS, SS, SF], [SS, BU, SS],
    [SF, BU, BU, SF], [SS, SS, SF, BU], [BU, SF, BU, SS], [SF, SS, SF, SS, BU, BU, SF],
]


@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "".join("FSB"[m] for m in s))
def test_mode_transitions(sched, rmat10):
    g, ctx, d = rmat10
    gg = Graph(*G.grid(20, 31))
    cg = Context(gg)
    for s in d["sources"][:4]:
        say("sched", sched, s)
        run_check(ctx, g, s, policy=POLICY_SCHEDULE, schedule=sched)
    for s in (0, 317, 619):
        say("sched grid", sched, s)
        run_check(cg, gg, s, policy=POLICY_SCHEDULE, schedule=sched)
    cg.close()


def test_random_schedules_rmat12():
    g = Graph.rmat(12, 16, seed=1012)
    ctx = Context(g)
    rng = np.random.default_rng(5)
    for s in seeded_sources(g, 6):
        sched = list(rng.integers(0, 3, 12))
        say("rand sched", s, sched)
        run_check(ctx, g, s, policy=POLICY_SCHEDULE, schedule=sched)


@pytest.mark.parametrize("kw", [dict(bin_thread_max=1, bin_wave_max=4, bin_block_max=16),
                                dict(bin_thread_max=2, bin_wave_max=2, bin_block_max=2),
                                dict(grid_dim=1),
                                dict(grid_dim=3, bin_wave_max=32, bin_block_max=64),
                                dict(grid_dim=7, bin_thread_max=4, bin_wave_max=64,
                                     bin_block_max=100)],
                         ids=["bins1-4-16", "bins2-2-2", "g1", "g3", "g7"])
def test_launch_shapes_and_bins(kw):
    g = Graph.rmat(11, 16, seed=31)
    ctx = Context(g)
    for s in seeded_sources(g, 3):
        for pol in POLICIES[:4]:
            say("shape", kw, s, pol)
            run_check(ctx, g, s, **kw, **pol)


@pytest.mark.parametrize("grid_dim", [32, 33, 48])
def test_small_frontier_block_counts(grid_dim):
    """Persistent grids around the small-frontier block count (32 active
    blocks): at 32 small levels stay on the whole grid, at 33 one block parks
    and rejoins on every level that needs it, at 48 sixteen do."""
    for g in (Graph.grid(64, 64), Graph.rmat(12, 16, seed=1012)):
        ctx = Context(g)
        for s in seeded_sources(g, 3):
            for pol in POLICIES[:4]:
                say("small blocks", grid_dim, g.n, s, pol)
                run_check(ctx, g, s, grid_dim=grid_dim, **pol)
        ctx.close()


def test_config1_rmat16_16_sources():
    """RMAT scale 16, ef 16, A/B/C/D = .57/.19/.19/.05, 16 seeded sources."""
    g = Graph.rmat(16, 16, seed=1016)
    ctx = Context(g)
    for s in seeded_sources(g, 16):
        say("rmat16", s)
        _, recs, st = run_check(ctx, g, s)
        want, _, _ = O.reference_tags(g.beg_pos, g.adj, int(s), 0.1)
        got = np.array([r["tag"] for r in recs], np.int32)
        assert_tags_prefix(got, want, st)


def test_er_and_grid_small():
    for g in (Graph.er(20000, 80000, seed=7), Graph.grid(64, 64)):
        ctx = Context(g)
        for s in seeded_sources(g, 4):
            for pol in POLICIES:
                say("er/grid", g.n, s, pol)
                run_check(ctx, g, s, **pol)
        ctx.close()


def test_repeat_is_deterministic(rmat10):
    g, ctx, d = rmat10
    s = int(d["sources"][0])
    a = ctx.bfs(s)[0]
    for _ in range(5):
        assert (ctx.bfs(s)[0] == a).all()


def test_device_scan_matches_exclusive_scan(rmat10):
    """The single-pass look-back scan (4096 elements per tile) against the
    oracle's: one and several tiles, ragged ends, 2^24 elements (4096 tiles
    looking back over each other), non-negative and signed values (the
    status words carry signed 62-bit sums), and sums past 2^31."""
    g, ctx, d = rmat10
    rng = np.random.default_rng(11)
    for n in (1, 2, 63, 4095, 4096, 4097, 8191, 100_003, 3_000_000, (1 << 24) - 3, 1 << 24):
        for lo in (0, -(1 << 20)):
            say("scan", n, lo)
            x = rng.integers(lo, 1 << 20, n, dtype=np.int32)
            assert (ctx.exclusive_scan(x) == O.exclusive_scan(x)).all()
    x = np.full(1 << 22, (1 << 31) - 1, np.int32)
    assert (ctx.exclusive_scan(x) == O.exclusive_scan(x)).all()


def test_unvisited_list_is_sorted_unvisited_set():
    g = Graph.rmat(14, 8, seed=99)
    ctx = Context(g)
    rng = np.random.default_rng(1)
    deg = g.degrees()
    for frac in (0.0, 0.3, 0.97, 1.0):
        say("unvisited", frac)
        lv = np.where(rng.random(g.n) < frac, -1, rng.integers(0, 9, g.n)).astype(np.int32)
        want = np.flatnonzero((lv == -1) & (deg > 0)).astype(np.int32)
        assert np.array_equal(ctx.unvisited_list(lv), want)


def test_config2_rmat22_full_size():
    """Full-size RMAT scale 22 (config 2): two sources checked bit-exact
    against the oracle, every source in a seeded set Graph500-validated."""
    g = Graph.rmat(22, 16, seed=1022)
    ctx = Context(g)
    srcs = seeded_sources(g, 6)
    for i, s in enumerate(srcs):
        say("rmat22", s)
        lv, par, recs, st = ctx.bfs(int(s), want_parent=True, rec_cap=64)
        assert O.validate(g.beg_pos, g.adj, int(s), lv, par) == 0
        if i < 2:
            want, _, _ = O.bfs_queue(g.beg_pos, g.adj, int(s))
            assert (lv == want).all()


def test_config3a_er_full_size():
    """Erdos-Renyi n = 2^24, m = 2^27 endpoint pairs (seed 7), symmetrized:
    one source bit-exact against the oracle under both drivers, two more
    Graph500-validated."""
    g = Graph.er(1 << 24, 1 << 27, seed=7)
    ctx = Context(g)
    srcs = seeded_sources(g, 3, seed=7)
    for i, s in enumerate(srcs):
        say("er24", s)
        lv, par, recs, st = ctx.bfs(int(s), want_parent=True, rec_cap=64)
        assert O.validate(g.beg_pos, g.adj, int(s), lv, par) == 0
        if i == 0:
            want, _, _ = O.bfs_queue(g.beg_pos, g.adj, int(s))
            assert (lv == want).all()
            lv0, _, _, _ = ctx.bfs(int(s), persistent=0)
            assert (lv0 == want).all()
            assert st["v_cc"] == (want >= 0).sum()
    ctx.close()


def test_config3b_grid4096_full_size():
    """4096 x 4096 grid (16.7M vertices, 8190 levels from a corner): levels
    equal the Manhattan distance, under both drivers; the persistent kernel
    runs all 8190 levels in one launch."""
    R = C = 4096
    g = Graph.grid(R, C)
    ctx = Context(g)
    rr, cc = np.divmod(np.arange(g.n, dtype=np.int64), C)
    for s, eng in ((0, 1), (R * C - 1, 0), ((R // 2) * C + C // 3, 1)):
        say("grid4096", s, "persistent", eng)
        lv, _, _, st = ctx.bfs(s, persistent=eng)
        sr, sc = divmod(s, C)
        want = (np.abs(rr - sr) + np.abs(cc - sc)).astype(np.int32)
        assert (lv == want).all()
        assert st["max_level"] == want.max()
        print(f"[grid4096] src {s} persistent {eng}: {st['n_levels']} levels, "
              f"{st['t_bfs_s'] * 1e3:.1f} ms, {st['t_bfs_s'] / max(st['n_levels'], 1) * 1e6:.1f} "
              f"us/level", flush=True)
    ctx.close()

"""Learned direction/queue policy: features (CPU), HIP MLP kernels against
the double-precision oracle, probe records and the LEARNED policy path (GPU)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import graphs as G
from xbfs_amd import _capi as C
from xbfs_amd import policy as P
from xbfs_amd import Context, Graph, POLICY_HEURISTIC, POLICY_LEARNED


def test_features_closed_form():
    x = C.policy_features(0, 0, 0, 0, 0, -1)
    assert np.allclose(x, [0, 0, 0, 0, 0, -1, 0, 1])
    x = C.policy_features(3, 1023, 4095, (1 << 20) - 1, 15, 2)
    # log2(1+v)/32 at exact powers of two; frontier-to-unvisited log2 ratios of
    # edges (12 - 4) and vertices (10 - 20); level/16, previous mode BU
    assert np.allclose(x, [10 / 32, 12 / 32, 8 / 32, -10 / 32, 3 / 16, 1, 0, 1], atol=2e-5)
    # the ratios do not depend on the graph's size: scaling every count by 2^10
    # leaves them unchanged (up to the +1 inside the logarithm)
    a = C.policy_features(2, 1 << 12, 1 << 20, 1 << 22, 1 << 28, 0)
    b = C.policy_features(2, 1 << 22, 1 << 30, 1 << 32, 1 << 38, 0)
    assert np.allclose(a[2:4], b[2:4], atol=1e-4) and abs(b[0] - a[0] - 10 / 32) < 1e-4
    x = C.policy_features(1, 5, 5, 5, 5, 0)
    assert x[5] == 0 and x[6] == 1
    # log2(1+v) approximation stays within 1e-4 of the true value
    for v in (2, 3, 6, 100, 12345, 10 ** 9, 2 ** 40 + 12345):
        assert abs(C.policy_features(0, v, 0, 0, 0, 1)[0] * 32 - np.log2(1 + v)) < 1e-4
    # on both sides of every power of two (the exponent comes from the top bit of 1+v)
    for k in range(1, 45):
        for v in (2 ** k - 2, 2 ** k - 1, 2 ** k):
            if v >= 0:
                assert abs(C.policy_features(0, v, 0, 0, 0, 1)[0] * 32 - np.log2(1 + v)) < 1e-4, v


def test_train_specs_are_seeded_and_in_range():
    s = P.train_specs(200)
    assert s == P.train_specs(200)
    assert all(18 <= t["scale"] <= 22 and 8 <= t["ef"] <= 32 and 0.45 <= t["a"] <= 0.62
               for t in s)
    assert all(abs(t["a"] + t["b"] + t["c"] + 0.05 - 1) < 1e-12 for t in s)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 8)).astype(np.float32)
    y = rng.standard_normal((n, 3)).astype(np.float32)
    m = (rng.random((n, 3)) < 0.8).astype(np.float32)
    return x, y, m


@pytest.mark.gpu
def test_mlp_forward_and_grad_match_oracle():
    rng = np.random.default_rng(7)
    p = (rng.standard_normal(C.MLP_NPARAM) * 0.3).astype(np.float32)
    for n in (1, 63, 64, 1000, 5000):
        x, y, m = _data(n, n)
        out = C.mlp_forward(p, x)
        want = O.mlp_forward(p, x)
        assert np.allclose(out, want, rtol=1e-4, atol=1e-4), n
        loss, g = C.mlp_loss_grad(p, x, y, m)
        wl, wg = O.mlp_loss_grad(p, x, y, m)
        assert abs(loss - wl) <= 1e-4 * max(1, abs(wl)), (n, loss, wl)
        # FP32 accumulation vs double: 1e-4 relative to the gradient's scale
        assert np.max(np.abs(g - wg)) <= 1e-4 * max(1.0, np.max(np.abs(wg))), n


@pytest.mark.gpu
def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2000, 8)).astype(np.float32)
    # a smooth target the network can fit
    y = np.stack([x[:, 0] + 0.5 * x[:, 1], np.abs(x[:, 2]), x[:, 3] * x[:, 4] * 0.3], 1)
    y = y.astype(np.float32)
    m = np.ones_like(y)
    p0 = P.init_params(0)
    l0, _ = C.mlp_loss_grad(p0, x, y, m)
    p, curve = C.mlp_train(p0, x, y, m, 1500, 3e-3)
    l1, _ = C.mlp_loss_grad(p, x, y, m)
    assert abs(curve[0] - l0) < 1e-4 * max(1, l0)
    assert l1 < 0.1 * l0 and curve[-1] < curve[0]


@pytest.mark.gpu
def test_probe_records_and_learned_path_exact():
    g = Graph.rmat(14, 16, seed=1014)
    ctx = Context(g)
    srcs = P.seeded_sources(g.degrees(), 3)
    for s in srcs:
        s = int(s)
        want, _, _ = O.bfs_queue(g.beg_pos, g.adj, s)
        recs, st = ctx.probe(s, reps=1)
        assert st["v_cc"] == (want >= 0).sum()
        cnt, nv, _ = O.level_counts(g.beg_pos, want)
        for r in recs:
            assert r["m_f"] == cnt[r["level"]] and r["n_f"] == nv[r["level"]]
            assert all(t > 0 for t in r["t_mode"])
        best, st2 = ctx.probe(s, follow_best=True, reps=1)
        assert all(r["mode"] == int(np.argmin(r["t_mode"])) for r in best)
        # learned policy without weights falls back to the heuristic
        _, _, rl, _ = ctx.bfs(s, rec_cap=64, policy=POLICY_LEARNED)
        _, _, rh, _ = ctx.bfs(s, rec_cap=64, policy=POLICY_HEURISTIC)
        assert [r["mode"] for r in rl] == [r["mode"] for r in rh]
    # with (random) weights the learned path still yields exact levels
    rng = np.random.default_rng(1)
    ctx.load_policy((rng.standard_normal(C.MLP_NPARAM) * 0.5).astype(np.float32))
    for s in srcs:
        for pers in (1, 0):
            lv, par, _, _ = ctx.bfs(int(s), want_parent=True, policy=POLICY_LEARNED,
                                    persistent=pers)
            want, _, _ = O.bfs_queue(g.beg_pos, g.adj, int(s))
            assert (lv == want).all()
            assert O.validate(g.beg_pos, g.adj, int(s), lv, par) == 0


@pytest.mark.gpu
def test_persistent_decisions_follow_the_model():
    """k_persistent evaluates the model with one wave from weights laid out
    transposed in LDS: at every level its choice is the argmin of the same
    model evaluated by xbfs_mlp_forward on that level's features (levels
    whose two best predictions lie within 1e-3 are skipped as ties)."""
    g = Graph.rmat(18, 16, seed=5)
    ctx = Context(g)
    srcs = P.seeded_sources(g.degrees(), 4, seed=3)
    rng = np.random.default_rng(2)
    weights = [(rng.standard_normal(C.MLP_NPARAM) * 0.5).astype(np.float32)]
    if os.path.exists(P.WEIGHTS):
        weights.append(P.load())
    checked = 0
    for w in weights:
        ctx.load_policy(w)
        for s in srcs:
            _, _, recs, _ = ctx.bfs(int(s), want_levels=False, rec_cap=64, policy=POLICY_LEARNED)
            X = np.stack([C.policy_features(r["level"], r["n_f"], r["m_f"], r["n_u"], r["m_u"],
                                            recs[i - 1]["mode"] if i else -1)
                          for i, r in enumerate(recs)])
            pred = C.mlp_forward(w, X)
            for r, y in zip(recs, pred):
                srt = np.sort(y)
                if r["policy"] != POLICY_LEARNED or srt[1] - srt[0] < 1e-3:
                    continue
                assert r["mode"] == int(np.argmin(y)), (s, r, y)
                checked += 1
    assert checked >= 8


@pytest.mark.gpu
def test_trained_policy_gain_on_held_out_graph():
    """The committed weights (trained by xbfs_amd/train_policy.py on RMAT
    scale 18-22) against the heuristic on a held-out graph (seed 777, outside
    the training set): the end-to-end time over 8 sources, each the median of
    5 traversals, must be at least 1.1x shorter, the target the learned
    policy is built for."""
    if not os.path.exists(P.WEIGHTS):
        pytest.skip("no trained weights committed yet")
    g = Graph.rmat(20, 16, seed=777)
    ctx = Context(g)
    srcs = P.seeded_sources(g.degrees(), 8, seed=9)
    tot = P.evaluate(ctx, srcs, P.load(), reps=5)
    gain = tot["heuristic"] / tot["learned"]
    print(f"[policy] held-out RMAT-20: heuristic {tot['heuristic']*1e3:.3f} ms, learned "
          f"{tot['learned']*1e3:.3f} ms, oracle-best {tot['oracle_best']*1e3:.3f} ms, "
          f"gain {gain:.3f}x", flush=True)
    assert gain >= 1.1


@pytest.mark.gpu
def test_learned_policy_domain_guard():
    """The learned policy decides only inside the domain it was trained on:
    skewed graphs (largest degree at least 16 times the mean degree of the
    non-isolated vertices) and levels below 16.  On a grid and an
    Erdos-Renyi graph, and on the deep levels of a skewed graph with a long
    path, the heuristic decides (records carry policy 0, modes equal the
    heuristic's); levels stay exact everywhere."""
    rng = np.random.default_rng(4)
    w = P.load() if os.path.exists(P.WEIGHTS) else \
        (rng.standard_normal(C.MLP_NPARAM) * 0.5).astype(np.float32)
    star_path = G.csr_from_edges(1040, [(0, i) for i in range(1, 1000)] +
                                 [(999 + i, 1000 + i) for i in range(40)])
    cases = {"grid": G.grid(64, 64), "er": None, "star_path": star_path}
    for name, csr in cases.items():
        g = Graph.er(1 << 14, 1 << 17, seed=3) if name == "er" else Graph(*csr)
        ctx = Context(g)
        ctx.load_policy(w)
        for s in (0, g.n // 2, g.n - 1):
            for pers in (1, 0):
                lv, _, rl, _ = ctx.bfs(s, rec_cap=128, policy=POLICY_LEARNED, persistent=pers)
                _, _, rh, _ = ctx.bfs(s, rec_cap=128, policy=POLICY_HEURISTIC, persistent=pers)
                want, _, _ = O.bfs_queue(g.beg_pos, g.adj, s)
                assert (lv == want).all()
                if name == "star_path":
                    assert len(rl) > 16
                    assert all(r["policy"] == POLICY_LEARNED for r in rl if r["level"] <= 15)
                    assert all(r["policy"] == 0 for r in rl if r["level"] > 15)
                else:
                    assert all(r["policy"] == 0 for r in rl)
                    assert [r["mode"] for r in rl] == [r["mode"] for r in rh]
        ctx.close()


def test_committed_samples_match_library_features():
    """The committed training samples carry their raw level states; their
    stored features must be what the library computes from those states, so
    the committed weights and the feature code cannot drift apart."""
    path = os.path.join(os.path.dirname(P.WEIGHTS), "policy_samples.npz")
    if not os.path.exists(path):
        pytest.skip("no committed samples")
    d = np.load(path)
    assert set(d.files) >= {"X", "Y", "M", "R"}
    idx = np.random.default_rng(0).choice(d["R"].shape[0], 500, replace=False)
    assert np.allclose(P.features_of(d["R"][idx]), d["X"][idx], atol=1e-6)
    assert d["Y"].shape == (d["X"].shape[0], 3) and np.isfinite(d["Y"]).all()

"""Learned direction/queue policy: training data from XBFS's own per-level
probes, training on the GPU with the library's MLP kernels, and evaluation
against the reference heuristic and the per-level oracle-best schedule.

No reference counterpart: the model replaces the decision block
bfs_adaptiveFQ_wb_async.cuh:1047-1078 when it is loaded (policy LEARNED);
without weights, or on a non-finite prediction, the library falls back to
that heuristic.

Training set (seeded): 200 RMAT graphs, graph i drawn with
numpy.random.default_rng(i): scale ~ U{18..22}, A ~ U[0.45, 0.62], D = 0.05,
B = C = (1 - A - D) / 2, edge factor ~ U{8..32}, generator seed i.
Labels: log2 of the microseconds each mode takes at each level, measured by
xbfs_bfs_probe (three modes from one device snapshot, median of `reps`
cooperative single-level launches).  Half of each graph's sources follow the
heuristic's choices, half the fastest measured mode, so the data covers the
states either policy visits.
"""
import json
import os
import time

import numpy as np

from ._capi import (MLP_NPARAM, POLICY_HEURISTIC, POLICY_LEARNED, POLICY_SCHEDULE, Context,
                    Graph, mlp_train, policy_features)

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS = os.path.join(DATA_DIR, "policy_mlp.f32")
META = os.path.join(DATA_DIR, "policy_mlp.json")


def train_specs(count=200, scales=(18, 22)):
    specs = []
    for i in range(count):
        rng = np.random.default_rng(i)
        scale = int(rng.integers(scales[0], scales[1] + 1))
        A = float(rng.uniform(0.45, 0.62))
        D = 0.05
        B = C = (1.0 - A - D) / 2.0
        ef = int(rng.integers(8, 33))
        specs.append(dict(scale=scale, ef=ef, a=A, b=B, c=C, seed=i))
    return specs


def seeded_sources(deg, k, seed=42):
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(deg > 0)
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


RAW_FIELDS = ("level", "n_f", "m_f", "n_u", "m_u", "prev_mode")


def records_to_samples(recs):
    """Probe records -> (features, labels, raw level state).  The raw state
    (RAW_FIELDS per level) lets recorded samples be re-featurised later."""
    X, Y, R = [], [], []
    for r in recs:
        X.append(policy_features(*(r[k] for k in RAW_FIELDS)))
        Y.append([np.log2(max(t, 1e-7) * 1e6) for t in r["t_mode"]])
        R.append([r[k] for k in RAW_FIELDS])
    return X, Y, R


def features_of(R):
    """Features of recorded raw level states, by the library's definition."""
    return np.asarray([policy_features(*(int(v) for v in row)) for row in R], np.float32)


def collect(specs, sources=4, reps=3, log=None):
    """Returns (X, Y, M, R): features, labels, label mask, raw level states."""
    X, Y, R = [], [], []
    for k, sp in enumerate(specs):
        t0 = time.time()
        g = Graph.rmat(sp["scale"], sp["ef"], sp["a"], sp["b"], sp["c"], seed=sp["seed"])
        ctx = Context(g)
        srcs = seeded_sources(g.degrees(), sources, seed=1000 + sp["seed"])
        for j, s in enumerate(srcs):
            recs, _ = ctx.probe(int(s), follow_best=(j % 2 == 1), reps=reps)
            x, y, raw = records_to_samples(recs)
            X += x
            Y += y
            R += raw
        ctx.close()
        del g
        if log:
            log(f"[policy] graph {k}: scale {sp['scale']} ef {sp['ef']} A {sp['a']:.3f}: "
                f"{len(X)} samples ({time.time() - t0:.1f}s)")
    X = np.asarray(X, np.float32)
    Y = np.asarray(Y, np.float32)
    return X, Y, np.ones_like(Y), np.asarray(R, np.int64)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_NPARAM, np.float32)
    o = 0
    for fan_in, fan_out in ((8, 32), (32, 32), (32, 3)):
        w = rng.standard_normal((fan_out, fan_in)) * np.sqrt(2.0 / fan_in)
        p[o:o + w.size] = w.ravel()
        o += w.size + fan_out  # biases start at zero
    return p


def train(X, Y, M, steps=4000, lr=3e-3, seed=0, init=None):
    """Full-batch Adam on the GPU (library kernels).  Returns (params, curve).
    init: starting weights (fine-tuning); default a fresh initialisation."""
    if init is not None:
        p = np.array(init, dtype=np.float32, copy=True)
    else:
        # start the output biases at the label means so early steps fit structure
        p = init_params(seed)
        p[-3:] = (Y * M).sum(0) / np.maximum(M.sum(0), 1)
    return mlp_train(p, X, Y, M, steps, lr)


def save(params, meta):
    os.makedirs(DATA_DIR, exist_ok=True)
    np.asarray(params, np.float32).tofile(WEIGHTS)
    with open(META, "w") as f:
        json.dump(meta, f, indent=1)


def load(path=WEIGHTS):
    w = np.fromfile(path, dtype=np.float32)
    if w.size != MLP_NPARAM:
        raise ValueError(f"{path}: {w.size} floats, expected {MLP_NPARAM}")
    return w


def evaluate(ctx, sources, weights, reps=3):
    """Total traversal time (persistent kernel) over `sources` for the
    heuristic, the learned policy, and the per-level oracle-best schedule
    (the fastest measured mode at every level, replayed as a schedule).
    Returns a dict of seconds (each the median over `reps` runs per source)."""
    ctx.load_policy(weights)
    tot = {"heuristic": 0.0, "learned": 0.0, "oracle_best": 0.0}
    for s in sources:
        s = int(s)
        recs, _ = ctx.probe(s, follow_best=True, reps=reps)
        sched = [r["mode"] for r in recs] or [0]
        for name, kw in (("heuristic", dict(policy=POLICY_HEURISTIC)),
                         ("learned", dict(policy=POLICY_LEARNED)),
                         ("oracle_best", dict(policy=POLICY_SCHEDULE, schedule=sched))):
            ts = [ctx.bfs(s, want_levels=False, **kw)[3]["t_bfs_s"] for _ in range(reps)]
            tot[name] += float(np.median(ts))
    return tot

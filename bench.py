"""XBFS benchmark: adaptive direction-optimizing BFS on one MI355X.

Workload (the configuration the headline metric is quoted on): RMAT scale 24,
edge factor 16, A/B/C/D = 0.57/0.19/0.19/0.05, seeded, symmetrized,
self-loops and duplicates removed, vertices relabelled; 64 seeded sources
drawn from non-isolated vertices.  One step = one BFS from each of the 64
sources with the graph resident in HBM.

value = harmonic mean over the 64 sources of GTEPS = m_cc / t_bfs / 1e9,
where m_cc is the sum of degrees of the vertices reached and t_bfs is the
device time (HIP events) from level initialisation to the final frontier
check, averaged over the timed steps.  Each step runs every source under the
reference heuristic and, when trained weights are present
(xbfs_amd/data/policy_mlp.f32), under the learned policy; value is the
learned policy's figure and "policies" reports both, their ratio, and the
per-level oracle-best schedule measured on a subset of the sources.
"without_component_end" times the same sources on a context built with
XBFS_COMPONENT=off (every traversal expands levels until its frontier is
empty instead of stopping once the source's component is covered).

roofline: the bytes one traversal's algorithm reads and writes as executed
(direction-optimised: the work each level's mode actually did, from the level
records: top-down levels expand m_f edges, bottom-up levels report the
candidates they probed and the rows and adjacency words they scanned; byte
costs per unit in traversal_bytes()), divided by the device time of the
launch that runs the traversal, against the 8 TB/s HBM peak.  The byte count
that charges every edge of the reached component, 8*m_cc + 20*V_cc + 4*V (the
TEPS-equivalent figure), is reported beside it.  roofline.traffic is the HBM
traffic per launch measured by rocprofv3 counters on this workload (the
profile is named in traffic_source); it is reported only while the launch
time stays within 10 % of the profiled one.

cpu_baseline: OpenMP top-down BFS (the oracle's CPU baseline) on the same
sources, on every core this process may run on (its affinity set), for as
many sources as fit a time budget (--cpu-budget seconds).  cpu_serial: the
reference's own CPU path (the level-synchronous scan BFS of
bfs_gpu_update.cu:104-127, restated in the oracle) on one core, first source.

batch64 (reported beside the headline, --batch 1): RMAT scale 22 ef16 (seed
1022), 64 seeded sources; device time of one 64-source bit-parallel batch (one
launch) against the sum of the 64 single-source traversals of the same
sources under the learned policy, the fastest single-source path (medians of
3; the heuristic's sum beside it).  multigraph: 256 RMAT graphs (scale
U{10..16}, edge factor U{4..32}, seeds 10000+i), one seeded source each; one
multi-source traversal of the packed graphs against 256 single-source
traversals on per-graph contexts, likewise under both policies, with the
autotuner's setting for the packed graph.

Other graphs, parity and figures only, never the headline: --graph er
(Erdos-Renyi, 2^scale vertices, 2^scale * ef / 2 edges) and --graph grid
(2^(scale/2) x 2^(scale/2) lattice, thousands of levels), e.g.
`python bench.py --graph grid --sources 4 --cpu-serial 0 --oracle-best 0
--no-component 0 --batch 0`.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12

# HBM bytes per k_persistent launch on the default workload (RMAT-24 ef16, the
# 64 seeded sources, learned policy), from two rocprofv3 --pmc passes
# (profiles/k_persistent_rmat24_pmc_traffic.md): mean FETCH_SIZE and
# WRITE_SIZE in KiB per dispatch over the traversal launches (the context's
# create-time traversal left out), each scaled by what the same counter reads
# for a plain 1 GiB device copy (FETCH_SIZE reads half of the bytes a copy
# moves: factor 2.0; WRITE_SIZE reads all of them: factor 1.0).  kernel_us is
# the mean launch time of the kernel trace taken with them; the figure is
# reported only while the measured launch time stays within 10 % of it.
PMC_TRAFFIC = {"workload": (24, 16, 64, "learned"), "fetch_kib": 405026.9, "fetch_factor": 1.99994,
               "write_kib": 258182.3, "write_factor": 1.0, "kernel_us": 405.5,
               "profile": "profiles/k_persistent_rmat24_pmc_traffic.md"}

MODE_SF, MODE_SS, MODE_BU = 0, 1, 2


def traversal_bytes(recs, n, v_cc, parents=False, per_level=None):
    """Bytes one traversal reads and writes as executed, from its level
    records.  W = one bitmap (n/8 bytes, rounded up to 64-bit words).
      initialisation: 4n level words (+4n parents), visited + 4 frontier rings (5W)
      every level: one frontier ring zeroed (W)
      sweep of F_L into the binned queue (a top-down level whose queue is not
        exact): W of bitmap, 16 B offsets + 4 B queue entry per frontier vertex;
        after a single-scan level the sweep reads the 4n level words instead
        and writes F_L and the visited words (2W)
      top-down: 16 B offsets per frontier vertex, 8 B per expanded edge (4 B
        adjacency word + 4 B level probe), 4 B level write per discovered
        vertex; scan-free also 16 B of bitmap atomics, 16 B offsets and a 4 B
        queue entry per discovered vertex
      bottom-up: candidate words (mask and visited, 2W), 12 B per candidate
        probed (8 B hub record + 4 B bitmap word), 16 B offsets per scanned row,
        8 B per scanned adjacency word (word + bitmap probe), 4 B level write
        per discovered vertex, visited and next-frontier words written (2W)
      parents: 4 B per discovered vertex.
    per_level (a list): receives (level, mode, bytes) of each level; the
    initialisation is the entry of level -1, so the entries sum to the total."""
    W = 8 * ((n + 63) // 64)
    b = 4 * n + 5 * W + (4 * n if parents else 0)
    if per_level is not None:
        per_level.append((-1, -1, b))
    prev = -1
    seen = sum(r["n_f"] for r in recs)
    for i, r in enumerate(recs):
        b0 = b
        mode, nf, mf = r["mode"], r["n_f"], r["m_f"]
        nd = recs[i + 1]["n_f"] if i + 1 < len(recs) else max(v_cc - seen, 0)
        b += W
        if prev == MODE_SS:
            b += 4 * n + 2 * W + 20 * nf
        elif mode != MODE_BU and i > 0 and prev != MODE_SF:
            b += W + 20 * nf
        if mode == MODE_BU:
            b += 4 * W + 12 * r["bu_cand"] + 16 * r["bu_scan_v"] + 8 * r["bu_scan_e"] + 4 * nd
        else:
            b += 16 * nf + 8 * mf + 4 * nd
            if mode == MODE_SF:
                b += 36 * nd
        if parents:
            b += 4 * nd
        prev = mode
        if per_level is not None:
            per_level.append((r["level"], mode, b - b0))
    return b


def hmean(x):
    x = np.asarray(x, np.float64)
    return float(len(x) / np.sum(1.0 / x))


def pick_sources(deg, k, seed=42):
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(deg > 0)
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


def batch_fields(reps=3):
    """The two batched workloads; returns (batch64 dict, multigraph dict).
    Each batch is compared with one-at-a-time traversals of the same sources
    under the fastest single-source path (the learned policy when its weights
    are present) and, beside it, under the reference heuristic."""
    from xbfs_amd import Context, Graph, default_opts, POLICY_HEURISTIC, POLICY_LEARNED
    from xbfs_amd import policy as P
    w = P.load() if os.path.exists(P.WEIGHTS) else None
    pols = [("heuristic", default_opts(policy=POLICY_HEURISTIC))]
    if w is not None:
        pols.append(("learned", default_opts(policy=POLICY_LEARNED)))
    best = pols[-1][0]

    def seq_times(ctx, srcs):
        out = {}
        for name, o in pols:
            out[name] = sum(float(np.median([ctx.bfs(int(s), opts=o, want_levels=False)[3]["t_bfs_s"]
                                             for _ in range(reps)])) for s in srcs)
        return out

    g = Graph.rmat(22, 16, 0.57, 0.19, 0.19, seed=1022)
    ctx = Context(g, device=0)
    if w is not None:
        ctx.load_policy(w)
    srcs = pick_sources(g.degrees(), 64)
    ctx.bfs_batch64(srcs, want_levels=False)
    for s in srcs:
        ctx.bfs(int(s), opts=pols[-1][1], want_levels=False)
    seq = seq_times(ctx, srcs)
    tb, st = [], None
    for _ in range(reps):
        _, st = ctx.bfs_batch64(srcs, want_levels=False)
        tb.append(st["t_bfs_s"])
    tb = float(np.median(tb))
    b64 = {"workload": "RMAT scale 22 ef16 seed 1022, 64 seeded sources", "n": g.n, "m": g.m,
           "batch_ms": round(tb * 1e3, 3), "sequential_ms": round(seq[best] * 1e3, 3),
           "sequential_policy": best, "speedup": round(seq[best] / tb, 3),
           "sequential_heuristic_ms": round(seq["heuristic"] * 1e3, 3),
           "speedup_vs_heuristic": round(seq["heuristic"] / tb, 3),
           "aggregate_gteps": round(st["m_cc"] / tb / 1e9, 1), "levels": st["n_levels"],
           "launches_per_batch": 1}
    ctx.close()
    del g, ctx

    gs, ls = [], []
    for i in range(256):
        rng = np.random.default_rng(10000 + i)
        scale = int(rng.integers(10, 17))
        ef = int(rng.integers(4, 33))
        gi = Graph.rmat(scale, ef, 0.57, 0.19, 0.19, seed=10000 + i)
        gs.append(gi)
        ls.append(int(pick_sources(gi.degrees(), 1, seed=10000 + i)[0]))
    seq = {name: 0.0 for name, _ in pols}
    for gi, s in zip(gs, ls):
        ci = Context(gi, device=0)
        if w is not None:
            ci.load_policy(w)
        ci.bfs(s, want_levels=False)
        for name, t in seq_times(ci, [s]).items():
            seq[name] += t
        ci.close()
    u, off = Graph.concat(gs)
    cu = Context(u, device=0)
    seeds = (off[:-1] + np.asarray(ls)).astype(np.int32)
    tuned, trials = cu.autotune(seeds, multi=True, reps=reps)
    tm, st = [], None
    for _ in range(reps):
        _, st = cu.bfs_multigraph(off, ls, opts=tuned, want_levels=False)
        tm.append(st["t_bfs_s"])
    tm = float(np.median(tm))
    t_default = trials[0]["t_s"]
    mg = {"workload": "256 RMAT graphs, scale U{10..16}, ef U{4..32}, seeds 10000+i, one source each",
          "n": u.n, "m": u.m, "batch_ms": round(tm * 1e3, 3), "sequential_ms": round(seq[best] * 1e3, 3),
          "sequential_policy": best, "speedup": round(seq[best] / tm, 3),
          "sequential_heuristic_ms": round(seq["heuristic"] * 1e3, 3),
          "aggregate_gteps": round(st["m_cc"] / tm / 1e9, 1),
          "levels": st["n_levels"], "autotune": {
              "bin_thread_max": tuned.bin_thread_max, "bin_wave_max": tuned.bin_wave_max,
              "bin_block_max": tuned.bin_block_max, "grid_dim": tuned.grid_dim,
              "trials": len(trials), "default_ms": round(t_default * 1e3, 3)}}
    cu.close()
    return b64, mg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--graph", choices=("rmat", "er", "grid"), default="rmat",
                    help="rmat: RMAT(scale, ef), A/B/C/D = .57/.19/.19/.05, seed 1000 + scale; "
                         "er: Erdos-Renyi, 2^scale vertices and 2^scale * ef / 2 edges, seed 7; "
                         "grid: 2^(scale/2) x 2^(scale/2) lattice (use --cpu-serial 0 "
                         "--oracle-best 0: both walk every level)")
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--persistent", type=int, default=-1,
                    help="1 = device-side level loop, 0 = host-driven, -1 = library default")
    ap.add_argument("--cpu-sample", type=int, default=64,
                    help="sources timed for the CPU baseline at most (0 = skip)")
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU baseline work: sources are timed in order until "
                         "the budget is spent (at least one)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (0 = every core this process "
                         "may run on)")
    ap.add_argument("--cpu-serial", type=int, default=1,
                    help="sources timed for the single-core serial reference path (0 = skip)")
    ap.add_argument("--oracle-best", type=int, default=16,
                    help="sources on which the per-level oracle-best schedule is measured")
    ap.add_argument("--no-component", type=int, default=1,
                    help="also time the sources with XBFS_COMPONENT=off (reported beside)")
    ap.add_argument("--compare", type=int, default=1,
                    help="1 = also time the heuristic policy; 0 = only the main policy "
                         "(profiling runs, so every k_persistent launch is the measured one)")
    ap.add_argument("--batch", type=int, default=1,
                    help="1 = also measure the 64-source batch and the 256-graph multigraph batch")
    ap.add_argument("--batch-only", type=int, default=0, metavar="N",
                    help="run only N launches of the 64-source batch on RMAT-22 ef16 (after one "
                         "warmup launch), print their times, and exit (profiling k_b64_persistent)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="run only a plain device-to-device copy of BYTES bytes, 5 times "
                         "(scale factor for the FETCH_SIZE / WRITE_SIZE counters), and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(sharding a traversal over several GPUs is not built)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy > 0:
        import torch
        a = torch.empty(args.calibrate_copy // 4, dtype=torch.int32, device="cuda:0").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": a.numel() * 4, "launches": 5}))
        return

    if args.batch_only > 0:
        from xbfs_amd import Context, Graph
        g = Graph.rmat(22, 16, 0.57, 0.19, 0.19, seed=1022)
        ctx = Context(g, device=0)
        srcs = pick_sources(g.degrees(), 64)
        ctx.bfs_batch64(srcs, want_levels=False)
        ts, st = [], None
        for _ in range(args.batch_only):
            _, st = ctx.bfs_batch64(srcs, want_levels=False)
            ts.append(st["t_bfs_s"])
        ctx.close()
        print(json.dumps({"batch64_only": {
            "workload": "RMAT scale 22 ef16 seed 1022, 64 seeded sources", "n": g.n, "m": g.m,
            "launches": len(ts) + 1, "batch_ms_median": round(float(np.median(ts)) * 1e3, 3),
            "levels": st["n_levels"], "level_rows_bytes": 64 * 4 * g.n}}))
        return

    from xbfs_amd import Context, Graph, default_opts, POLICY_HEURISTIC, POLICY_LEARNED
    from xbfs_amd import policy as P
    t0 = time.time()
    if args.graph == "rmat":
        g = Graph.rmat(args.scale, args.ef, 0.57, 0.19, 0.19, seed=1000 + args.scale)
        gname = f"RMAT-{args.scale} ef{args.ef}"
        gdesc = f"RMAT scale {args.scale} ef{args.ef} A/B/C/D=.57/.19/.19/.05"
    elif args.graph == "er":
        g = Graph.er(1 << args.scale, (1 << args.scale) * args.ef // 2, seed=7)
        gname = f"ER-{args.scale} ef{args.ef}"
        gdesc = f"Erdos-Renyi n=2^{args.scale} m=2^{args.scale}*{args.ef}/2 seed 7"
    else:
        side = 1 << (args.scale // 2)
        g = Graph.grid(side, side)
        gname = gdesc = f"{side}x{side} grid"
    t_gen = time.time() - t0
    deg = g.degrees()
    srcs = pick_sources(deg, args.sources)
    t_c0 = time.perf_counter()
    ctx = Context(g, device=0)
    t_create = time.perf_counter() - t_c0
    kw = {}
    if args.persistent >= 0:
        kw["persistent"] = args.persistent
    have_model = os.path.exists(P.WEIGHTS)
    if have_model:
        ctx.load_policy(P.load())
    pols = [("heuristic", default_opts(policy=POLICY_HEURISTIC, **kw))]
    if have_model:
        pols.append(("learned", default_opts(policy=POLICY_LEARNED, **kw)))
        if not args.compare:
            pols = pols[1:]
    opts = pols[-1][1]
    print(f"[bench] {gname}: n={g.n} m={g.m} gen {t_gen:.1f}s; "
          f"{len(srcs)} sources; policies {[p for p, _ in pols]}", file=sys.stderr, flush=True)

    def one_step():
        ts = {p: [] for p, _ in pols}
        st_all = []
        for s in srcs:
            for p, o in pols:
                _, _, _, st = ctx.bfs(int(s), opts=o, want_levels=False)
                ts[p].append(st["t_bfs_s"])
            st_all.append(st)
        return ts, st_all

    for _ in range(args.warmup):
        one_step()
    t_dev = {p: np.zeros(len(srcs)) for p, _ in pols}
    st_ref = None
    w0 = time.perf_counter()
    for k in range(args.steps):
        ts, sts = one_step()
        for p in ts:
            t_dev[p] += np.asarray(ts[p])
        st_ref = sts
        print(f"[bench] step {k} done", file=sys.stderr, flush=True)
    wall = time.perf_counter() - w0
    for p in t_dev:
        t_dev[p] /= max(args.steps, 1)
    m_cc = np.array([s["m_cc"] for s in st_ref], np.float64)
    v_cc = np.array([s["v_cc"] for s in st_ref], np.float64)
    main_pol = pols[-1][0]
    gteps = m_cc / t_dev[main_pol] / 1e9
    value = hmean(gteps)
    # bytes as executed, from each traversal's level records (one untimed
    # pass; the decisions and the work counts do not depend on timing)
    b_do = np.zeros(len(srcs))
    lvl_bytes = {}
    for i, s in enumerate(srcs):
        _, _, recs, st1 = ctx.bfs(int(s), opts=opts, want_levels=False, rec_cap=256)
        split = []
        b_do[i] = traversal_bytes(recs, g.n, st1["v_cc"], per_level=split)
        for lv_, md_, by_ in split:
            key = (lv_, md_)
            lvl_bytes[key] = lvl_bytes.get(key, 0) + by_
    b_teps = 8 * m_cc + 20 * v_cc + 4 * g.n
    t_mean = float(t_dev[main_pol].mean())
    achieved = float(b_do.mean()) / t_mean
    traffic, traffic_note = None, None
    if args.graph == "rmat" and PMC_TRAFFIC["workload"] == (args.scale, args.ef, len(srcs), main_pol):
        drift = abs(t_mean * 1e6 - PMC_TRAFFIC["kernel_us"]) / PMC_TRAFFIC["kernel_us"]
        if drift <= 0.10:
            traffic = 1024.0 * (PMC_TRAFFIC["fetch_kib"] * PMC_TRAFFIC["fetch_factor"] +
                                PMC_TRAFFIC["write_kib"] * PMC_TRAFFIC["write_factor"])
        else:
            traffic_note = (f"stale: launch time {t_mean * 1e6:.1f} us differs from the profiled "
                            f"{PMC_TRAFFIC['kernel_us']} us by {drift * 100:.0f} %")
    policy_cmp = {}
    if "heuristic" in t_dev:
        policy_cmp["heuristic_gteps_hmean"] = round(hmean(m_cc / t_dev["heuristic"] / 1e9), 3)
        policy_cmp["heuristic_total_ms"] = round(float(t_dev["heuristic"].sum()) * 1e3, 3)
    if have_model and "heuristic" in t_dev:
        policy_cmp["learned_gteps_hmean"] = round(value, 3)
        policy_cmp["learned_total_ms"] = round(float(t_dev["learned"].sum()) * 1e3, 3)
        policy_cmp["learned_speedup_vs_heuristic"] = round(
            float(t_dev["heuristic"].sum() / t_dev["learned"].sum()), 4)
        if args.oracle_best > 0:
            sub = srcs[: args.oracle_best]
            tot = P.evaluate(ctx, sub, P.load(), reps=3)
            policy_cmp["oracle_best_sample"] = {
                "sources": int(len(sub)), "heuristic_ms": round(tot["heuristic"] * 1e3, 3),
                "learned_ms": round(tot["learned"] * 1e3, 3),
                "oracle_best_ms": round(tot["oracle_best"] * 1e3, 3)}

    # The same sources on a context built with XBFS_COMPONENT=off, where every
    # traversal runs until its frontier is empty (no create-time component of
    # the largest-degree vertex): reported beside the headline.
    no_comp = None
    if args.no_component:
        prev = os.environ.get("XBFS_COMPONENT")
        os.environ["XBFS_COMPONENT"] = "off"
        try:
            ctx_nc = Context(g, device=0)
        finally:
            if prev is None:
                del os.environ["XBFS_COMPONENT"]
            else:
                os.environ["XBFS_COMPONENT"] = prev
        if have_model:
            ctx_nc.load_policy(P.load())
        t_nc = np.zeros(len(srcs))
        for _ in range(max(args.steps, 1)):
            for i, s in enumerate(srcs):
                t_nc[i] += ctx_nc.bfs(int(s), opts=opts, want_levels=False)[3]["t_bfs_s"]
        t_nc /= max(args.steps, 1)
        ctx_nc.close()
        no_comp = {"gteps_hmean": round(hmean(m_cc / t_nc / 1e9), 3),
                   "total_ms": round(float(t_nc.sum()) * 1e3, 3),
                   "note": "XBFS_COMPONENT=off: no create-time component, every traversal "
                           "expands levels until its frontier is empty"}

    # correctness spot check of one source against the oracle validator
    from oracle import oracle as O
    s0 = int(srcs[0])
    lv, par, _, _ = ctx.bfs(s0, opts=opts, want_parent=True)
    valid = O.validate(g.beg_pos, g.adj, s0, lv, par) == 0

    cpu = None
    if args.cpu_sample > 0:
        usable = len(os.sched_getaffinity(0))
        nthr = args.cpu_threads or usable
        cg, spent = [], 0.0
        for s in srcs[: args.cpu_sample]:
            lvc, tc = O.omp_td_bfs(g.beg_pos, g.adj, int(s), nthr)
            cg.append(float(deg[lvc >= 0].sum()) / tc / 1e9)
            spent += tc
            if spent >= args.cpu_budget:
                break
        cpu = {"value": hmean(cg), "unit": "GTEPS", "cores": nthr, "kind": "port",
               "sample": f"OpenMP top-down BFS (frontier queue, CAS on levels), first {len(cg)} "
                         f"of the {len(srcs)} sources ({spent:.1f} s of traversal, budget "
                         f"{args.cpu_budget:g} s), same graph",
               "host_cpus": os.cpu_count(), "affinity_cpus": usable}
    cpu_serial = None
    if args.cpu_serial > 0:
        cs = []
        for s in srcs[: args.cpu_serial]:
            t0 = time.perf_counter()
            lvs, _ = O.bfs_scan(g.beg_pos, g.adj, int(s))
            ts = time.perf_counter() - t0
            cs.append(float(deg[lvs >= 0].sum()) / ts / 1e9)
        cpu_serial = {"value": hmean(cs), "unit": "GTEPS", "cores": 1, "kind": "port",
                      "sample": f"reference serial scan BFS (bfs_gpu_update.cu:104-127), "
                                f"first {len(cs)} of the {len(srcs)} sources, same graph"}

    ctx.close()  # before the batch contexts are made, and never left to interpreter teardown

    b64 = mg = None
    if args.batch:
        b64, mg = batch_fields()
        print(f"[bench] batch64 {b64['speedup']}x, multigraph {mg['speedup']}x", file=sys.stderr,
              flush=True)

    out = {
        "metric": "GTEPS (harm. mean, 64 src) RMAT-24 ef16; learned vs heuristic switch; % HBM",
        "value": round(value, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / max(args.steps, 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": f"synthetic ({'seeded ' if args.graph != 'grid' else ''}{args.graph.upper()} "
                f"graph generated in-process)",
        "config": {"workload": f"{gdesc}, {len(srcs)} seeded sources, single-source BFS each",
                   "n": g.n, "m": g.m, "persistent": int(opts.persistent),
                   "policy": ("learned MLP cost model (heuristic fallback)" if have_model
                              else "heuristic (reference alpha rule, alpha=0.1)"),
                   "levels_valid": bool(valid),
                   "preprocessing": {
                       "context_create_s": round(t_create, 3),
                       "note": "once per graph, outside the timed region: upload, the "
                               "degree-ordered bottom-up copy of the adjacency, the hub table, "
                               "and one traversal from the largest-degree vertex whose "
                               "component every traversal from a member then uses to stop "
                               "early and to skip other components' bottom-up candidates; "
                               "without_component_end is the same run without that component"}},
        "policies": policy_cmp,
        "without_component_end": no_comp,
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "algorithmic_bytes_per_launch": round(float(b_do.mean())),
                     "bytes_model": "as executed, per level from the level records "
                                    "(bench.traversal_bytes)",
                     "bytes_per_level_mode": {("init" if k[0] < 0 else
                                               f"L{k[0]}_{['td_sf', 'td_ss', 'bu'][k[1]]}"):
                                              round(v / len(srcs))
                                              for k, v in sorted(lvl_bytes.items())[:40]},
                     "bytes_per_level_note": "mean over the sources of each (level, mode)'s bytes "
                                             "(sources that skip a level or run it in another mode "
                                             "count zero); the entries sum to "
                                             "algorithmic_bytes_per_launch (first 40 entries)",
                     "teps_equivalent": {"bytes_per_launch": round(float(b_teps.mean())),
                                         "gbs": round(float(b_teps.mean()) / t_mean / 1e9, 2),
                                         "model": "8*m_cc + 20*V_cc + 4*V (every edge of the "
                                                  "reached component charged)"},
                     "measured_hbm_gbs": (None if traffic is None else round(traffic / t_mean / 1e9, 1)),
                     "traffic_source": (f"{PMC_TRAFFIC['profile']} (mean launch "
                                        f"{PMC_TRAFFIC['kernel_us']} us when profiled)"
                                        if traffic is not None else traffic_note),
                     "kernel": "k_persistent" if opts.persistent else "host-driven level loop",
                     "launches_per_bfs": 1 if opts.persistent else None,
                     "avg_launch_us": round(t_mean * 1e6, 2)},
        "cpu_baseline": cpu,
        "cpu_serial": cpu_serial,
        "batch64": b64,
        "multigraph": mg,
        "gteps_min": round(float(gteps.min()), 3),
        "gteps_max": round(float(gteps.max()), 3),
        "graph500_gteps_hmean": round(value / 2, 3),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

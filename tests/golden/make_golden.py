"""Writes tests/golden/rmat10_golden.npz: a seeded RMAT scale-10 ef16 graph
from libxbfs's generator, 8 seeded sources, and for each source the levels
produced by both CPU BFS implementations of the oracle (they must agree) plus
the reference-heuristic tag sequences at two alphas.

The reference ships no golden vectors; these fixtures pin regressions of the
generator, the oracle and the heuristic restatement.  Re-run only on purpose:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from xbfs_amd import Graph  # noqa: E402


def main():
    g = Graph.rmat(10, 16, seed=1010)
    rng = np.random.default_rng(42)
    deg = g.degrees()
    cand = np.flatnonzero(deg > 0)
    srcs = rng.choice(cand, size=8, replace=False).astype(np.int32)
    levels = []
    tags01, tags001 = [], []
    for s in srcs:
        a, _, _ = O.bfs_queue(g.beg_pos, g.adj, int(s))
        b, _ = O.bfs_scan(g.beg_pos, g.adj, int(s))
        assert (a == b).all()
        levels.append(a)
        t1, _, _ = O.reference_tags(g.beg_pos, g.adj, int(s), 0.1)
        t2, _, _ = O.reference_tags(g.beg_pos, g.adj, int(s), 0.01)
        tags01.append(np.pad(t1, (0, 32 - len(t1)), constant_values=-9))
        tags001.append(np.pad(t2, (0, 32 - len(t2)), constant_values=-9))
    np.savez_compressed(os.path.join(HERE, "rmat10_golden.npz"), beg_pos=g.beg_pos, adj=g.adj,
                        sources=srcs, levels=np.stack(levels), tags_a0p1=np.stack(tags01),
                        tags_a0p01=np.stack(tags001))
    print("wrote", g.n, g.m, srcs)


if __name__ == "__main__":
    main()

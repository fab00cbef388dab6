"""Closed-form and small random graphs (numpy CSR) used by the tests."""
import numpy as np


def csr_from_edges(n, edges, symmetric=True):
    """CSR (beg int64[n+1], adj int32[m]) from undirected edge pairs; rows
    sorted, duplicates and self-loops removed."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    if e.size:
        e = e[e[:, 0] != e[:, 1]]
        if symmetric:
            e = np.concatenate([e, e[:, ::-1]])
        key = np.unique(e[:, 0] * n + e[:, 1])
        u, v = key // n, key % n
    else:
        u = v = np.zeros(0, np.int64)
    beg = np.zeros(n + 1, np.int64)
    np.add.at(beg, u + 1, 1)
    beg = np.cumsum(beg)
    return beg, v.astype(np.int32)


def raw_csr(n, pairs):
    """CSR straight from directed pairs, keeping duplicates and self-loops
    (rows sorted): the loader accepts such files as they are."""
    e = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    e = e[np.lexsort((e[:, 1], e[:, 0]))]
    beg = np.zeros(n + 1, np.int64)
    np.add.at(beg, e[:, 0] + 1, 1)
    return np.cumsum(beg), e[:, 1].astype(np.int32)


def loops_and_duplicates():
    """Self-loops, duplicate entries and degree-1 chains around a small core:
    every edge of the path 0..19 twice; 20, 21 with only a self-loop; 22 with
    a self-loop and an edge to 0; the pair 23-24; a star on 25 (leaves 26..60,
    even leaves joined twice); 61-62-63 hanging off leaf 26 (61 and 63 have
    degree 1 after 62 gets a self-loop too); 64 joined to 19 twice and to
    itself."""
    und = []
    for i in range(19):
        und += [(i, i + 1), (i, i + 1)]
    und += [(22, 0), (23, 24)]
    for leaf in range(26, 61):
        und += [(25, leaf)] * (2 if leaf % 2 == 0 else 1)
    und += [(26, 61), (61, 62), (62, 63), (64, 19), (64, 19)]
    pairs = [(u, v) for u, v in und] + [(v, u) for u, v in und]
    pairs += [(20, 20), (21, 21), (22, 22), (62, 62), (64, 64)]
    return raw_csr(65, pairs)


def path(n):
    return csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])


def star(n):
    return csr_from_edges(n, [(0, i) for i in range(1, n)])


def complete(n):
    return csr_from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])


def binary_tree(n):
    return csr_from_edges(n, [(i, (i - 1) // 2) for i in range(1, n)])


def grid(r, c):
    e = []
    for i in range(r):
        for j in range(c):
            v = i * c + j
            if j + 1 < c:
                e.append((v, v + 1))
            if i + 1 < r:
                e.append((v, v + c))
    return csr_from_edges(r * c, e)


def disjoint(*graphs):
    """Disjoint union of CSR graphs (ids shifted)."""
    begs, adjs, off, eoff = [np.zeros(1, np.int64)], [], 0, 0
    for beg, adj in graphs:
        n = beg.shape[0] - 1
        begs.append(beg[1:] + eoff)
        adjs.append(adj.astype(np.int64) + off)
        off += n
        eoff += beg[-1]
    return np.concatenate(begs), np.concatenate(adjs).astype(np.int32)


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    a = rng.random((n, n)) < p
    iu = np.argwhere(np.triu(a, 1))
    return csr_from_edges(n, iu)


def floyd_warshall_levels(beg, adj, src):
    n = beg.shape[0] - 1
    INF = 10 ** 9
    d = np.full((n, n), INF, np.int64)
    np.fill_diagonal(d, 0)
    for u in range(n):
        for j in range(beg[u], beg[u + 1]):
            d[u, adj[j]] = 1
    for k in range(n):
        d = np.minimum(d, d[:, k:k + 1] + d[k:k + 1, :])
    r = d[src]
    return np.where(r >= INF, -1, r).astype(np.int32)

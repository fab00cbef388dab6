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

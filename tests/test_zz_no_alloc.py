"""xbfs_bfs_noalloc, the allocation-free entry point, through the C ABI: on
device buffers the caller allocated, its levels must equal xbfs_bfs's on the
same graph and source bit for bit (and the oracle's), its parents must be
tree edges one level up, and its reach statistics must match.  Two calls in
a row, the second on a smaller graph, check that the shared scratch array
keeps nothing from the previous call that the next one reads.

Device buffers come from the HIP runtime libxbfs.so itself links (loaded by
soname through ctypes), not from torch, whose wheel bundles another runtime."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import oracle as O
from xbfs_amd import (POLICY_FORCE_BU, POLICY_FORCE_TD_SF, POLICY_FORCE_TD_SS, POLICY_HEURISTIC,
                      POLICY_LEARNED, POLICY_SCHEDULE, MODE_BU, MODE_TD_SF, MODE_TD_SS, Context,
                      Graph, default_opts)
from xbfs_amd import _capi
from xbfs_amd import policy as P

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def hip():
    _capi.lib()  # loads the runtime libxbfs.so links against
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipFree.argtypes = [ctypes.c_void_p]
    return H


class DevBuf:
    def __init__(self, H, nbytes):
        self.H = H
        self.p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), max(nbytes, 4)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.H.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.H.hipFree(self.p)
            self.p = ctypes.c_void_p()


def noalloc_bfs(H, g, src, weights=None, want_parent=True, **kw):
    o = default_opts(**kw)
    b = DevBuf(H, g.beg_pos.nbytes)
    a = DevBuf(H, g.adj.nbytes)
    lv = DevBuf(H, 4 * g.n)
    par = DevBuf(H, 4 * g.n) if want_parent else None
    try:
        b.put(g.beg_pos)
        a.put(g.adj)
        st = _capi.Stats()
        w = np.ascontiguousarray(weights, np.float32) if weights is not None else None
        rc = _capi.lib().xbfs_bfs_noalloc(0, b.p, a.p, g.n, g.m, int(src), ctypes.byref(o),
                                          w.ctypes.data if w is not None else None, lv.p,
                                          par.p if par is not None else None, ctypes.byref(st))
        _capi.check(rc, "xbfs_bfs_noalloc")
        levels = lv.get(np.int32, g.n)
        parents = par.get(np.int32, g.n) if par is not None else None
        return levels, parents, {k: getattr(st, k) for k, _ in _capi.Stats._fields_}
    finally:
        for d in (b, a, lv, par):
            if d is not None:
                d.free()


def seeded_sources(g, k, seed=42):
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(g.degrees() > 0)
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


CASES = [dict(policy=POLICY_HEURISTIC), dict(policy=POLICY_FORCE_BU),
         dict(policy=POLICY_FORCE_TD_SS), dict(policy=POLICY_FORCE_TD_SF, bin_thread_max=2,
                                               bin_wave_max=8, bin_block_max=32),
         dict(policy=POLICY_SCHEDULE, schedule=[MODE_TD_SF, MODE_BU, MODE_TD_SS, MODE_BU])]


def check_against_entry_point(H, g, ctx, src, weights=None, **kw):
    lv, par, st = noalloc_bfs(H, g, src, weights=weights, **kw)
    ref, _, _, st_ref = ctx.bfs(int(src), **kw)
    bad = np.flatnonzero(lv != ref)
    assert bad.size == 0, f"src={src} {kw}: {bad.size} levels differ from xbfs_bfs, first " \
                          f"{bad[:5]} got {lv[bad[:5]]} want {ref[bad[:5]]}"
    want, _, _ = O.bfs_queue(g.beg_pos, g.adj, int(src))
    assert (lv == want).all()
    assert O.validate(g.beg_pos, g.adj, int(src), lv, par) == 0
    assert st["v_cc"] == st_ref["v_cc"] and st["m_cc"] == st_ref["m_cc"]
    assert st["max_level"] == st_ref["max_level"]


def test_noalloc_two_calls_second_smaller():
    """RMAT-16 then RMAT-12 (and back), every policy: each call equals xbfs_bfs."""
    H = hip()
    big = Graph.rmat(16, 16, seed=1016)
    small = Graph.rmat(12, 8, seed=1212)
    cb, cs = Context(big), Context(small)
    try:
        for kw in CASES:
            for g, ctx, name in ((big, cb, "rmat16"), (small, cs, "rmat12")):
                for src in seeded_sources(g, 2, seed=len(name)):
                    say("noalloc", name, int(src), kw)
                    check_against_entry_point(H, g, ctx, src, **kw)
    finally:
        cb.close()
        cs.close()


def test_noalloc_learned_weights_and_no_parents():
    """Policy LEARNED with the committed weights (same decisions as a context
    with them loaded) and without a parent array."""
    H = hip()
    w = P.load()
    g = Graph.rmat(16, 16, seed=1016)
    ctx = Context(g)
    ctx.load_policy(w)
    try:
        for src in seeded_sources(g, 3, seed=3):
            say("noalloc learned", int(src))
            lv, par, st = noalloc_bfs(H, g, src, weights=w, want_parent=False,
                                      policy=POLICY_LEARNED)
            assert par is None
            ref, _, _, st_ref = ctx.bfs(int(src), policy=POLICY_LEARNED)
            assert (lv == ref).all()
            assert st["m_cc"] == st_ref["m_cc"]
    finally:
        ctx.close()


def test_noalloc_limit_and_degree0_source():
    """n beyond the maximum is refused before any GPU work; a degree-0
    source reaches only itself."""
    H = hip()
    o = default_opts()
    st = _capi.Stats()
    one = ctypes.c_void_p(1)
    rc = _capi.lib().xbfs_bfs_noalloc(0, one, one, _capi.NOALLOC_MAX_N + 1, 1, 0,
                                      ctypes.byref(o), None, one, None, ctypes.byref(st))
    assert rc == 7  # XBFS_ERR_LIMIT
    beg = np.array([0, 0, 1, 2], np.int64)
    adj = np.array([2, 1], np.int32)
    g = Graph(beg, adj)
    lv, par, st = noalloc_bfs(H, g, 0)
    assert lv.tolist() == [0, -1, -1] and par.tolist() == [0, -1, -1]
    assert st["v_cc"] == 1

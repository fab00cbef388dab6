"""ctypes binding of libxbfs.so (the C ABI declared in include/xbfs.h).

The library is loaded from this package directory only; when it is missing
every call raises, so nothing here can silently fall back to a CPU path.
"""
import atexit
import ctypes
import os
import weakref
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_float, c_int, c_int8,
                    c_int32, c_int64, c_uint64, c_void_p)

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libxbfs.so")

XBFS_OK = 0
ERR_NAMES = {
    1: "XBFS_ERR_ARG", 2: "XBFS_ERR_HIP", 3: "XBFS_ERR_NOMEM", 4: "XBFS_ERR_IO",
    5: "XBFS_ERR_FORMAT", 6: "XBFS_ERR_UNSUPPORTED", 7: "XBFS_ERR_LIMIT",
}

MODE_TD_SF, MODE_TD_SS, MODE_BU = 0, 1, 2
POLICY_HEURISTIC, POLICY_LEARNED = 0, 1
POLICY_FORCE_TD_SF, POLICY_FORCE_TD_SS, POLICY_FORCE_BU, POLICY_SCHEDULE = 2, 3, 4, 5


class XbfsError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {ERR_NAMES.get(code, code)}")


class HostGraph(Structure):
    _fields_ = [("n", c_int64), ("m", c_int64), ("beg_pos", POINTER(c_int64)),
                ("adj", POINTER(c_int32)), ("weight", POINTER(c_int8))]


class DegreeInfo(Structure):
    _fields_ = [("small", c_int64), ("medium", c_int64), ("large", c_int64),
                ("max_degree", c_int64), ("max_vertex", c_int64), ("hub_num", c_int64),
                ("isolated", c_int64)]


class Opts(Structure):
    _fields_ = [("block_dim", c_int32), ("grid_dim", c_int32), ("alpha", c_float),
                ("beta", c_float), ("padding", c_int32), ("policy", c_int32),
                ("persistent", c_int32), ("want_parent", c_int32),
                ("bin_thread_max", c_int32), ("bin_wave_max", c_int32),
                ("schedule", POINTER(c_int8)), ("schedule_len", c_int32),
                ("time_levels", c_int32), ("bin_block_max", c_int32),
                ("max_levels", c_int32), ("wait_limit_us", c_int32)]


class LevelRec(Structure):
    _fields_ = [("level", c_int32), ("mode", c_int32), ("tag", c_int32), ("policy", c_int32),
                ("n_f", c_int64), ("m_f", c_int64), ("n_u", c_int64), ("m_u", c_int64),
                ("ratio", c_float), ("pad", c_float), ("t_s", c_double),
                ("bu_cand", c_int64), ("bu_scan_v", c_int64), ("bu_scan_e", c_int64)]


class Stats(Structure):
    _fields_ = [("t_bfs_s", c_double), ("n_levels", c_int32), ("max_level", c_int32),
                ("v_cc", c_int64), ("m_cc", c_int64), ("td_s", c_double), ("bu_s", c_double),
                ("gen_s", c_double)]


class TuneRec(Structure):
    _fields_ = [("bin_thread_max", c_int32), ("bin_wave_max", c_int32),
                ("bin_block_max", c_int32), ("grid_dim", c_int32), ("t_s", c_double)]


class ProbeRec(Structure):
    _fields_ = [("level", c_int32), ("prev_mode", c_int32), ("mode", c_int32), ("htag", c_int32),
                ("n_f", c_int64), ("m_f", c_int64), ("n_u", c_int64), ("m_u", c_int64),
                ("t_mode", c_double * 3)]


MLP_NPARAM = 1443

# every symbol include/xbfs.h declares, with its ctypes signature
SIGNATURES = {
    "xbfs_strerror": (c_char_p, [c_int]),
    "xbfs_graph_load": (c_int, [c_char_p, c_char_p, c_char_p, POINTER(HostGraph)]),
    "xbfs_graph_write": (c_int, [c_char_p, POINTER(HostGraph)]),
    "xbfs_graph_from_edge_text": (c_int, [c_char_p, c_int, POINTER(HostGraph)]),
    "xbfs_graph_from_tuples": (c_int, [c_char_p, c_int, c_int64, POINTER(HostGraph)]),
    "xbfs_graph_free": (None, [POINTER(HostGraph)]),
    "xbfs_gen_rmat": (c_int, [c_int32, c_int32, c_double, c_double, c_double, c_uint64,
                              c_int32, POINTER(HostGraph)]),
    "xbfs_gen_er": (c_int, [c_int64, c_int64, c_uint64, POINTER(HostGraph)]),
    "xbfs_gen_grid": (c_int, [c_int32, c_int32, POINTER(HostGraph)]),
    "xbfs_degree_stats": (c_int, [POINTER(HostGraph), POINTER(DegreeInfo)]),
    "xbfs_create": (c_int, [c_int, c_void_p, c_void_p, c_int64, c_int64, POINTER(c_void_p)]),
    "xbfs_destroy": (None, [c_void_p]),
    "xbfs_num_vertices": (c_int64, [c_void_p]),
    "xbfs_num_edges": (c_int64, [c_void_p]),
    "xbfs_launch_fallbacks": (c_int64, [c_void_p]),
    "xbfs_opts_default": (None, [POINTER(Opts)]),
    "xbfs_bfs": (c_int, [c_void_p, c_int32, POINTER(Opts), c_void_p, c_void_p,
                         POINTER(LevelRec), c_int32, POINTER(Stats)]),
    "xbfs_exclusive_scan_i32": (c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
    "xbfs_unvisited_list": (c_int, [c_void_p, c_void_p, c_void_p, POINTER(c_int64)]),
    "xbfs_policy_features": (c_int, [c_int32, c_int64, c_int64, c_int64, c_int64, c_int32,
                                     c_void_p]),
    "xbfs_policy_load": (c_int, [c_void_p, c_void_p, c_int32]),
    "xbfs_mlp_forward": (c_int, [c_int, c_void_p, c_void_p, c_int32, c_void_p]),
    "xbfs_mlp_loss_grad": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                   c_void_p, c_void_p]),
    "xbfs_mlp_train": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32,
                               c_float, c_void_p]),
    "xbfs_bfs_probe": (c_int, [c_void_p, c_int32, POINTER(Opts), c_int32, c_int32,
                               POINTER(ProbeRec), c_int32, POINTER(Stats)]),
    "xbfs_bfs_multi": (c_int, [c_void_p, c_void_p, c_int32, POINTER(Opts), c_void_p, c_void_p,
                               POINTER(LevelRec), c_int32, POINTER(Stats)]),
    "xbfs_graph_concat": (c_int, [POINTER(HostGraph), c_int32, POINTER(HostGraph), c_void_p]),
    "xbfs_bfs_multigraph": (c_int, [c_void_p, c_void_p, c_int32, c_void_p, POINTER(Opts),
                                    c_void_p, POINTER(Stats)]),
    "xbfs_bfs_batch64": (c_int, [c_void_p, c_void_p, c_int32, POINTER(Opts), c_void_p,
                                 POINTER(Stats)]),
    "xbfs_autotune": (c_int, [c_void_p, c_void_p, c_int32, c_int32, c_int32, POINTER(Opts),
                              POINTER(TuneRec), c_int32, POINTER(c_int32)]),
    "xbfs_bfs_noalloc": (c_int, [c_int, c_void_p, c_void_p, c_int64, c_int64, c_int32,
                                 POINTER(Opts), c_void_p, c_void_p, c_void_p, POINTER(Stats)]),
}

NOALLOC_MAX_N = 1 << 20

_lib = None


def lib():
    """Loads libxbfs.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C xbfs_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc, what):
    if rc != XBFS_OK:
        raise XbfsError(rc, what)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Graph:
    """A host CSR graph: beg_pos int64[n+1], adj int32[m] (numpy arrays)."""

    def __init__(self, beg_pos, adj):
        self.beg_pos = np.ascontiguousarray(beg_pos, dtype=np.int64)
        self.adj = np.ascontiguousarray(adj, dtype=np.int32)
        self.n = int(self.beg_pos.shape[0] - 1)
        self.m = int(self.adj.shape[0])

    @staticmethod
    def _from_host(hg):
        n, m = hg.n, hg.m
        beg = np.ctypeslib.as_array(hg.beg_pos, shape=(n + 1,)).copy()
        adj = np.ctypeslib.as_array(hg.adj, shape=(m,)).copy() if m > 0 else np.zeros(0, np.int32)
        lib().xbfs_graph_free(byref(hg))
        return Graph(beg, adj)

    @staticmethod
    def rmat(scale, ef, a=0.57, b=0.19, c=0.19, seed=1, relabel=True):
        hg = HostGraph()
        check(lib().xbfs_gen_rmat(scale, ef, a, b, c, seed, int(relabel), byref(hg)), "gen_rmat")
        return Graph._from_host(hg)

    @staticmethod
    def er(n, m, seed=7):
        hg = HostGraph()
        check(lib().xbfs_gen_er(n, m, seed, byref(hg)), "gen_er")
        return Graph._from_host(hg)

    @staticmethod
    def grid(rows, cols):
        hg = HostGraph()
        check(lib().xbfs_gen_grid(rows, cols, byref(hg)), "gen_grid")
        return Graph._from_host(hg)

    @staticmethod
    def load(beg_file, csr_file, weight_file=None):
        hg = HostGraph()
        check(lib().xbfs_graph_load(beg_file.encode(), csr_file.encode(),
                                    weight_file.encode() if weight_file else None, byref(hg)),
              "graph_load")
        return Graph._from_host(hg)

    @staticmethod
    def from_tuples(path, reverse=True, skip_lines=0):
        """The reference converter's semantics (xbfs_graph_from_tuples).
        Returns (graph, weights int8[m])."""
        hg = HostGraph()
        check(lib().xbfs_graph_from_tuples(path.encode(), int(reverse), int(skip_lines), byref(hg)),
              "graph_from_tuples")
        m = hg.m
        w = np.ctypeslib.as_array(hg.weight, shape=(m,)).copy() if m > 0 else np.zeros(0, np.int8)
        return Graph._from_host(hg), w

    @staticmethod
    def from_edge_text(path, reverse=True):
        hg = HostGraph()
        check(lib().xbfs_graph_from_edge_text(path.encode(), int(reverse), byref(hg)),
              "graph_from_edge_text")
        return Graph._from_host(hg)

    def _host(self):
        hg = HostGraph()
        hg.n, hg.m = self.n, self.m
        hg.beg_pos = self.beg_pos.ctypes.data_as(POINTER(c_int64))
        hg.adj = self.adj.ctypes.data_as(POINTER(c_int32))
        hg.weight = POINTER(c_int8)()
        return hg

    def write(self, prefix):
        hg = self._host()
        check(lib().xbfs_graph_write(prefix.encode(), byref(hg)), "graph_write")

    def degree_stats(self):
        hg = self._host()
        di = DegreeInfo()
        check(lib().xbfs_degree_stats(byref(hg), byref(di)), "degree_stats")
        return {k: getattr(di, k) for k, _ in DegreeInfo._fields_}

    def degrees(self):
        return np.diff(self.beg_pos)

    @staticmethod
    def concat(graphs):
        """Disjoint union (xbfs_graph_concat).  Returns (graph, vert_off) where
        graph i's vertex v is vert_off[i] + v."""
        k = len(graphs)
        hosts = (HostGraph * k)()
        for i, g in enumerate(graphs):
            hosts[i] = g._host()
        off = np.empty(k + 1, np.int64)
        out = HostGraph()
        check(lib().xbfs_graph_concat(hosts, k, byref(out), _ptr(off)), "graph_concat")
        return Graph._from_host(out), off


def default_opts(**kw):
    o = Opts()
    lib().xbfs_opts_default(byref(o))
    sched = kw.pop("schedule", None)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    if sched is not None:
        arr = (c_int8 * len(sched))(*sched)
        o.schedule = ctypes.cast(arr, POINTER(c_int8))
        o.schedule_len = len(sched)
        o._sched_keepalive = arr
    return o


# Contexts still open when the interpreter exits are destroyed by an atexit
# handler, which runs before module teardown and before the C runtime's own
# exit handlers (a profiler's among them): device memory, streams and events
# are then released while the HIP runtime is still fully up, never from a
# finalizer during interpreter teardown.
_live = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    for c in list(_live):
        c.close()


class Context:
    """Device context: the CSR uploaded to one GPU plus all traversal state."""

    def __init__(self, graph, device=0):
        self.graph = graph
        self.h = None
        h = c_void_p()
        check(lib().xbfs_create(device, _ptr(graph.beg_pos), _ptr(graph.adj), graph.n, graph.m,
                                byref(h)), "xbfs_create")
        self.h = h
        self.n, self.m = graph.n, graph.m
        _live.add(self)

    def launch_fallbacks(self):
        """Persistent traversals relaunched cooperatively after a plain
        launch's residency check failed (xbfs_launch_fallbacks)."""
        return int(lib().xbfs_launch_fallbacks(self.h))

    def close(self):
        if self.h:
            lib().xbfs_destroy(self.h)
            self.h = None
        _live.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bfs(self, src, opts=None, want_levels=True, want_parent=False, rec_cap=0, **kw):
        """One BFS from src.  Returns (levels, parents, records, stats)."""
        o = opts if opts is not None else default_opts(**kw)
        lv = np.empty(self.n, np.int32) if want_levels else None
        par = np.empty(self.n, np.int32) if want_parent else None
        recs = (LevelRec * rec_cap)() if rec_cap > 0 else None
        st = Stats()
        rc = lib().xbfs_bfs(self.h, int(src), byref(o), _ptr(lv) if lv is not None else None,
                            _ptr(par) if par is not None else None, recs, rec_cap, byref(st))
        check(rc, "xbfs_bfs")
        out_recs = []
        if recs is not None:
            for i in range(min(rec_cap, st.n_levels)):
                r = recs[i]
                out_recs.append({k: getattr(r, k) for k, _ in LevelRec._fields_})
        stats = {k: getattr(st, k) for k, _ in Stats._fields_}
        return lv, par, out_recs, stats

    def bfs_multi(self, srcs, opts=None, want_levels=True, want_parent=False, rec_cap=0, **kw):
        """Multi-source BFS (distance to the nearest seed).  Returns
        (levels, parents, records, stats)."""
        o = opts if opts is not None else default_opts(**kw)
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        lv = np.empty(self.n, np.int32) if want_levels else None
        par = np.empty(self.n, np.int32) if want_parent else None
        recs = (LevelRec * rec_cap)() if rec_cap > 0 else None
        st = Stats()
        check(lib().xbfs_bfs_multi(self.h, _ptr(s), s.shape[0], byref(o),
                                   _ptr(lv) if lv is not None else None,
                                   _ptr(par) if par is not None else None, recs, rec_cap,
                                   byref(st)), "xbfs_bfs_multi")
        out_recs = []
        if recs is not None:
            for i in range(min(rec_cap, st.n_levels)):
                r = recs[i]
                out_recs.append({k: getattr(r, k) for k, _ in LevelRec._fields_})
        return lv, par, out_recs, {k: getattr(st, k) for k, _ in Stats._fields_}

    def bfs_multigraph(self, vert_off, local_src, opts=None, want_levels=True, **kw):
        """One BFS per packed graph (xbfs_bfs_multigraph).  Returns (levels, stats);
        levels[vert_off[i]:vert_off[i+1]] is graph i's level array."""
        o = opts if opts is not None else default_opts(**kw)
        off = np.ascontiguousarray(vert_off, dtype=np.int64)
        ls = np.ascontiguousarray(local_src, dtype=np.int32)
        lv = np.empty(self.n, np.int32) if want_levels else None
        st = Stats()
        check(lib().xbfs_bfs_multigraph(self.h, _ptr(off), ls.shape[0], _ptr(ls), byref(o),
                                        _ptr(lv) if lv is not None else None, byref(st)),
              "xbfs_bfs_multigraph")
        return lv, {k: getattr(st, k) for k, _ in Stats._fields_}

    def bfs_batch64(self, srcs, opts=None, want_levels=True, **kw):
        """Up to 64 single-source BFSs in one bit-parallel traversal.  Returns
        (levels [nsrc, n] or None, stats)."""
        o = opts if opts is not None else default_opts(**kw)
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        lv = np.empty((s.shape[0], self.n), np.int32) if want_levels else None
        st = Stats()
        check(lib().xbfs_bfs_batch64(self.h, _ptr(s), s.shape[0], byref(o),
                                     _ptr(lv) if lv is not None else None, byref(st)),
              "xbfs_bfs_batch64")
        return lv, {k: getattr(st, k) for k, _ in Stats._fields_}

    def autotune(self, srcs, multi=False, reps=3, opts=None, cap=256, **kw):
        """Coordinate-descent search of bin thresholds and grid size
        (xbfs_autotune).  Returns (best opts, list of trial dicts)."""
        o = opts if opts is not None else default_opts(**kw)
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        trials = (TuneRec * cap)()
        nt = c_int32()
        check(lib().xbfs_autotune(self.h, _ptr(s), s.shape[0], int(bool(multi)), int(reps),
                                  byref(o), trials, cap, byref(nt)), "xbfs_autotune")
        out = [{k: getattr(trials[i], k) for k, _ in TuneRec._fields_} for i in range(nt.value)]
        return o, out

    def load_policy(self, weights):
        w = np.ascontiguousarray(weights, dtype=np.float32)
        check(lib().xbfs_policy_load(self.h, _ptr(w), w.shape[0]), "policy_load")

    def probe(self, src, follow_best=False, reps=3, rec_cap=256, **kw):
        """Per-level timing of all three modes (see xbfs_bfs_probe)."""
        o = default_opts(**kw)
        recs = (ProbeRec * rec_cap)()
        st = Stats()
        check(lib().xbfs_bfs_probe(self.h, int(src), byref(o), int(follow_best), int(reps), recs,
                                   rec_cap, byref(st)), "xbfs_bfs_probe")
        out = []
        for i in range(min(rec_cap, st.n_levels)):
            r = recs[i]
            d = {k: getattr(r, k) for k, _ in ProbeRec._fields_ if k != "t_mode"}
            d["t_mode"] = [r.t_mode[j] for j in range(3)]
            out.append(d)
        return out, {k: getattr(st, k) for k, _ in Stats._fields_}

    def exclusive_scan(self, x):
        x = np.ascontiguousarray(x, dtype=np.int32)
        out = np.empty(x.shape[0], np.int64)
        check(lib().xbfs_exclusive_scan_i32(self.h, _ptr(x), _ptr(out), x.shape[0]), "scan")
        return out

    def unvisited_list(self, level):
        level = np.ascontiguousarray(level, dtype=np.int32)
        out = np.empty(self.n, np.int32)
        cnt = c_int64()
        check(lib().xbfs_unvisited_list(self.h, _ptr(level), _ptr(out), byref(cnt)),
              "unvisited_list")
        return out[: cnt.value].copy()


def policy_features(level, n_f, m_f, n_u, m_u, prev_mode):
    x = np.empty(8, np.float32)
    check(lib().xbfs_policy_features(int(level), int(n_f), int(m_f), int(n_u), int(m_u),
                                     int(prev_mode), _ptr(x)), "policy_features")
    return x


def mlp_forward(params, x, device=0):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty((x.shape[0], 3), np.float32)
    check(lib().xbfs_mlp_forward(device, _ptr(params), _ptr(x), x.shape[0], _ptr(out)),
          "mlp_forward")
    return out


def mlp_loss_grad(params, x, y, mask, device=0):
    params = np.ascontiguousarray(params, np.float32)
    x, y, mask = (np.ascontiguousarray(a, np.float32) for a in (x, y, mask))
    g = np.empty(MLP_NPARAM, np.float32)
    loss = np.empty(1, np.float32)
    check(lib().xbfs_mlp_loss_grad(device, _ptr(params), _ptr(x), _ptr(y), _ptr(mask),
                                   x.shape[0], _ptr(g), _ptr(loss)), "mlp_loss_grad")
    return float(loss[0]), g


def mlp_train(params, x, y, mask, steps, lr, device=0):
    p = np.array(params, dtype=np.float32, copy=True)
    x, y, mask = (np.ascontiguousarray(a, np.float32) for a in (x, y, mask))
    curve = np.empty(steps, np.float32)
    check(lib().xbfs_mlp_train(device, _ptr(p), _ptr(x), _ptr(y), _ptr(mask), x.shape[0],
                               int(steps), float(lr), _ptr(curve)), "mlp_train")
    return p, curve

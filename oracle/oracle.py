"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as the thing measured or shipped.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LOADER = os.path.join(HERE, "_ref", "ref_loader")
_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-B", "all"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-C", HERE, "-B", "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", HERE, "all"], check=True)
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_bfs_scan.restype = c_int32
        L.oracle_bfs_scan.argtypes = [c_int64, c_void_p, c_void_p, c_int32, c_void_p]
        L.oracle_bfs_queue.restype = c_int32
        L.oracle_bfs_queue.argtypes = [c_int64, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]
        L.oracle_validate.restype = c_int64
        L.oracle_validate.argtypes = [c_int64, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]
        L.oracle_level_counts.restype = c_int32
        L.oracle_level_counts.argtypes = [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int32]
        L.oracle_heuristic_tags.restype = c_int32
        L.oracle_heuristic_tags.argtypes = [c_void_p, c_int32, c_int64, c_float, c_void_p, c_void_p]
        L.oracle_omp_td_bfs.restype = c_double
        L.oracle_omp_td_bfs.argtypes = [c_int64, c_void_p, c_void_p, c_int32, c_void_p, c_int32]
        L.oracle_exclusive_scan.restype = None
        L.oracle_exclusive_scan.argtypes = [c_int64, c_void_p, c_void_p]
        L.oracle_bfs_queue_rows.restype = None
        L.oracle_bfs_queue_rows.argtypes = [c_int64, c_void_p, c_void_p, c_void_p, c_int32,
                                            c_void_p]
        L.oracle_bfs_multi.restype = c_int32
        L.oracle_bfs_multi.argtypes = [c_int64, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                                       c_void_p]
        L.oracle_validate_multi.restype = c_int64
        L.oracle_validate_multi.argtypes = [c_int64, c_void_p, c_void_p, c_void_p, c_int32,
                                            c_void_p, c_void_p]
        L.oracle_max_threads.restype = c_int32
        L.oracle_mlp_nparam.restype = c_int32
        L.oracle_mlp_forward.restype = None
        L.oracle_mlp_forward.argtypes = [c_void_p, c_void_p, c_int32, c_void_p]
        L.oracle_mlp_loss.restype = c_double
        L.oracle_mlp_loss.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32]
        L.oracle_mlp_loss_grad.restype = c_double
        L.oracle_mlp_loss_grad.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                           c_void_p]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _csr(beg, adj):
    return (np.ascontiguousarray(beg, np.int64), np.ascontiguousarray(adj, np.int32))


def bfs_scan(beg, adj, src):
    """Literal restatement of the reference serial BFS (bfs_gpu_update.cu:104-127).
    Returns (levels, k)."""
    beg, adj = _csr(beg, adj)
    n = beg.shape[0] - 1
    lv = np.empty(n, np.int32)
    k = lib().oracle_bfs_scan(n, _p(beg), _p(adj), int(src), _p(lv))
    return lv, k


def bfs_queue(beg, adj, src, parents=False):
    """Serial FIFO BFS.  Returns (levels, parents or None, nlevels)."""
    beg, adj = _csr(beg, adj)
    n = beg.shape[0] - 1
    lv = np.empty(n, np.int32)
    par = np.empty(n, np.int32) if parents else None
    k = lib().oracle_bfs_queue(n, _p(beg), _p(adj), int(src), _p(lv), _p(par))
    return lv, par, k


def bfs_queue_rows(beg, adj, srcs):
    """oracle_bfs_queue from each source (rows computed in parallel threads).
    Returns int32 levels [len(srcs), n]."""
    beg, adj = _csr(beg, adj)
    n = beg.shape[0] - 1
    s = np.ascontiguousarray(srcs, np.int32)
    out = np.empty((s.shape[0], n), np.int32)
    lib().oracle_bfs_queue_rows(n, _p(beg), _p(adj), _p(s), s.shape[0], _p(out))
    return out


def bfs_multi(beg, adj, srcs, parents=False):
    """Serial FIFO BFS from every seed at once.  Returns (levels, parents or None, nlevels)."""
    beg, adj = _csr(beg, adj)
    n = beg.shape[0] - 1
    s = np.ascontiguousarray(srcs, np.int32)
    lv = np.empty(n, np.int32)
    par = np.empty(n, np.int32) if parents else None
    k = lib().oracle_bfs_multi(n, _p(beg), _p(adj), _p(s), s.shape[0], _p(lv), _p(par))
    return lv, par, k


def validate_multi(beg, adj, srcs, level, parent=None):
    """Number of violated checks of a multi-source result (0 = valid)."""
    beg, adj = _csr(beg, adj)
    n = beg.shape[0] - 1
    s = np.ascontiguousarray(srcs, np.int32)
    level = np.ascontiguousarray(level, np.int32)
    parent = None if parent is None else np.ascontiguousarray(parent, np.int32)
    return int(lib().oracle_validate_multi(n, _p(beg), _p(adj), _p(s), s.shape[0], _p(level),
                                           _p(parent)))


def validate(beg, adj, src, level, parent=None):
    """Number of violated Graph500 checks (0 = valid BFS result)."""
    beg, adj = _csr(beg, adj)
    n = beg.shape[0] - 1
    level = np.ascontiguousarray(level, np.int32)
    parent = None if parent is None else np.ascontiguousarray(parent, np.int32)
    return int(lib().oracle_validate(n, _p(beg), _p(adj), int(src), _p(level), _p(parent)))


def level_counts(beg, level, cap=None):
    """(count[L] = sum deg over level==L, nverts[L]) for L < cap."""
    beg = np.ascontiguousarray(beg, np.int64)
    level = np.ascontiguousarray(level, np.int32)
    n = beg.shape[0] - 1
    if cap is None:
        cap = int(level.max()) + 1 if n else 0
    cnt = np.zeros(max(cap, 1), np.int64)
    nv = np.zeros(max(cap, 1), np.int64)
    nl = lib().oracle_level_counts(n, _p(beg), _p(level), _p(cnt), _p(nv), cap)
    return cnt[:cap], nv[:cap], nl


def heuristic_tags(count, E, alpha):
    """Reference tag sequence (bfs_adaptiveFQ_wb_async.cuh:1047-1078) for the
    frontier-edge counts count[0..k-1].  Returns (tags, subs, TD_adaptive)."""
    count = np.ascontiguousarray(count, np.int64)
    k = count.shape[0]
    tags = np.zeros(max(k, 1), np.int32)
    subs = np.zeros(max(k, 1), np.int32)
    tda = lib().oracle_heuristic_tags(_p(count), k, int(E), float(alpha), _p(tags), _p(subs))
    return tags[:k], subs[:k], tda


def reference_tags(beg, adj, src, alpha):
    """Tags the reference emits from src: its loop runs levels 0..maxl-1."""
    lv, _, nl = bfs_queue(beg, adj, src)
    cnt, _, _ = level_counts(beg, lv, cap=nl)
    E = int(np.asarray(beg)[-1])
    k = max(nl - 1, 0)
    return heuristic_tags(cnt[:k], E, alpha)


def omp_td_bfs(beg, adj, src, nthreads=0):
    """OpenMP top-down BFS (CPU baseline).  Returns (levels, seconds)."""
    beg, adj = _csr(beg, adj)
    n = beg.shape[0] - 1
    lv = np.empty(n, np.int32)
    t = lib().oracle_omp_td_bfs(n, _p(beg), _p(adj), int(src), _p(lv), int(nthreads))
    return lv, t


def max_threads():
    return int(lib().oracle_max_threads())


def exclusive_scan(x):
    x = np.ascontiguousarray(x, np.int32)
    out = np.empty(x.shape[0], np.int64)
    lib().oracle_exclusive_scan(x.shape[0], _p(x), _p(out))
    return out


MLP_NPARAM = 1443


def mlp_forward(params, x):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    B = x.shape[0]
    out = np.empty((B, 3), np.float64)
    lib().oracle_mlp_forward(_p(params), _p(x), B, _p(out))
    return out


def mlp_loss(params64, x, t, mask):
    params64 = np.ascontiguousarray(params64, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    mask = np.ascontiguousarray(mask, np.float32)
    return lib().oracle_mlp_loss(_p(params64), _p(x), _p(t), _p(mask), x.shape[0])


def mlp_loss_grad(params, x, t, mask):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    mask = np.ascontiguousarray(mask, np.float32)
    g = np.empty(MLP_NPARAM, np.float64)
    loss = lib().oracle_mlp_loss_grad(_p(params), _p(x), _p(t), _p(mask), x.shape[0], _p(g))
    return loss, g


def ref_loader(beg_file, csr_file, weight_file="/nonexistent"):
    """Runs the reference's own loader (built into oracle/_ref/ from the
    reference tree) and returns (n, m, max_degree, printed_vertex, h_beg, h_csr),
    or None when it was not built."""
    if not os.path.exists(REF_LOADER):
        return None
    out = subprocess.run([REF_LOADER, beg_file, csr_file, weight_file], check=True,
                         capture_output=True, text=True).stdout.split()
    return tuple(int(v) for v in out)

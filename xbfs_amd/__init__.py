"""xbfs_amd: MI355X-native adaptive direction-optimizing BFS (XBFS).

The product is libxbfs.so (HIP kernels for gfx950 behind the C ABI in
include/xbfs.h) and the ./bfs driver.  This package is the Python host
binding used by the tests and the benchmark.
"""
from ._capi import (Context, Graph, XbfsError, default_opts, lib, LIB_PATH,  # noqa: F401
                    MODE_TD_SF, MODE_TD_SS, MODE_BU, POLICY_HEURISTIC, POLICY_LEARNED,
                    POLICY_FORCE_TD_SF, POLICY_FORCE_TD_SS, POLICY_FORCE_BU, POLICY_SCHEDULE)

// Device context behind the opaque xbfs_ctx handle.
#ifndef XBFS_CTX_H
#define XBFS_CTX_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>
#include "xbfs_device.h"

#define XB_DEFAULT_T1 32   /* thread bin: 0 < deg <= 32    */
#define XB_DEFAULT_T2 1024 /* wave bin: deg <= 1024, block bin above */

struct xbfs_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0, m = 0;       // vertices, directed edges
    int64_t nwords = 0;         // ceil(n / 64) bitmap words
    int64_t ntiles = 0;         // ceil(nwords / 64) bottom-up tiles
    int64_t n_nz = 0;           // vertices with degree > 0
    int64_t *d_beg = nullptr;   // int64[n+1]
    int32_t *d_adj = nullptr;   // int32[m]
    uint64_t *d_nz = nullptr;   // bit v: deg(v) > 0
    int32_t *d_level = nullptr; // int32[n]
    int32_t *d_parent = nullptr;
    int32_t *d_qbuf = nullptr;  // 2 queue buffers x 3 degree bins x n
    int32_t *d_list = nullptr;  // bottom-up unvisited list, int32[n]
    uint64_t *d_bits = nullptr; // frontier, next, unvisited bitmaps (3 x nwords)
    int32_t *d_tile = nullptr;  // tile counts + offsets (2 x ntiles)
    XbDevCtl *d_ctl = nullptr;
    XbDevCtl *h_ctl = nullptr;  // pinned mirror
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::vector<hipEvent_t> ev_pool;
    int32_t *q(int b, int k) { return d_qbuf + ((int64_t)b * 3 + k) * n; }
};

int xb_ctx_free_device(xbfs_ctx *c);
#endif

// Device context behind the opaque xbfs_ctx handle.
#ifndef XBFS_CTX_H
#define XBFS_CTX_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>
#include "xbfs_engine.h"

#define XB_DEFAULT_T1 32     /* thread bin: 0 < deg <= 32                     */
#define XB_DEFAULT_T2 1024   /* wave bin: deg <= 1024                          */
#define XB_DEFAULT_T3 16384  /* block bin: deg <= 16384, grid bin above        */
#define XB_REC_CAP 4096      /* device-side level records kept per traversal   */

struct xbfs_ctx {
    int device = 0;
    int num_cu = 0;
    int coop_blocks = 0;        // co-resident blocks of the persistent kernel
    double wall_hz = 1e8;       // device wall clock rate
    hipStream_t stream = nullptr;
    int64_t n = 0, m = 0;       // vertices, directed edges
    int64_t nwords = 0;         // ceil(n / 64) bitmap words
    int64_t ntiles_bu = 0;      // bottom-up tiles (XB_BU_TILE_WORDS words each)
    int64_t fstride = 0;        // words per frontier ring (tile multiple)
    int64_t n_nz = 0;           // vertices with degree > 0
    double skew = 0;            // largest degree / mean degree of the non-isolated vertices
    int64_t *d_beg = nullptr;   // int64[n+1]
    int32_t *d_adj = nullptr;   // int32[m]
    uint64_t *d_nz = nullptr;   // bit v: deg(v) > 0
    // The component of the largest-degree vertex (the giant component of a
    // skewed graph), found at create by one traversal from that vertex: a
    // traversal from a member ends once n_comp vertices have levels, and its
    // bottom-up candidates leave out the vertices no path reaches.  NULL when
    // the component holds every non-isolated vertex (nothing to leave out).
    uint64_t *d_comp = nullptr;
    std::vector<uint64_t> h_comp; // host copy (membership of a source)
    int64_t n_comp = 0;
    int2 *d_hd = nullptr;       // {largest-degree neighbour, degree} per vertex
    int32_t *d_adj_bu = nullptr; // bottom-up copy of d_adj, rows by neighbour degree (or NULL)
    int32_t hub_first = 0;      // d_adj_bu is built (each row starts with its hub)
    int launch_plain = 1;       // persistent kernel: plain launch with a residency check
    double res_us = 20000.0;    // bound of that check (first barrier wait), microseconds
    int nt_bu = -1;             // bottom-up non-temporal loads: -1 by size (XB_NT_HUB_BYTES), 0 off, 1 on
    int64_t plain_fallbacks = 0; // plain launches that were relaunched cooperatively
    uint64_t *d_vis = nullptr;  // visited bitmap
    uint64_t *d_fbm = nullptr;  // 4 frontier bitmaps (ring by level & 3)
    int32_t *d_level = nullptr; // int32[n]
    int32_t *d_parent = nullptr;
    int32_t *d_qbuf = nullptr;  // 3 queue buffers x 4 degree bins x n
    XbCtl *d_ctl = nullptr;
    XbLoopShared *d_lsg = nullptr; // persistent kernel: loop state for parked blocks
    long long *d_part = nullptr; // XB_NFIELDS x XB_PMAX partial statistics
    long long *s_part = nullptr; // probe snapshot of d_part
    XbCtl *h_ctl = nullptr;     // pinned mirror
    XbDevRec *d_recs = nullptr; // XB_REC_CAP records
    int8_t *d_sched = nullptr;  // schedule copy (XB_REC_CAP entries)
    float *d_mlp = nullptr;     // learned policy weights
    int mlp_loaded = 0;
    std::vector<float> h_mlp;   // host copy (the host-driven loop decides on the CPU)
    // ordered unvisited list / generic scan scratch
    int32_t *d_list = nullptr;
    uint64_t *d_ubits = nullptr;
    int32_t *d_tile = nullptr;
    int32_t *d_list_cnt = nullptr;
    int64_t ntiles_list = 0;
    // snapshot of the traversal state for per-level probes (allocated lazily)
    int32_t *s_level = nullptr;
    uint64_t *s_vis = nullptr, *s_fbm = nullptr;
    int32_t *s_qbuf = nullptr;
    XbCtl *s_ctl = nullptr;
    // multi-source seeds (grown on demand)
    int32_t *d_srcs = nullptr;
    int64_t srcs_cap = 0;
    // 64-source batch engine (allocated on the first xbfs_bfs_batch64): the
    // graph relabelled by degree (rank order) and the batch state
    int b64_ready = 0;
    int32_t b64_H = 0;             // history slots
    int32_t b64_grid = 0;          // co-resident blocks of k_b64_persistent
    int64_t b64_nseg = 0;          // 32-edge segments of the relabelled rows
    int32_t *d_b64_rank = nullptr; // original id -> rank
    int32_t *d_b64_ord = nullptr;  // rank -> original id
    int32_t *d_b64_rdeg = nullptr; // degree by rank
    int64_t *d_b64_rbeg = nullptr; // int64[n+1]
    int32_t *d_b64_radj = nullptr; // int32[m]
    int32_t *d_b64_segr = nullptr; // int32[nseg]: rank owning segment i
    int64_t *d_b64_sege = nullptr; // int64[nseg+1]: first edge of segment i
    uint64_t *d_b64_F = nullptr;   // H x uint64[n]
    uint64_t *d_b64_vis = nullptr; // uint64[n]
    uint64_t *d_b64_acc = nullptr; // uint64[n]
    uint64_t *d_b64_rec = nullptr; // 8 x uint64[n]
    int32_t *d_b64_lv = nullptr;   // 64 x int32[n]
    long long *d_b64_out = nullptr;// 8 results
    int32_t *d_b64_srcs = nullptr; // 64 sources
    hipEvent_t ev_b64[2] = {nullptr, nullptr};
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::vector<hipEvent_t> ev_pool;
};

int xb_ctx_free_device(xbfs_ctx *c);
#endif

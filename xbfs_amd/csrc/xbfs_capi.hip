// C ABI of libxbfs: device context, the two drivers of the adaptive level loop
// (BFS_ControlFlow, bfs_adaptiveFQ_wb_async.cuh:934-1258, with the decision
// made from device-side frontier statistics instead of O(V) host scans), and
// the test hooks for the scan / list kernels.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "xbfs.h"
#include "xbfs_policy.h"
#include "xbfs_host_internal.h"
#include "xbfs_ctx.h"
#include "xbfs_kernels.hip"
#include "xbfs_engine.hip"
#include "xbfs_mlp.hip"
#include "xbfs_batch.hip"

// dynamic LDS: one XbWaveLds per wave of the block
#define XB_LDS_PHASE (sizeof(XbWaveLds) * (XB_PBLOCK / 64))
#define XB_LDS_PERSIST (sizeof(XbWaveLds) * (XB_PERSIST_BLOCK / 64))

#define XB_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return XBFS_ERR_HIP;                                     \
    } while (0)
#define XB_LAUNCHED() XB_HIP(hipGetLastError())

namespace {

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

int grid_for(int64_t items, int block, int cap)
{
    int64_t g = ceil_div(items, block);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

bool pow2_block(int b) { return b >= 64 && b <= 1024 && (b & (b - 1)) == 0; }

} // namespace

// Exclusive scan of n int32 into int64 on the device, one launch of
// k_scan_lookback.  scratch: ceil(n / XB_SCAN_CHUNK) status words, then one
// word for the tile ticket and the error flag, zeroed here.  *err_word points
// at the error flag (nonzero: a look-back wait ran out).
static hipError_t xb_scan_i32(hipStream_t s, const int32_t *in, int64_t n, long long *out,
                              unsigned long long *scratch, int32_t **err_word)
{
    const int64_t T = (n + XB_SCAN_CHUNK - 1) / XB_SCAN_CHUNK;
    hipError_t e = hipMemsetAsync(scratch, 0, sizeof(unsigned long long) * (size_t)(T + 1), s);
    if (e != hipSuccess) return e;
    int32_t *tk = reinterpret_cast<int32_t *>(scratch + T);
    *err_word = tk + 1;
    hipLaunchKernelGGL(k_scan_lookback, dim3((unsigned)T), dim3(256), 0, s, in, n, scratch, tk, tk + 1,
                       out);
    return hipGetLastError();
}

int xb_ctx_free_device(xbfs_ctx *c)
{
    (void)hipSetDevice(c->device);
    void *bufs[] = {c->d_beg,   c->d_adj,  c->d_adj_bu, c->d_nz,  c->d_comp, c->d_hd,   c->d_vis,
                    c->d_fbm,   c->d_level, c->d_parent, c->d_qbuf, c->d_ctl,  c->d_lsg,
                    c->d_recs,  c->d_sched, c->d_mlp,
                    c->d_list,  c->d_ubits, c->d_tile, c->d_list_cnt, c->s_level,
                    c->s_vis,   c->s_fbm,   c->s_qbuf, c->s_ctl, c->d_part, c->s_part,
                    c->d_srcs,  c->d_b64_rank, c->d_b64_ord, c->d_b64_rdeg, c->d_b64_rbeg,
                    c->d_b64_radj, c->d_b64_segr, c->d_b64_sege, c->d_b64_F, c->d_b64_vis,
                    c->d_b64_acc, c->d_b64_rec,
                    c->d_b64_lv, c->d_b64_out, c->d_b64_srcs};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
    for (hipEvent_t e : c->ev_b64)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    return XBFS_OK;
}

extern "C" {

// Bound of one grid-barrier wait in the persistent kernel, in wall-clock
// ticks: opts.wait_limit_us, or 4 s.
// Bound of the first barrier's wait under a plain launch (see bfs_persistent).
static unsigned long long res_ticks(const xbfs_ctx *c)
{
    const double t = c->res_us * 1e-6 * c->wall_hz;
    return t < 1.0 ? 1ull : (unsigned long long)t;
}

static unsigned long long wait_ticks(const xbfs_ctx *c, const xbfs_opts &o)
{
    const double us = o.wait_limit_us > 0 ? (double)o.wait_limit_us : 4.0e6;
    const double t = us * 1e-6 * c->wall_hz;
    return t < 1.0 ? 1ull : (unsigned long long)t;
}

void xbfs_opts_default(xbfs_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->block_dim = XB_PBLOCK;
    o->grid_dim = 0;
    o->alpha = 0.1f; // run.sh:36
    o->beta = 0.1f;  // bfs_adaptiveFQ_wb_async.cuh:1018 (declared, unused)
    o->padding = 1;
    o->policy = XBFS_POLICY_HEURISTIC;
    o->persistent = 1;
    o->want_parent = 0;
    o->bin_thread_max = XB_DEFAULT_T1;
    o->bin_wave_max = XB_DEFAULT_T2;
    o->schedule = nullptr;
    o->schedule_len = 0;
    o->time_levels = 0;
    o->bin_block_max = XB_DEFAULT_T3;
    o->max_levels = 0;
    o->wait_limit_us = 0;
}

int64_t xbfs_num_vertices(const xbfs_ctx *c) { return c ? c->n : -1; }
int64_t xbfs_launch_fallbacks(const xbfs_ctx *c) { return c ? c->plain_fallbacks : -1; }
int64_t xbfs_num_edges(const xbfs_ctx *c) { return c ? c->m : -1; }

// Builds the bottom-up copy of the adjacency, c->d_adj_bu: every row ordered
// by neighbour degree, descending, keeping the CSR order among equal degrees
// (stable radix sort on the device).  A row then starts with its hub, so a
// bottom-up scan that probed the hub first skips it, and the other
// high-degree neighbours (the likeliest frontier members) come early.  The
// top-down phases keep the uploaded order: there, a degree-first order sends
// many lanes to the same few hubs at once.  Levels do not depend on the
// order; parents stay tree edges.  Returns false (no copy) when the sizes
// exceed the sort's 32-bit counts or the memory is not available.
static bool sort_rows_by_degree(xbfs_ctx *c, int64_t max_deg)
{
    if (c->m < 2 || c->m > INT32_MAX || c->n > INT32_MAX) return false;
    int end_bit = 1;
    while (end_bit < 32 && (max_deg >> end_bit) != 0) end_bit++;
    uint32_t *kin = nullptr, *kout = nullptr;
    int32_t *adj2 = nullptr;
    void *tmp = nullptr;
    size_t tb = 0;
    const size_t m = (size_t)c->m;
    bool ok = rocprim::segmented_radix_sort_pairs_desc(nullptr, tb, kin, kout, c->d_adj, adj2,
                                                       (unsigned int)m, (unsigned int)c->n,
                                                       c->d_beg, c->d_beg + 1, 0, end_bit,
                                                       c->stream) == hipSuccess &&
              hipMalloc(&kin, sizeof(uint32_t) * m) == hipSuccess &&
              hipMalloc(&kout, sizeof(uint32_t) * m) == hipSuccess &&
              hipMalloc(&adj2, sizeof(int32_t) * m) == hipSuccess &&
              hipMalloc(&tmp, tb > 0 ? tb : 1) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_nbr_deg, dim3(grid_for(c->m, 256, 65536)), dim3(256), 0, c->stream,
                           c->d_beg, c->d_adj, c->m, kin);
        ok = hipGetLastError() == hipSuccess &&
             rocprim::segmented_radix_sort_pairs_desc(tmp, tb, kin, kout, c->d_adj, adj2,
                                                      (unsigned int)m, (unsigned int)c->n,
                                                      c->d_beg, c->d_beg + 1, 0, end_bit,
                                                      c->stream) == hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
    }
    if (ok) {
        c->d_adj_bu = adj2;
        adj2 = nullptr;
    }
    if (kin) (void)hipFree(kin);
    if (kout) (void)hipFree(kout);
    if (adj2) (void)hipFree(adj2);
    if (tmp) (void)hipFree(tmp);
    (void)hipGetLastError(); // a failed allocation must not linger as the next call's error
    return ok;
}

static int find_hub_component(xbfs_ctx *c, int32_t hub);

int xbfs_create(int device, const int64_t *beg, const int32_t *adj, int64_t n, int64_t m,
                xbfs_ctx **out)
{
    if (!out) return XBFS_ERR_ARG;
    *out = nullptr;
    int rc = xbfs_check_csr(n, m, beg, adj);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return XBFS_ERR_HIP;
    if (device < 0 || device >= ndev) return XBFS_ERR_ARG;
    XB_HIP(hipSetDevice(device));
    xbfs_ctx *c = new (std::nothrow) xbfs_ctx();
    if (!c) return XBFS_ERR_NOMEM;
    c->device = device;
    c->n = n;
    c->m = m;
    c->nwords = ceil_div(n, 64);
    c->ntiles_bu = ceil_div(c->nwords, XB_BU_TILE_WORDS);
    c->fstride = c->ntiles_bu * XB_BU_TILE_WORDS;
    c->ntiles_list = ceil_div(c->nwords, XB_TILE_WORDS);
    int64_t nz = 0, max_deg = 0, hub = -1;
    for (int64_t i = 0; i < n; i++) {
        nz += beg[i + 1] > beg[i];
        if (beg[i + 1] - beg[i] > max_deg) {
            max_deg = beg[i + 1] - beg[i];
            hub = i;
        }
    }
    c->n_nz = nz;
    c->skew = nz > 0 && m > 0 ? (double)max_deg * (double)nz / (double)m : 0.0;
    auto fail = [&](int code) {
        xb_ctx_free_device(c);
        delete c;
        return code;
    };
    int clk_khz = 0;
    if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess)
        return fail(XBFS_ERR_HIP);
    if (clk_khz > 0) c->wall_hz = (double)clk_khz * 1e3;
    if (const char *e = getenv("XBFS_LAUNCH")) c->launch_plain = strcmp(e, "coop") != 0;
    if (const char *e = getenv("XBFS_RESIDENCY_US")) c->res_us = atof(e);
    if (const char *e = getenv("XBFS_NT_BU")) c->nt_bu = atoi(e) != 0 ? 1 : 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persistent, XB_PERSIST_BLOCK,
                                                     XB_LDS_PERSIST) !=
            hipSuccess ||
        per_cu < 1)
        return fail(XBFS_ERR_HIP);
    c->coop_blocks = std::min(per_cu * c->num_cu, XB_PMAX);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(XBFS_ERR_HIP);
    const size_t mm = (size_t)(m > 0 ? m : 1);
    const size_t nw = (size_t)c->nwords;
    bool ok = hipMalloc(&c->d_beg, sizeof(int64_t) * (size_t)(n + 1)) == hipSuccess &&
              hipMalloc(&c->d_adj, sizeof(int32_t) * mm) == hipSuccess &&
              hipMalloc(&c->d_nz, sizeof(uint64_t) * nw) == hipSuccess &&
              hipMalloc(&c->d_hd, sizeof(int2) * (size_t)(n > 0 ? n : 1)) == hipSuccess &&
              hipMalloc(&c->d_vis, sizeof(uint64_t) * nw) == hipSuccess &&
              hipMalloc(&c->d_fbm, sizeof(uint64_t) * (size_t)c->fstride * XB_FRINGS) == hipSuccess &&
              hipMalloc(&c->d_level, sizeof(int32_t) * (size_t)n) == hipSuccess &&
              hipMalloc(&c->d_parent, sizeof(int32_t) * (size_t)n) == hipSuccess &&
              hipMalloc(&c->d_qbuf, sizeof(int32_t) * (size_t)n * XB_QBUFS * XB_BINS) ==
                  hipSuccess &&
              hipMalloc(&c->d_ctl, sizeof(XbCtl)) == hipSuccess &&
              hipMalloc(&c->d_lsg, sizeof(XbLoopShared)) == hipSuccess &&
              hipMalloc(&c->d_part, sizeof(long long) * XB_NFIELDS * XB_PMAX) == hipSuccess &&
              hipMalloc(&c->d_recs, sizeof(XbDevRec) * XB_REC_CAP) == hipSuccess &&
              hipMalloc(&c->d_sched, XB_REC_CAP) == hipSuccess &&
              hipMalloc(&c->d_mlp, sizeof(float) * XB_MLP_NPARAM) == hipSuccess &&
              hipMalloc(&c->d_list, sizeof(int32_t) * (size_t)n) == hipSuccess &&
              hipMalloc(&c->d_ubits, sizeof(uint64_t) * nw) == hipSuccess &&
              hipMalloc(&c->d_tile, sizeof(int32_t) * (size_t)c->ntiles_list * 2) == hipSuccess &&
              hipMalloc(&c->d_list_cnt, sizeof(int32_t)) == hipSuccess &&
              hipHostMalloc(&c->h_ctl, sizeof(XbCtl), hipHostMallocDefault) == hipSuccess;
    if (!ok) return fail(XBFS_ERR_NOMEM);
    if (hipEventCreate(&c->ev_start) != hipSuccess || hipEventCreate(&c->ev_stop) != hipSuccess)
        return fail(XBFS_ERR_HIP);
    if (hipMemcpy(c->d_beg, beg, sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (m > 0 && hipMemcpy(c->d_adj, adj, sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice) !=
                      hipSuccess))
        return fail(XBFS_ERR_HIP);
    hipLaunchKernelGGL(k_build_nz, dim3(grid_for(c->nwords * 64, 256, 65536)), dim3(256), 0,
                       c->stream, c->d_beg, n, c->d_nz, c->nwords);
    // XBFS_ROW_ORDER=csr keeps the uploaded row order (bottom-up scans then
    // start at the row's first entry)
    const char *row_order = getenv("XBFS_ROW_ORDER");
    const bool keep_csr = row_order && strcmp(row_order, "csr") == 0;
    c->hub_first = (!keep_csr && sort_rows_by_degree(c, max_deg)) ? 1 : 0;
    if (n > 0 && c->hub_first) {
        hipLaunchKernelGGL(k_build_hd_sorted, dim3(grid_for(n, 256, 65536)), dim3(256), 0,
                           c->stream, c->d_beg, c->d_adj_bu, n, c->d_hd);
    } else if (n > 0) {
        hipLaunchKernelGGL(k_build_hd, dim3(grid_for(n, 256, 65536)), dim3(256), 0, c->stream,
                           c->d_beg, c->d_adj, n, c->d_hd);
        hipLaunchKernelGGL(k_build_hd_big, dim3(c->num_cu * 8), dim3(256), 0, c->stream, c->d_beg,
                           c->d_adj, n, c->d_hd);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(XBFS_ERR_HIP);
    // XBFS_COMPONENT=off: every traversal runs until its frontier is empty
    const char *comp = getenv("XBFS_COMPONENT");
    if (hub >= 0 && !(comp && strcmp(comp, "off") == 0)) {
        if ((rc = find_hub_component(c, (int32_t)hub)) != XBFS_OK) return fail(rc);
    }
    *out = c;
    return XBFS_OK;
}

void xbfs_destroy(xbfs_ctx *c)
{
    if (!c) return;
    xb_ctx_free_device(c);
    delete c;
}

} // extern "C"

// Resolves opts against defaults; XBFS_OK or XBFS_ERR_ARG.
int xb_resolve_opts(const xbfs_opts *in, xbfs_opts *o)
{
    if (in) *o = *in;
    else xbfs_opts_default(o);
    if (o->block_dim == 0) o->block_dim = XB_PBLOCK;
    if (!pow2_block(o->block_dim)) return XBFS_ERR_ARG;
    if (o->grid_dim < 0 || o->grid_dim > XB_PMAX) return XBFS_ERR_ARG;
    if (o->bin_thread_max <= 0) o->bin_thread_max = XB_DEFAULT_T1;
    if (o->bin_wave_max <= 0) o->bin_wave_max = XB_DEFAULT_T2;
    if (o->bin_block_max <= 0) o->bin_block_max = std::max<int32_t>(XB_DEFAULT_T3, o->bin_wave_max);
    if (o->bin_wave_max < o->bin_thread_max || o->bin_block_max < o->bin_wave_max)
        return XBFS_ERR_ARG;
    if (o->policy < XBFS_POLICY_HEURISTIC || o->policy > XBFS_POLICY_SCHEDULE) return XBFS_ERR_ARG;
    if (o->policy == XBFS_POLICY_SCHEDULE) {
        if (!o->schedule || o->schedule_len <= 0 || o->schedule_len > XB_REC_CAP)
            return XBFS_ERR_ARG;
        for (int i = 0; i < o->schedule_len; i++)
            if (o->schedule[i] < 0 || o->schedule[i] > 2) return XBFS_ERR_ARG;
    }
    if (!(o->alpha >= 0.0f)) return XBFS_ERR_ARG;
    if (o->max_levels < 0 || o->wait_limit_us < 0) return XBFS_ERR_ARG;
    return XBFS_OK;
}

static int ensure_events(xbfs_ctx *c, size_t need)
{
    while (c->ev_pool.size() < need) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return XBFS_ERR_HIP;
        c->ev_pool.push_back(e);
    }
    return XBFS_OK;
}

static XbArgs make_args(xbfs_ctx *c, int32_t src, const xbfs_opts &o)
{
    XbArgs a;
    a.level = c->d_level;
    a.parent = o.want_parent ? c->d_parent : nullptr;
    a.beg = c->d_beg;
    a.adj = c->d_adj;
    a.adj_bu = c->hub_first ? c->d_adj_bu : c->d_adj;
    a.nz = c->d_nz;
    a.hd = c->d_hd;
    a.vis = c->d_vis;
    a.fbm = c->d_fbm;
    a.qbuf = c->d_qbuf;
    a.ctl = c->d_ctl;
    a.lsg = c->d_lsg;
    a.part = c->d_part;
    a.recs = c->d_recs;
    a.rec_cap = XB_REC_CAP;
    a.sched = c->d_sched;
    a.sched_len = o.policy == XBFS_POLICY_SCHEDULE ? o.schedule_len : 0;
    // the learned policy decides only on graphs of the skew it was trained on
    a.mlp = (o.policy == XBFS_POLICY_LEARNED && c->mlp_loaded && c->skew >= XBFS_MLP_MIN_SKEW)
                ? c->d_mlp
                : nullptr;
    a.n = c->n;
    a.m = c->m;
    a.nwords = c->nwords;
    a.ntiles = c->ntiles_bu;
    a.fstride = c->fstride;
    a.n_nz = c->n_nz;
    a.n_target = c->n_nz;
    if (c->d_comp && src >= 0 && (int64_t)src < c->n &&
        ((c->h_comp[(size_t)src >> 6] >> (src & 63)) & 1ull)) {
        a.nz = c->d_comp;
        a.n_target = c->n_comp;
    }
    a.srcs = nullptr;
    a.nsrc = 0;
    a.hub_first = c->hub_first;
    a.nt_bu = c->nt_bu >= 0 ? c->nt_bu : (8 * (long long)c->n > XB_NT_HUB_BYTES ? 1 : 0);
    a.src = src;
    a.t1 = o.bin_thread_max;
    a.t2 = o.bin_wave_max;
    a.t3 = o.bin_block_max;
    a.policy = o.policy;
    a.alpha = o.alpha;
    a.max_levels = (int32_t)std::min<int64_t>(c->n + 1, INT32_MAX - 4);
    if (o.max_levels > 0) a.max_levels = std::min(a.max_levels, o.max_levels);
    return a;
}

static void fill_rec(xbfs_level_rec &r, int32_t L, int32_t mode, int32_t htag, int32_t used,
                     long long n_f, long long m_f, long long n_u, long long m_u, int64_t E,
                     double t_s)
{
    r.level = L;
    r.mode = mode;
    r.tag = xbfs_mode_to_tag(mode);
    r.policy = used;
    r.n_f = n_f;
    r.m_f = m_f;
    r.n_u = n_u;
    r.m_u = m_u;
    r.ratio = (float)m_f / (float)E;
    r.pad = (float)htag;
    r.t_s = t_s;
    r.bu_cand = r.bu_scan_v = r.bu_scan_e = 0;
}

// Enqueues the kernels of level L in `mode` on the context stream: zeroing
// schedule, the frontier queue rebuild a top-down mode needs, the expansion.
static hipError_t launch_level(xbfs_ctx *c, const XbArgs &a, int grid, int32_t L, int32_t mode,
                               int32_t q_valid, int32_t need_sweep, bool enqueue)
{
    hipLaunchKernelGGL(k_ph_prep, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L);
    if (mode != XBFS_MODE_BU && !q_valid && !need_sweep)
        hipLaunchKernelGGL(k_ph_sweep, dim3(grid), dim3(XB_PBLOCK), XB_LDS_PHASE, c->stream, a, L,
                           0, 0);
    if (mode == XBFS_MODE_TD_SF)
        hipLaunchKernelGGL(k_ph_td_sf, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L,
                           (int32_t)enqueue);
    else if (mode == XBFS_MODE_TD_SS)
        hipLaunchKernelGGL(k_ph_td_ss, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L);
    else
        hipLaunchKernelGGL(k_ph_bu, dim3(grid), dim3(XB_PBLOCK), XB_LDS_PHASE, c->stream, a, L);
    return hipGetLastError();
}

// Host-driven level loop: one kernel per phase and one small D2H copy of the
// control block per level; the decision runs on the host from the same
// device statistics the persistent kernel reads.
static int bfs_host_loop(xbfs_ctx *c, const XbArgs &a, const xbfs_opts &o, xbfs_level_rec *rec,
                         int32_t rec_cap, xbfs_stats *st)
{
    const int grid = o.grid_dim > 0 ? o.grid_dim : c->num_cu * 8;
    XbCtl *ctl = c->d_ctl;
    if (ensure_events(c, 4) != XBFS_OK) return XBFS_ERR_HIP;
    XB_HIP(hipEventRecord(c->ev_start, c->stream));
    hipLaunchKernelGGL(k_ph_init, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a);
    if (a.nsrc > 0) hipLaunchKernelGGL(k_ph_seed, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a);
    XB_LAUNCHED();
    XbLoop s;
    xb_loop_init(&s);
    if (a.nsrc > 0) s.q_valid = 0; // the level-0 queue comes from a sweep of F_0
    double td_s = 0, bu_s = 0, gen_s = 0;
    long long pend_n = 0, pend_m = 0, n_f = 0;
    for (;;) {
        const int32_t L = s.L;
        double sweep_s = 0;
        if (s.need_sweep) {
            if (o.time_levels) XB_HIP(hipEventRecord(c->ev_pool[0], c->stream));
            hipLaunchKernelGGL(k_ph_sweep, dim3(grid), dim3(XB_PBLOCK), XB_LDS_PHASE, c->stream, a,
                               L, 1, 1);
            XB_LAUNCHED();
            if (o.time_levels) XB_HIP(hipEventRecord(c->ev_pool[1], c->stream));
        }
        hipLaunchKernelGGL(k_part_reduce, dim3(1), dim3(1024), 0, c->stream, a);
        XB_HIP(hipMemcpyAsync(c->h_ctl, ctl, sizeof(XbCtl), hipMemcpyDeviceToHost, c->stream));
        XB_HIP(hipStreamSynchronize(c->stream));
        if (o.time_levels && s.need_sweep) {
            float ms = 0;
            XB_HIP(hipEventElapsedTime(&ms, c->ev_pool[0], c->ev_pool[1]));
            sweep_s = ms * 1e-3;
            gen_s += sweep_s;
        }
        const XbCtl &h = *c->h_ctl;
        long long m_f;
        if (s.need_sweep) {
            n_f = h.sum[xb_f_sweep(L, 0)];
            m_f = h.sum[xb_f_sweep(L, 1)];
        } else {
            n_f = h.sum[xb_f_ring(L, 0)];
            m_f = h.sum[xb_f_ring(L, 1)];
        }
        pend_n = h.sum[xb_f_ring(L + 1, 0)];
        pend_m = h.sum[xb_f_ring(L + 1, 1)];
        const long long p2_next = h.sum[xb_f_ring(L + 1, 2)];
        if (rec && L > 0 && L - 1 < rec_cap) {
            rec[L - 1].bu_cand = h.sum[xb_f_bu(L - 1, 0)];
            rec[L - 1].bu_scan_v = h.sum[xb_f_bu(L - 1, 1)];
            rec[L - 1].bu_scan_e = h.sum[xb_f_bu(L - 1, 2)];
        }
        long long n_u = 0, m_u = 0;
        if (!xb_level_open(s, a, n_f, m_f, pend_n, pend_m, &n_u, &m_u)) break;
        if (L >= a.max_levels) return XBFS_ERR_LIMIT;
        float y[XB_MLP_NO];
        const float *pred = nullptr;
        if (a.mlp && L <= XBFS_MLP_MAX_LEVEL) {
            float x[XB_MLP_NI];
            xbfs_mlp_features(L, n_f, m_f, n_u, m_u, s.prev_mode, x);
            xbfs_mlp_forward1(c->h_mlp.data(), x, y);
            pred = y;
        }
        int32_t htag = 0, used = 0;
        // host copy of the schedule (a.sched is the device copy)
        const int32_t mode = xbfs_choose_mode(a.policy, o.schedule, a.sched_len, pred, &s.hs, L,
                                              m_f, a.m, a.alpha, &htag, &used);
        if (o.time_levels) XB_HIP(hipEventRecord(c->ev_pool[2], c->stream));
        XB_HIP(launch_level(c, a, grid, L, mode, s.q_valid, s.need_sweep, p2_next == 0));
        double lvl_s = sweep_s;
        if (o.time_levels) {
            XB_HIP(hipEventRecord(c->ev_pool[3], c->stream));
            XB_HIP(hipEventSynchronize(c->ev_pool[3]));
            float ms = 0;
            XB_HIP(hipEventElapsedTime(&ms, c->ev_pool[2], c->ev_pool[3]));
            lvl_s += ms * 1e-3;
            if (mode == XBFS_MODE_BU) bu_s += ms * 1e-3;
            else td_s += ms * 1e-3;
        }
        if (rec && L < rec_cap) fill_rec(rec[L], L, mode, htag, used, n_f, m_f, n_u, m_u, c->m, lvl_s);
        s.q_valid = (mode == XBFS_MODE_TD_SF && p2_next == 0);
        s.need_sweep = (mode == XBFS_MODE_TD_SS);
        s.prev_mode = mode;
        s.L = L + 1;
    }
    XB_HIP(hipEventRecord(c->ev_stop, c->stream));
    XB_HIP(hipEventSynchronize(c->ev_stop));
    float ms = 0.0f;
    XB_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
    if (st) {
        const int32_t L = s.L;
        st->t_bfs_s = ms * 1e-3;
        st->n_levels = L;
        st->max_level = pend_n > 0 ? L + 1 : (n_f > 0 ? L : L - 1);
        st->v_cc = s.cum_n + pend_n + c->h_ctl->seed_iso;
        st->m_cc = s.cum_m + pend_m;
        if (st->max_level < 0 && c->h_ctl->seed_iso > 0) st->max_level = 0; // isolated seeds only
        st->td_s = td_s;
        st->bu_s = bu_s;
        st->gen_s = gen_s;
    }
    return XBFS_OK;
}

// Persistent driver: the whole level loop in one launch of at most one block
// per CU.  A plain launch costs about 6 us on MI355X, a cooperative one about
// 19 us (an empty kernel of the same shape, event to event), so the kernel is
// launched plainly while the GPU shows it every block at once: its first
// barrier then bounds its wait by res_ticks (XBFS_RESIDENCY_US, default
// 20 ms) and, when some block did not arrive in time, every block leaves with
// code 3; the traversal is then relaunched cooperatively (co-residency
// guaranteed), and the context keeps launching cooperatively from then on.
// XBFS_LAUNCH=coop makes every launch cooperative.  The device time of a
// relaunched traversal includes the first attempt.
static int bfs_persistent(xbfs_ctx *c, const XbArgs &a, const xbfs_opts &o, xbfs_level_rec *rec,
                          int32_t rec_cap, xbfs_stats *st)
{
    int grid = c->coop_blocks;
    if (o.grid_dim > 0 && o.grid_dim < grid) grid = o.grid_dim;
    const unsigned long long res = c->launch_plain ? res_ticks(c) : 0ull;
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, c->stream, c->d_ctl, wait_ticks(c, o), res);
    XB_LAUNCHED();
    XB_HIP(hipEventRecord(c->ev_start, c->stream));
    XbArgs args = a;
    void *kargs[] = {&args};
    if (res)
        XB_HIP(hipLaunchKernel((const void *)k_persistent, dim3(grid), dim3(XB_PERSIST_BLOCK), kargs,
                               XB_LDS_PERSIST, c->stream));
    else
        XB_HIP(hipLaunchCooperativeKernel((const void *)k_persistent, dim3(grid),
                                          dim3(XB_PERSIST_BLOCK), kargs, XB_LDS_PERSIST, c->stream));
    XB_HIP(hipEventRecord(c->ev_stop, c->stream));
    XB_HIP(hipEventSynchronize(c->ev_stop));
    XB_HIP(hipMemcpy(c->h_ctl, c->d_ctl, sizeof(XbCtl), hipMemcpyDeviceToHost));
    if (res && c->h_ctl->error == 3) {
        // some block was not resident with the others: cooperative from now on
        c->launch_plain = 0;
        c->plain_fallbacks++;
        hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, c->stream, c->d_ctl, wait_ticks(c, o),
                           0ull);
        XB_LAUNCHED();
        XB_HIP(hipLaunchCooperativeKernel((const void *)k_persistent, dim3(grid),
                                          dim3(XB_PERSIST_BLOCK), kargs, XB_LDS_PERSIST, c->stream));
        XB_HIP(hipEventRecord(c->ev_stop, c->stream));
        XB_HIP(hipEventSynchronize(c->ev_stop));
        XB_HIP(hipMemcpy(c->h_ctl, c->d_ctl, sizeof(XbCtl), hipMemcpyDeviceToHost));
    }
    const XbCtl &h = *c->h_ctl;
    if (h.error == 1) return XBFS_ERR_DEVICE;
    if (h.error == 2) return XBFS_ERR_LIMIT;
    float ms = 0.0f;
    XB_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
    double td_s = 0, bu_s = 0;
    const int32_t nrec = std::min<int32_t>(h.n_levels, XB_REC_CAP);
    if (nrec > 0 && (rec || o.time_levels || st)) {
        std::vector<XbDevRec> dr((size_t)nrec);
        XB_HIP(hipMemcpy(dr.data(), c->d_recs, sizeof(XbDevRec) * (size_t)nrec,
                         hipMemcpyDeviceToHost));
        for (int32_t L = 0; L < nrec; L++) {
            const XbDevRec &d = dr[(size_t)L];
            const double t_s = (double)(d.t1 - d.t0) / c->wall_hz;
            if (d.mode == XBFS_MODE_BU) bu_s += t_s;
            else td_s += t_s;
            if (rec && L < rec_cap) {
                fill_rec(rec[L], L, d.mode, (int8_t)(d.htag & 0xff), (d.htag >> 8) & 0xff, d.n_f, d.m_f,
                         d.n_u, d.m_u, c->m, t_s);
                rec[L].bu_cand = d.bu_cand;
                rec[L].bu_scan_v = d.bu_scan_v;
                rec[L].bu_scan_e = d.bu_scan_e;
            }
        }
    }
    if (st) {
        st->t_bfs_s = ms * 1e-3;
        st->n_levels = h.n_levels;
        st->max_level = h.max_level;
        st->v_cc = h.v_cc + h.seed_iso;
        st->m_cc = h.m_cc;
        if (st->max_level < 0 && h.seed_iso > 0) st->max_level = 0; // isolated seeds only
        st->td_s = td_s;
        st->bu_s = bu_s;
        st->gen_s = 0;
    }
    return XBFS_OK;
}

extern "C" {

} // extern "C"

// Shared body of xbfs_bfs (nsrc == 0: single source src) and xbfs_bfs_multi
// (nsrc seeds in srcs, host memory, already range-checked).
static int bfs_run(xbfs_ctx *c, int32_t src, const int32_t *srcs, int32_t nsrc,
                   const xbfs_opts *opts, int32_t *level_out, int32_t *parent_out,
                   xbfs_level_rec *rec, int32_t rec_cap, xbfs_stats *st)
{
    if (rec_cap < 0 || (rec_cap > 0 && !rec)) return XBFS_ERR_ARG;
    xbfs_opts o;
    int rc = xb_resolve_opts(opts, &o);
    if (rc) return rc;
    if (parent_out) o.want_parent = 1;
    XB_HIP(hipSetDevice(c->device));
    if (o.policy == XBFS_POLICY_SCHEDULE)
        XB_HIP(hipMemcpyAsync(c->d_sched, o.schedule, (size_t)o.schedule_len,
                              hipMemcpyHostToDevice, c->stream));
    XbArgs a = make_args(c, src, o);
    if (nsrc > 0) {
        if (nsrc > c->srcs_cap) {
            if (c->d_srcs) XB_HIP(hipFree(c->d_srcs));
            c->d_srcs = nullptr;
            c->srcs_cap = 0;
            if (hipMalloc(&c->d_srcs, sizeof(int32_t) * (size_t)nsrc) != hipSuccess)
                return XBFS_ERR_NOMEM;
            c->srcs_cap = nsrc;
        }
        XB_HIP(hipMemcpyAsync(c->d_srcs, srcs, sizeof(int32_t) * (size_t)nsrc,
                              hipMemcpyHostToDevice, c->stream));
        a.srcs = c->d_srcs;
        a.nsrc = nsrc;
        a.src = srcs[0];
        a.nz = c->d_nz; // seeds may lie in several components
        a.n_target = c->n_nz;
    }
    rc = o.persistent ? bfs_persistent(c, a, o, rec, rec_cap, st)
                      : bfs_host_loop(c, a, o, rec, rec_cap, st);
    if (rc) return rc;
    if (level_out)
        XB_HIP(hipMemcpy(level_out, c->d_level, sizeof(int32_t) * (size_t)c->n,
                         hipMemcpyDeviceToHost));
    if (parent_out)
        XB_HIP(hipMemcpy(parent_out, c->d_parent, sizeof(int32_t) * (size_t)c->n,
                         hipMemcpyDeviceToHost));
    return XBFS_OK;
}

extern "C" {

int xbfs_bfs(xbfs_ctx *c, int32_t src, const xbfs_opts *opts, int32_t *level_out,
             int32_t *parent_out, xbfs_level_rec *rec, int32_t rec_cap, xbfs_stats *st)
{
    if (!c) return XBFS_ERR_ARG;
    if (src < 0 || (int64_t)src >= c->n) return XBFS_ERR_ARG;
    return bfs_run(c, src, nullptr, 0, opts, level_out, parent_out, rec, rec_cap, st);
}

} // extern "C"

// One traversal from the largest-degree vertex marks its component (the giant
// component of a skewed graph).  When the component leaves out some
// non-isolated vertex, a later traversal from a member ends as soon as every
// member has its level instead of expanding its last level to find nothing,
// and its bottom-up candidates skip the vertices outside (no path reaches
// them).  When memory for the bitmap is not available every traversal runs
// until its frontier is empty, as without the component.
static int find_hub_component(xbfs_ctx *c, int32_t hub)
{
    xbfs_opts o;
    xbfs_opts_default(&o);
    xbfs_stats st;
    int rc = bfs_run(c, hub, nullptr, 0, &o, nullptr, nullptr, nullptr, 0, &st);
    if (rc) return rc;
    if (st.v_cc >= c->n_nz) return XBFS_OK; // nothing to leave out
    if (hipMalloc(&c->d_comp, sizeof(uint64_t) * (size_t)c->nwords) != hipSuccess) {
        (void)hipGetLastError();
        c->d_comp = nullptr;
        return XBFS_OK;
    }
    hipLaunchKernelGGL(k_reached_bits, dim3(grid_for(c->nwords * 64, 256, 65536)), dim3(256), 0,
                       c->stream, c->d_level, c->n, c->d_comp, c->nwords);
    XB_LAUNCHED();
    c->h_comp.resize((size_t)c->nwords);
    XB_HIP(hipMemcpyAsync(c->h_comp.data(), c->d_comp, sizeof(uint64_t) * (size_t)c->nwords,
                          hipMemcpyDeviceToHost, c->stream));
    XB_HIP(hipStreamSynchronize(c->stream));
    c->n_comp = st.v_cc;
    return XBFS_OK;
}

extern "C" {

int xbfs_bfs_multi(xbfs_ctx *c, const int32_t *srcs, int32_t nsrc, const xbfs_opts *opts,
                   int32_t *level_out, int32_t *parent_out, xbfs_level_rec *rec, int32_t rec_cap,
                   xbfs_stats *st)
{
    if (!c || !srcs || nsrc <= 0) return XBFS_ERR_ARG;
    for (int32_t i = 0; i < nsrc; i++)
        if (srcs[i] < 0 || (int64_t)srcs[i] >= c->n) return XBFS_ERR_ARG;
    return bfs_run(c, srcs[0], srcs, nsrc, opts, level_out, parent_out, rec, rec_cap, st);
}

int xbfs_bfs_multigraph(xbfs_ctx *c, const int64_t *vert_off, int32_t k, const int32_t *local_src,
                        const xbfs_opts *opts, int32_t *level_out, xbfs_stats *st)
{
    if (!c || !vert_off || k <= 0 || !local_src) return XBFS_ERR_ARG;
    if (vert_off[0] != 0 || vert_off[k] != c->n) return XBFS_ERR_ARG;
    std::vector<int32_t> g((size_t)k);
    for (int32_t i = 0; i < k; i++) {
        if (vert_off[i + 1] < vert_off[i]) return XBFS_ERR_ARG;
        if (local_src[i] < 0 || vert_off[i] + local_src[i] >= vert_off[i + 1]) return XBFS_ERR_ARG;
        g[(size_t)i] = (int32_t)(vert_off[i] + local_src[i]);
    }
    return bfs_run(c, g[0], g.data(), k, opts, level_out, nullptr, nullptr, 0, st);
}

} // extern "C"

// Median over reps runs of the objective of xbfs_autotune.
static int tune_measure(xbfs_ctx *c, const int32_t *srcs, int32_t nsrc, int32_t multi,
                        int32_t reps, const xbfs_opts &o, double *t_out)
{
    std::vector<double> ts((size_t)reps);
    double total = 0.0;
    const int32_t nrun = multi ? 1 : nsrc;
    for (int32_t i = 0; i < nrun; i++) {
        for (int32_t r = 0; r < reps; r++) {
            xbfs_stats st;
            const int rc = multi ? bfs_run(c, srcs[0], srcs, nsrc, &o, nullptr, nullptr, nullptr, 0, &st)
                                 : bfs_run(c, srcs[i], nullptr, 0, &o, nullptr, nullptr, nullptr, 0, &st);
            if (rc) return rc;
            ts[(size_t)r] = st.t_bfs_s;
        }
        std::sort(ts.begin(), ts.end());
        total += ts[(size_t)reps / 2];
    }
    *t_out = total;
    return XBFS_OK;
}

extern "C" {

int xbfs_autotune(xbfs_ctx *c, const int32_t *srcs, int32_t nsrc, int32_t multi, int32_t reps,
                  xbfs_opts *opts, xbfs_tune_rec *trials, int32_t cap, int32_t *ntrials)
{
    if (!c || !srcs || nsrc <= 0 || !opts || reps <= 0 || cap < 0 || (cap > 0 && !trials))
        return XBFS_ERR_ARG;
    for (int32_t i = 0; i < nsrc; i++)
        if (srcs[i] < 0 || (int64_t)srcs[i] >= c->n) return XBFS_ERR_ARG;
    xbfs_opts best;
    int rc = xb_resolve_opts(opts, &best);
    if (rc) return rc;
    const int32_t full = best.persistent ? c->coop_blocks : c->num_cu * 8;
    if (best.grid_dim <= 0 || best.grid_dim > full) best.grid_dim = full;
    const std::vector<int32_t> cand_t1 = {8, 16, 32, 64, 128};
    const std::vector<int32_t> cand_t2 = {256, 512, 1024, 2048, 4096};
    const std::vector<int32_t> cand_t3 = {4096, 16384, 65536, 262144};
    std::vector<int32_t> cand_g;
    for (int32_t f : {8, 6, 4})
        if (full * f / 8 >= 1) cand_g.push_back(full * f / 8);
    int32_t nt = 0;
    auto record = [&](const xbfs_opts &o, double t) {
        if (nt < cap) {
            xbfs_tune_rec &r = trials[nt];
            r.bin_thread_max = o.bin_thread_max;
            r.bin_wave_max = o.bin_wave_max;
            r.bin_block_max = o.bin_block_max;
            r.grid_dim = o.grid_dim;
            r.t_s = t;
        }
        nt++;
    };
    double t_best = 0.0;
    rc = tune_measure(c, srcs, nsrc, multi, reps, best, &t_best);
    if (rc) return rc;
    record(best, t_best);
    for (int pass = 0; pass < 2; pass++) {
        for (int p = 0; p < 4; p++) {
            const std::vector<int32_t> &cand = p == 0 ? cand_t1 : p == 1 ? cand_t2 : p == 2 ? cand_t3 : cand_g;
            for (int32_t v : cand) {
                xbfs_opts o = best;
                int32_t *field = p == 0 ? &o.bin_thread_max : p == 1 ? &o.bin_wave_max
                               : p == 2 ? &o.bin_block_max : &o.grid_dim;
                if (*field == v) continue;
                *field = v;
                if (o.bin_wave_max < o.bin_thread_max || o.bin_block_max < o.bin_wave_max) continue;
                double t = 0.0;
                rc = tune_measure(c, srcs, nsrc, multi, reps, o, &t);
                if (rc) return rc;
                record(o, t);
                if (t < t_best) {
                    t_best = t;
                    best = o;
                }
            }
        }
    }
    *opts = best;
    if (ntrials) *ntrials = std::min(nt, cap);
    return XBFS_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// 64-source batch (kernels in xbfs_batch.hip)
// ---------------------------------------------------------------------------
#define XB_B64_HIST_BUDGET (16ll << 30) /* bytes of frontier history at most */

// The relabelled graph (vertices by degree, descending; rows by neighbour
// rank, ascending) and the batch buffers, once per context.
static int b64_setup(xbfs_ctx *c)
{
    if (c->b64_ready) return XBFS_OK;
    const int64_t n = c->n, m = c->m;
    if (n > INT32_MAX) return XBFS_ERR_UNSUPPORTED;
    const int grid = grid_for(n, 256, 65536);
    uint32_t *kin = nullptr, *kout = nullptr;
    int32_t *ids = nullptr, *radj_in = nullptr;
    long long *spart = nullptr, *soff = nullptr;
    void *tmp = nullptr;
    const int64_t T = ceil_div(n, XB_SCAN_CHUNK);
    int end_bit = 1;
    while (end_bit < 32 && (n >> end_bit) != 0) end_bit++;
    int rc = XBFS_OK;
    auto cleanup = [&]() {
        for (void *p : {(void *)kin, (void *)kout, (void *)ids, (void *)radj_in, (void *)spart,
                        (void *)soff, tmp})
            if (p) (void)hipFree(p);
        (void)hipGetLastError();
    };
    if (hipMalloc(&c->d_b64_rank, sizeof(int32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&c->d_b64_ord, sizeof(int32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&c->d_b64_rdeg, sizeof(int32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&c->d_b64_rbeg, sizeof(int64_t) * (size_t)(n + 1)) != hipSuccess ||
        hipMalloc(&c->d_b64_radj, sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1)) != hipSuccess ||
        hipMalloc(&kin, sizeof(uint32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&kout, sizeof(uint32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&ids, sizeof(int32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&spart, sizeof(long long) * (size_t)(T + 1)) != hipSuccess ||
        hipMalloc(&soff, sizeof(long long) * (size_t)n) != hipSuccess) {
        cleanup();
        return XBFS_ERR_NOMEM;
    }
    // ranks: a stable radix sort of the vertex ids by degree, descending
    size_t tb = 0;
    if (rocprim::radix_sort_pairs_desc(nullptr, tb, kin, kout, ids, c->d_b64_ord, (unsigned int)n, 0,
                                       32, c->stream) != hipSuccess ||
        hipMalloc(&tmp, tb > 0 ? tb : 1) != hipSuccess) {
        cleanup();
        return XBFS_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_b64_deg_key, dim3(grid), dim3(256), 0, c->stream, c->d_beg, n, kin, ids);
    if (rocprim::radix_sort_pairs_desc(tmp, tb, kin, kout, ids, c->d_b64_ord, (unsigned int)n, 0, 32,
                                       c->stream) != hipSuccess)
        rc = XBFS_ERR_HIP;
    (void)hipFree(tmp);
    tmp = nullptr;
    if (rc == XBFS_OK) {
        hipLaunchKernelGGL(k_b64_rank, dim3(grid), dim3(256), 0, c->stream, c->d_b64_ord, kout, n,
                           c->d_b64_rank, c->d_b64_rdeg);
        // row offsets: exclusive scan of the degrees by rank
        int32_t *d_err = nullptr;
        if (xb_scan_i32(c->stream, c->d_b64_rdeg, n, reinterpret_cast<long long *>(c->d_b64_rbeg),
                        reinterpret_cast<unsigned long long *>(spart), &d_err) != hipSuccess)
            rc = XBFS_ERR_HIP;
        hipLaunchKernelGGL(k_b64_set_end, dim3(1), dim3(64), 0, c->stream, c->d_b64_rbeg, n, m);
        int32_t h_err = 0;
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&h_err, d_err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = XBFS_ERR_HIP;
        else if (h_err)
            rc = XBFS_ERR_DEVICE;
    }
    // 32-edge segments: count per rank (into the ids buffer, no longer
    // needed), exclusive scan, fill
    if (rc == XBFS_OK) {
        int32_t *scnt = ids;
        hipLaunchKernelGGL(k_b64_seg_count, dim3(grid), dim3(256), 0, c->stream, c->d_b64_rdeg, n, scnt);
        int32_t *d_err = nullptr, h_err = 0;
        if (xb_scan_i32(c->stream, scnt, n, soff, reinterpret_cast<unsigned long long *>(spart),
                        &d_err) != hipSuccess ||
            hipMemcpyAsync(&h_err, d_err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess || h_err)
            rc = h_err ? XBFS_ERR_DEVICE : XBFS_ERR_HIP;
        long long last_off = 0;
        int32_t last_cnt = 0;
        if (rc != XBFS_OK || hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&last_off, soff + (n - 1), sizeof(long long), hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipMemcpyAsync(&last_cnt, scnt + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = XBFS_ERR_HIP;
        const int64_t nseg = last_off + last_cnt;
        if (rc == XBFS_OK &&
            (nseg > INT32_MAX ||
             hipMalloc(&c->d_b64_segr, sizeof(int32_t) * (size_t)std::max<int64_t>(nseg, 1)) != hipSuccess ||
             hipMalloc(&c->d_b64_sege, sizeof(int64_t) * (size_t)(nseg + 1)) != hipSuccess))
            rc = nseg > INT32_MAX ? XBFS_ERR_UNSUPPORTED : XBFS_ERR_NOMEM;
        if (rc == XBFS_OK) {
            hipLaunchKernelGGL(k_b64_seg_fill, dim3(grid), dim3(256), 0, c->stream, c->d_b64_rbeg, n, scnt,
                               soff, c->d_b64_segr, c->d_b64_sege, nseg, m);
            if (hipGetLastError() != hipSuccess) rc = XBFS_ERR_HIP;
            c->b64_nseg = nseg;
        }
    }
    // rows in rank ids, then each row sorted ascending (largest degree first);
    // without the sort's memory or past its 32-bit counts the rows keep their
    // input order (the traversal does not depend on it)
    if (rc == XBFS_OK && m > 0) {
        const bool sortable = m <= INT32_MAX && hipMalloc(&radj_in, sizeof(int32_t) * (size_t)m) == hipSuccess;
        (void)hipGetLastError();
        int32_t *dst = sortable ? radj_in : c->d_b64_radj;
        hipLaunchKernelGGL(k_b64_rows, dim3(grid_for(n * 64, 256, 65536)), dim3(256), 0, c->stream,
                           c->d_beg, c->d_adj, c->d_b64_ord, c->d_b64_rank, c->d_b64_rbeg, n, dst);
        if (hipGetLastError() != hipSuccess) rc = XBFS_ERR_HIP;
        if (rc == XBFS_OK && sortable) {
            bool ok = rocprim::segmented_radix_sort_keys(nullptr, tb, radj_in, c->d_b64_radj,
                                                         (unsigned int)m, (unsigned int)n, c->d_b64_rbeg,
                                                         c->d_b64_rbeg + 1, 0, end_bit,
                                                         c->stream) == hipSuccess &&
                      hipMalloc(&tmp, tb > 0 ? tb : 1) == hipSuccess &&
                      rocprim::segmented_radix_sort_keys(tmp, tb, radj_in, c->d_b64_radj,
                                                         (unsigned int)m, (unsigned int)n, c->d_b64_rbeg,
                                                         c->d_b64_rbeg + 1, 0, end_bit,
                                                         c->stream) == hipSuccess;
            if (!ok) {
                // keep the unsorted rows
                (void)hipGetLastError();
                if (hipMemcpyAsync(c->d_b64_radj, radj_in, sizeof(int32_t) * (size_t)m,
                                   hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                    rc = XBFS_ERR_HIP;
            }
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) rc = XBFS_ERR_HIP;
    }
    cleanup();
    if (rc) return rc;
    const int64_t H = std::max<int64_t>(8, std::min<int64_t>(XB_B64_HMAX, XB_B64_HIST_BUDGET / (8 * std::max<int64_t>(n, 1))));
    c->b64_H = (int32_t)H;
    if (hipMalloc(&c->d_b64_F, sizeof(uint64_t) * (size_t)n * (size_t)H) != hipSuccess ||
        hipMalloc(&c->d_b64_vis, sizeof(uint64_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&c->d_b64_acc, sizeof(uint64_t) * (size_t)(n + 15 * XB_B64_ACC_HOT)) != hipSuccess ||
        hipMalloc(&c->d_b64_rec, sizeof(uint64_t) * 8 * (size_t)n) != hipSuccess ||
        hipMalloc(&c->d_b64_lv, sizeof(int32_t) * (size_t)n * 64) != hipSuccess ||
        hipMalloc(&c->d_b64_out, sizeof(long long) * XB_B64_OUT) != hipSuccess ||
        hipMalloc(&c->d_b64_srcs, sizeof(int32_t) * 64) != hipSuccess)
        return XBFS_ERR_NOMEM;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_b64_persistent, XB_B64_BLOCK, 0) !=
            hipSuccess ||
        per_cu < 1)
        return XBFS_ERR_HIP;
    c->b64_grid = std::min(per_cu * c->num_cu, XB_PMAX);
    for (hipEvent_t &e : c->ev_b64)
        if (!e) XB_HIP(hipEventCreate(&e));
    c->b64_ready = 1;
    return XBFS_OK;
}

extern "C" {

int xbfs_bfs_batch64(xbfs_ctx *c, const int32_t *srcs, int32_t nsrc, const xbfs_opts *opts,
                     int32_t *levels_out, xbfs_stats *st)
{
    if (!c || !srcs || nsrc <= 0 || nsrc > 64 || c->n <= 0) return XBFS_ERR_ARG;
    for (int32_t i = 0; i < nsrc; i++)
        if (srcs[i] < 0 || (int64_t)srcs[i] >= c->n) return XBFS_ERR_ARG;
    xbfs_opts o;
    int rc = xb_resolve_opts(opts, &o);
    if (rc) return rc;
    XB_HIP(hipSetDevice(c->device));
    rc = b64_setup(c);
    if (rc) return rc;
    const int64_t n = c->n;
    XbB64Args a;
    a.rbeg = c->d_b64_rbeg;
    a.radj = c->d_b64_radj;
    a.rdeg = c->d_b64_rdeg;
    a.rank = c->d_b64_rank;
    a.ord = c->d_b64_ord;
    a.F = c->d_b64_F;
    a.vis = c->d_b64_vis;
    a.acc = c->d_b64_acc;
    a.rec = c->d_b64_rec;
    a.lv = c->d_b64_lv;
    a.part = c->d_part;
    a.ctl = c->d_ctl;
    a.out = c->d_b64_out;
    a.srcs = c->d_b64_srcs;
    a.sched = c->d_sched;
    a.n = n;
    a.m = c->m;
    a.n_nz = c->n_nz;
    a.seg_r = c->d_b64_segr;
    a.seg_e = c->d_b64_sege;
    a.nseg = c->b64_nseg;
    a.all = nsrc == 64 ? ~0ull : ((1ull << nsrc) - 1ull);
    a.nsrc = nsrc;
    a.H = c->b64_H;
    a.policy = o.policy;
    a.sched_len = o.policy == XBFS_POLICY_SCHEDULE ? o.schedule_len : 0;
    a.alpha = o.alpha;
    int grid = c->b64_grid;
    if (o.grid_dim > 0 && o.grid_dim < grid) grid = o.grid_dim;
    XB_HIP(hipMemcpyAsync(c->d_b64_srcs, srcs, sizeof(int32_t) * (size_t)nsrc,
                          hipMemcpyHostToDevice, c->stream));
    if (o.policy == XBFS_POLICY_SCHEDULE)
        XB_HIP(hipMemcpyAsync(c->d_sched, o.schedule, (size_t)o.schedule_len,
                              hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, c->stream, c->d_ctl, wait_ticks(c, o), 0ull);
    XB_LAUNCHED();
    XB_HIP(hipEventRecord(c->ev_b64[0], c->stream));
    void *kargs[] = {&a};
    XB_HIP(hipLaunchCooperativeKernel((const void *)k_b64_persistent, dim3(grid), dim3(XB_B64_BLOCK),
                                      kargs, 0, c->stream));
    XB_HIP(hipEventRecord(c->ev_b64[1], c->stream));
    XB_HIP(hipEventSynchronize(c->ev_b64[1]));
    float ms = 0.0f;
    XB_HIP(hipEventElapsedTime(&ms, c->ev_b64[0], c->ev_b64[1]));
    XB_HIP(hipMemcpy(c->h_ctl, c->d_ctl, sizeof(XbCtl), hipMemcpyDeviceToHost));
    if (c->h_ctl->error) return XBFS_ERR_DEVICE;
    long long h_out[XB_B64_OUT];
    XB_HIP(hipMemcpy(h_out, c->d_b64_out, sizeof(h_out), hipMemcpyDeviceToHost));
    const int32_t L = (int32_t)h_out[0];
    // per-direction device time from the kernel's wall-clock stamps
    double td_s = 0, bu_s = 0;
    for (int32_t l = 0; l < L; l++) {
        const double t = (double)(h_out[8 + l + 1] - h_out[8 + l]) / c->wall_hz;
        if (h_out[8 + XB_B64_HMAX + 2 + l]) bu_s += t;
        else td_s += t;
    }
    const double rows_s = (double)(h_out[8 + XB_B64_HMAX + 1] - h_out[8 + L]) / c->wall_hz;
    // XBFS_B64_TRACE=1: per-level device times of the batch on stderr
    if (const char *tr = getenv("XBFS_B64_TRACE")) {
        if (tr[0] == '1') {
            for (int32_t l = 0; l < L; l++)
                fprintf(stderr, "[b64] level %d %s %.1f us\n", l,
                        h_out[8 + XB_B64_HMAX + 2 + l] ? "pull" : "push",
                        (double)(h_out[8 + l + 1] - h_out[8 + l]) / c->wall_hz * 1e6);
            fprintf(stderr, "[b64] rows pass 1 %.1f us, pass 2 %.1f us\n",
                    (double)(h_out[8 + XB_B64_HMAX] - h_out[8 + L]) / c->wall_hz * 1e6,
                    (double)(h_out[8 + XB_B64_HMAX + 1] - h_out[8 + XB_B64_HMAX]) / c->wall_hz * 1e6);
        }
    }
    const uint64_t unfinished = (uint64_t)h_out[1];
    int32_t max_level = h_out[2] > 0 ? L : std::max(L - 1, 0);
    long long v_cc = h_out[3], m_cc = h_out[4];
    // sources whose traversal outlasts the history window: single-source engine
    for (int32_t s = 0; s < nsrc; s++) {
        if (!((unfinished >> s) & 1ull)) continue;
        xbfs_opts so;
        xbfs_opts_default(&so);
        xbfs_stats s1;
        rc = bfs_run(c, srcs[s], nullptr, 0, &so, nullptr, nullptr, nullptr, 0, &s1);
        if (rc) return rc;
        XB_HIP(hipMemcpyAsync(c->d_b64_lv + (size_t)s * (size_t)n, c->d_level,
                              sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
        v_cc += s1.v_cc;
        m_cc += s1.m_cc;
        max_level = std::max(max_level, s1.max_level);
    }
    XB_HIP(hipStreamSynchronize(c->stream));
    if (levels_out)
        XB_HIP(hipMemcpy(levels_out, c->d_b64_lv, sizeof(int32_t) * (size_t)n * (size_t)nsrc,
                         hipMemcpyDeviceToHost));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->t_bfs_s = ms * 1e-3;
        st->n_levels = L;
        st->max_level = max_level;
        st->v_cc = v_cc;
        st->m_cc = m_cc;
        st->td_s = td_s;
        st->bu_s = bu_s;
        st->gen_s = rows_s;
    }
    return XBFS_OK;
}

int xbfs_exclusive_scan_i32(xbfs_ctx *c, const int32_t *in, int64_t *out, int64_t n)
{
    if (!c || n < 0 || n > INT32_MAX || (n > 0 && (!in || !out))) return XBFS_ERR_ARG;
    if (n == 0) return XBFS_OK;
    XB_HIP(hipSetDevice(c->device));
    const int64_t T = ceil_div(n, XB_SCAN_CHUNK);
    int32_t *d_in = nullptr;
    long long *d_out = nullptr;
    unsigned long long *d_part = nullptr;
    int rc = XBFS_OK;
    if (hipMalloc(&d_in, sizeof(int32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&d_out, sizeof(long long) * (size_t)n) != hipSuccess ||
        hipMalloc(&d_part, sizeof(unsigned long long) * (size_t)(T + 1)) != hipSuccess) {
        rc = XBFS_ERR_NOMEM;
    } else if (hipMemcpy(d_in, in, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice) !=
               hipSuccess) {
        rc = XBFS_ERR_HIP;
    } else {
        int32_t *d_err = nullptr, h_err = 0;
        if (xb_scan_i32(c->stream, d_in, n, d_out, d_part, &d_err) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(&h_err, d_err, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(out, d_out, sizeof(long long) * (size_t)n, hipMemcpyDeviceToHost) !=
                hipSuccess)
            rc = XBFS_ERR_HIP;
        else if (h_err)
            rc = XBFS_ERR_DEVICE;
    }
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    if (d_part) (void)hipFree(d_part);
    return rc;
}

int xbfs_unvisited_list(xbfs_ctx *c, const int32_t *level, int32_t *list, int64_t *count)
{
    if (!c || !level || !list || !count) return XBFS_ERR_ARG;
    XB_HIP(hipSetDevice(c->device));
    XB_HIP(hipMemcpy(c->d_level, level, sizeof(int32_t) * (size_t)c->n, hipMemcpyHostToDevice));
    int32_t *tile_cnt = c->d_tile, *tile_off = c->d_tile + c->ntiles_list;
    XbBuCountArgs bc;
    bc.level = c->d_level;
    bc.nz = c->d_nz;
    bc.n = c->n;
    bc.nwords = c->nwords;
    bc.L = INT_MIN / 2;
    bc.fb = nullptr;
    bc.nb = nullptr;
    bc.ub = c->d_ubits;
    bc.tile_cnt = tile_cnt;
    hipLaunchKernelGGL(k_bu_count, dim3((unsigned)c->ntiles_list), dim3(256), 0, c->stream, bc);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, c->stream, tile_cnt, tile_off,
                       c->ntiles_list, c->d_list_cnt);
    hipLaunchKernelGGL(k_bu_scatter, dim3((unsigned)c->ntiles_list), dim3(256), 0, c->stream,
                       c->d_ubits, c->nwords, tile_off, c->d_list);
    XB_LAUNCHED();
    XB_HIP(hipStreamSynchronize(c->stream));
    int32_t cnt = 0;
    XB_HIP(hipMemcpy(&cnt, c->d_list_cnt, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (cnt > 0)
        XB_HIP(hipMemcpy(list, c->d_list, sizeof(int32_t) * (size_t)cnt, hipMemcpyDeviceToHost));
    *count = cnt;
    return XBFS_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// per-level probe (training data and the oracle-best schedule)
// ---------------------------------------------------------------------------
static int snap_alloc(xbfs_ctx *c)
{
    if (c->s_level) return XBFS_OK;
    const size_t qb = sizeof(int32_t) * (size_t)c->n * XB_QBUFS * XB_BINS;
    if (hipMalloc(&c->s_level, sizeof(int32_t) * (size_t)c->n) != hipSuccess ||
        hipMalloc(&c->s_vis, sizeof(uint64_t) * (size_t)c->nwords) != hipSuccess ||
        hipMalloc(&c->s_fbm, sizeof(uint64_t) * (size_t)c->fstride * XB_FRINGS) != hipSuccess ||
        hipMalloc(&c->s_qbuf, qb) != hipSuccess || hipMalloc(&c->s_ctl, sizeof(XbCtl)) != hipSuccess ||
        hipMalloc(&c->s_part, sizeof(long long) * XB_NFIELDS * XB_PMAX) != hipSuccess)
        return XBFS_ERR_NOMEM;
    return XBFS_OK;
}

// save = true: live -> snapshot; false: snapshot -> live (device copies on
// the context stream)
static hipError_t snap_copy(xbfs_ctx *c, bool save)
{
    struct { void *live, *snap; size_t bytes; } parts[] = {
        {c->d_level, c->s_level, sizeof(int32_t) * (size_t)c->n},
        {c->d_vis, c->s_vis, sizeof(uint64_t) * (size_t)c->nwords},
        {c->d_fbm, c->s_fbm, sizeof(uint64_t) * (size_t)c->fstride * XB_FRINGS},
        {c->d_qbuf, c->s_qbuf, sizeof(int32_t) * (size_t)c->n * XB_QBUFS * XB_BINS},
        {c->d_ctl, c->s_ctl, sizeof(XbCtl)},
        {c->d_part, c->s_part, sizeof(long long) * XB_NFIELDS * XB_PMAX},
    };
    for (auto &pt : parts) {
        hipError_t e = hipMemcpyAsync(save ? pt.snap : pt.live, save ? pt.live : pt.snap, pt.bytes,
                                      hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static hipError_t launch_coop_level(xbfs_ctx *c, const XbArgs &a, int32_t L, int32_t mode,
                                    int32_t q_valid, int32_t need_sweep, int32_t enqueue,
                                    int32_t extra_sweep)
{
    XbArgs args = a;
    void *kargs[] = {&args, &L, &mode, &q_valid, &need_sweep, &enqueue, &extra_sweep};
    return hipLaunchCooperativeKernel((const void *)k_coop_level, dim3(c->coop_blocks),
                                      dim3(XB_PERSIST_BLOCK), kargs, XB_LDS_PERSIST, c->stream);
}

extern "C" int xbfs_bfs_probe(xbfs_ctx *c, int32_t src, const xbfs_opts *opts, int32_t follow_best,
                              int32_t reps, xbfs_probe_rec *rec, int32_t rec_cap, xbfs_stats *st)
{
    if (!c || src < 0 || (int64_t)src >= c->n || reps < 1 || reps > 64) return XBFS_ERR_ARG;
    if (rec_cap < 0 || (rec_cap > 0 && !rec)) return XBFS_ERR_ARG;
    xbfs_opts o;
    int rc = xb_resolve_opts(opts, &o);
    if (rc) return rc;
    o.want_parent = 0;
    XB_HIP(hipSetDevice(c->device));
    if ((rc = snap_alloc(c)) != XBFS_OK) return rc;
    if (ensure_events(c, 2) != XBFS_OK) return XBFS_ERR_HIP;
    if (o.policy == XBFS_POLICY_SCHEDULE)
        XB_HIP(hipMemcpyAsync(c->d_sched, o.schedule, (size_t)o.schedule_len,
                              hipMemcpyHostToDevice, c->stream));
    const XbArgs a = make_args(c, src, o);
    const int grid = o.grid_dim > 0 ? o.grid_dim : c->num_cu * 8;
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, c->stream, c->d_ctl, wait_ticks(c, o), 0ull);
    hipLaunchKernelGGL(k_ph_init, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a);
    if (a.nsrc > 0) hipLaunchKernelGGL(k_ph_seed, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a);
    XB_LAUNCHED();
    XbLoop s;
    xb_loop_init(&s);
    if (a.nsrc > 0) s.q_valid = 0; // the level-0 queue comes from a sweep of F_0
    long long pend_n = 0, pend_m = 0, n_f = 0;
    double t_total = 0;
    std::vector<float> tr((size_t)reps);
    for (;;) {
        const int32_t L = s.L;
        if (s.need_sweep)
            hipLaunchKernelGGL(k_ph_sweep, dim3(grid), dim3(XB_PBLOCK), XB_LDS_PHASE, c->stream, a,
                               L, 1, 1);
        hipLaunchKernelGGL(k_part_reduce, dim3(1), dim3(1024), 0, c->stream, a);
        XB_HIP(hipMemcpyAsync(c->h_ctl, c->d_ctl, sizeof(XbCtl), hipMemcpyDeviceToHost, c->stream));
        XB_HIP(hipStreamSynchronize(c->stream));
        const XbCtl &h = *c->h_ctl;
        if (h.error) return XBFS_ERR_DEVICE;
        long long m_f;
        if (s.need_sweep) {
            n_f = h.sum[xb_f_sweep(L, 0)];
            m_f = h.sum[xb_f_sweep(L, 1)];
        } else {
            n_f = h.sum[xb_f_ring(L, 0)];
            m_f = h.sum[xb_f_ring(L, 1)];
        }
        pend_n = h.sum[xb_f_ring(L + 1, 0)];
        pend_m = h.sum[xb_f_ring(L + 1, 1)];
        const long long p2_next = h.sum[xb_f_ring(L + 1, 2)];
        long long n_u = 0, m_u = 0;
        if (!xb_level_open(s, a, n_f, m_f, pend_n, pend_m, &n_u, &m_u)) break;
        if (L >= a.max_levels) return XBFS_ERR_LIMIT;
        // time every mode from the same state
        double tm[3];
        XB_HIP(snap_copy(c, true));
        for (int m = 0; m < 3; m++) {
            for (int r = 0; r < reps; r++) {
                XB_HIP(hipEventRecord(c->ev_pool[0], c->stream));
                XB_HIP(launch_coop_level(c, a, L, m, s.q_valid, s.need_sweep, p2_next == 0,
                                         m == XBFS_MODE_TD_SS));
                XB_HIP(hipEventRecord(c->ev_pool[1], c->stream));
                XB_HIP(hipEventSynchronize(c->ev_pool[1]));
                XB_HIP(hipEventElapsedTime(&tr[(size_t)r], c->ev_pool[0], c->ev_pool[1]));
                XB_HIP(snap_copy(c, false));
            }
            std::sort(tr.begin(), tr.end());
            tm[m] = tr[(size_t)reps / 2] * 1e-3;
        }
        float y[XB_MLP_NO];
        const float *pred = nullptr;
        if (a.mlp && L <= XBFS_MLP_MAX_LEVEL) {
            float x[XB_MLP_NI];
            xbfs_mlp_features(L, n_f, m_f, n_u, m_u, s.prev_mode, x);
            xbfs_mlp_forward1(c->h_mlp.data(), x, y);
            pred = y;
        }
        int32_t htag = 0, used = 0;
        int32_t mode = xbfs_choose_mode(a.policy, o.schedule, a.sched_len, pred, &s.hs, L, m_f, a.m,
                                        a.alpha, &htag, &used);
        if (follow_best) {
            mode = 0;
            for (int m = 1; m < 3; m++)
                if (tm[m] < tm[mode]) mode = m;
        }
        XB_HIP(launch_coop_level(c, a, L, mode, s.q_valid, s.need_sweep, p2_next == 0, 0));
        t_total += tm[mode];
        if (rec && L < rec_cap) {
            xbfs_probe_rec &r = rec[L];
            r.level = L;
            r.prev_mode = s.prev_mode;
            r.mode = mode;
            r.htag = htag;
            r.n_f = n_f;
            r.m_f = m_f;
            r.n_u = n_u;
            r.m_u = m_u;
            for (int m = 0; m < 3; m++) r.t_mode[m] = tm[m];
        }
        s.q_valid = (mode == XBFS_MODE_TD_SF && p2_next == 0);
        s.need_sweep = (mode == XBFS_MODE_TD_SS);
        s.prev_mode = mode;
        s.L = L + 1;
    }
    XB_HIP(hipStreamSynchronize(c->stream));
    if (st) {
        const int32_t L = s.L;
        st->t_bfs_s = t_total; // sum of the executed modes' probed times
        st->n_levels = L;
        st->max_level = pend_n > 0 ? L + 1 : (n_f > 0 ? L : L - 1);
        st->v_cc = s.cum_n + pend_n + c->h_ctl->seed_iso;
        st->m_cc = s.cum_m + pend_m;
        st->td_s = st->bu_s = st->gen_s = 0;
    }
    return XBFS_OK;
}

// ---------------------------------------------------------------------------
// learned policy: features, weights, MLP kernels
// ---------------------------------------------------------------------------
extern "C" int xbfs_policy_features(int32_t level, int64_t n_f, int64_t m_f, int64_t n_u,
                                    int64_t m_u, int32_t prev_mode, float *x8)
{
    if (!x8 || level < 0) return XBFS_ERR_ARG;
    xbfs_mlp_features(level, n_f, m_f, n_u, m_u, prev_mode, x8);
    return XBFS_OK;
}

extern "C" int xbfs_policy_load(xbfs_ctx *c, const float *w, int32_t nw)
{
    if (!c || !w || nw != XB_MLP_NPARAM) return XBFS_ERR_ARG;
    for (int i = 0; i < nw; i++)
        if (!(w[i] > -1e30f && w[i] < 1e30f)) return XBFS_ERR_ARG;
    XB_HIP(hipSetDevice(c->device));
    XB_HIP(hipMemcpy(c->d_mlp, w, sizeof(float) * XB_MLP_NPARAM, hipMemcpyHostToDevice));
    c->h_mlp.assign(w, w + XB_MLP_NPARAM);
    c->mlp_loaded = 1;
    return XBFS_OK;
}

namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t bytes) { return hipMalloc(&p, bytes > 0 ? bytes : 1) == hipSuccess; }
    template <typename T> T *as() { return (T *)p; }
};
constexpr int kMlpGradBlocks = 64;
}

static int mlp_device(int device)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return XBFS_ERR_HIP;
    if (device < 0 || device >= ndev) return XBFS_ERR_ARG;
    return hipSetDevice(device) == hipSuccess ? XBFS_OK : XBFS_ERR_HIP;
}

extern "C" int xbfs_mlp_forward(int device, const float *params, const float *x, int32_t n,
                                float *out)
{
    if (!params || n < 0 || (n > 0 && (!x || !out))) return XBFS_ERR_ARG;
    int rc = mlp_device(device);
    if (rc) return rc;
    if (n == 0) return XBFS_OK;
    DevBuf dp, dx, dy;
    if (!dp.alloc(sizeof(float) * XB_MLP_NPARAM) || !dx.alloc(sizeof(float) * (size_t)n * 8) ||
        !dy.alloc(sizeof(float) * (size_t)n * 3))
        return XBFS_ERR_NOMEM;
    XB_HIP(hipMemcpy(dp.p, params, sizeof(float) * XB_MLP_NPARAM, hipMemcpyHostToDevice));
    XB_HIP(hipMemcpy(dx.p, x, sizeof(float) * (size_t)n * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n, XB_MLP_BLOCK, 1024)), dim3(XB_MLP_BLOCK), 0,
                       0, dp.as<float>(), dx.as<float>(), n, dy.as<float>());
    XB_LAUNCHED();
    XB_HIP(hipDeviceSynchronize());
    XB_HIP(hipMemcpy(out, dy.p, sizeof(float) * (size_t)n * 3, hipMemcpyDeviceToHost));
    return XBFS_OK;
}

struct MlpData {
    DevBuf p, x, y, mk, pg, pl, g, loss;
};

static int mlp_upload(MlpData &d, const float *params, const float *x, const float *y,
                      const float *mask, int32_t n)
{
    if (!d.p.alloc(sizeof(float) * XB_MLP_NPARAM) || !d.x.alloc(sizeof(float) * (size_t)n * 8) ||
        !d.y.alloc(sizeof(float) * (size_t)n * 3) || !d.mk.alloc(sizeof(float) * (size_t)n * 3) ||
        !d.pg.alloc(sizeof(float) * XB_MLP_NPARAM * kMlpGradBlocks) ||
        !d.pl.alloc(sizeof(float) * kMlpGradBlocks) || !d.g.alloc(sizeof(float) * XB_MLP_NPARAM) ||
        !d.loss.alloc(sizeof(float)))
        return XBFS_ERR_NOMEM;
    if (hipMemcpy(d.p.p, params, sizeof(float) * XB_MLP_NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.x.p, x, sizeof(float) * (size_t)n * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.y.p, y, sizeof(float) * (size_t)n * 3, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.mk.p, mask, sizeof(float) * (size_t)n * 3, hipMemcpyHostToDevice) != hipSuccess)
        return XBFS_ERR_HIP;
    return XBFS_OK;
}

static void mlp_enqueue_grad(MlpData &d, int32_t n)
{
    hipLaunchKernelGGL(k_mlp_grad, dim3(kMlpGradBlocks), dim3(XB_MLP_BLOCK), 0, 0, d.p.as<float>(),
                       d.x.as<float>(), d.y.as<float>(), d.mk.as<float>(), n, d.pg.as<float>(),
                       d.pl.as<float>());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((XB_MLP_NPARAM + 255) / 256), dim3(256), 0, 0,
                       d.pg.as<float>(), d.pl.as<float>(), kMlpGradBlocks, d.g.as<float>(),
                       d.loss.as<float>());
}

extern "C" int xbfs_mlp_loss_grad(int device, const float *params, const float *x, const float *y,
                                  const float *mask, int32_t n, float *grad, float *loss)
{
    if (!params || !x || !y || !mask || n < 1 || !grad || !loss) return XBFS_ERR_ARG;
    int rc = mlp_device(device);
    if (rc) return rc;
    MlpData d;
    if ((rc = mlp_upload(d, params, x, y, mask, n)) != XBFS_OK) return rc;
    mlp_enqueue_grad(d, n);
    XB_LAUNCHED();
    XB_HIP(hipDeviceSynchronize());
    XB_HIP(hipMemcpy(grad, d.g.p, sizeof(float) * XB_MLP_NPARAM, hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(loss, d.loss.p, sizeof(float), hipMemcpyDeviceToHost));
    return XBFS_OK;
}

extern "C" int xbfs_mlp_train(int device, float *params, const float *x, const float *y,
                              const float *mask, int32_t n, int32_t steps, float lr,
                              float *loss_curve)
{
    if (!params || !x || !y || !mask || n < 1 || steps < 1 || !(lr > 0.0f)) return XBFS_ERR_ARG;
    int rc = mlp_device(device);
    if (rc) return rc;
    MlpData d;
    if ((rc = mlp_upload(d, params, x, y, mask, n)) != XBFS_OK) return rc;
    DevBuf m, v, curve;
    if (!m.alloc(sizeof(float) * XB_MLP_NPARAM) || !v.alloc(sizeof(float) * XB_MLP_NPARAM) ||
        !curve.alloc(sizeof(float) * (size_t)steps))
        return XBFS_ERR_NOMEM;
    XB_HIP(hipMemset(m.p, 0, sizeof(float) * XB_MLP_NPARAM));
    XB_HIP(hipMemset(v.p, 0, sizeof(float) * XB_MLP_NPARAM));
    for (int t = 1; t <= steps; t++) {
        mlp_enqueue_grad(d, n);
        hipLaunchKernelGGL(k_mlp_adam, dim3((XB_MLP_NPARAM + 255) / 256), dim3(256), 0, 0,
                           d.p.as<float>(), d.g.as<float>(), m.as<float>(), v.as<float>(), t, lr,
                           0.9f, 0.999f, 1e-8f, d.loss.as<float>(), curve.as<float>());
    }
    XB_LAUNCHED();
    XB_HIP(hipDeviceSynchronize());
    XB_HIP(hipMemcpy(params, d.p.p, sizeof(float) * XB_MLP_NPARAM, hipMemcpyDeviceToHost));
    if (loss_curve)
        XB_HIP(hipMemcpy(loss_curve, curve.p, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost));
    return XBFS_OK;
}

// the allocation-free single-source entry point (its scratch array and set-up)
#include "xbfs_noalloc.hip"

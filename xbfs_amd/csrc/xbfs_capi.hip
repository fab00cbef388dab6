// C ABI of libxbfs: device context, the host-driven adaptive level loop
// (BFS_ControlFlow, bfs_adaptiveFQ_wb_async.cuh:934-1258, without the
// per-level O(V) host work) and the test hooks for the scan / list kernels.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstring>
#include <new>
#include <vector>

#include "xbfs.h"
#include "xbfs_policy.h"
#include "xbfs_host_internal.h"
#include "xbfs_ctx.h"
#include "xbfs_kernels.hip"

#define XB_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return XBFS_ERR_HIP;                                     \
    } while (0)
#define XB_LAUNCHED() XB_HIP(hipGetLastError())

namespace {

enum { FS_QUEUE = 0, FS_LIST = 1, FS_SWEEP = 2 };

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

int xb_ctx_free_device(xbfs_ctx *c)
{
    (void)hipSetDevice(c->device);
    void *bufs[] = {c->d_beg, c->d_adj, c->d_nz, c->d_level, c->d_parent, c->d_qbuf,
                    c->d_list, c->d_bits, c->d_tile, c->d_ctl};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    return XBFS_OK;
}

extern "C" {

void xbfs_opts_default(xbfs_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->block_dim = 256;
    o->grid_dim = 0;
    o->alpha = 0.1f; // run.sh:36
    o->beta = 0.1f;  // wb:1018 (unused)
    o->padding = 1;
    o->policy = XBFS_POLICY_HEURISTIC;
    o->persistent = 0;
    o->want_parent = 0;
    o->bin_thread_max = XB_DEFAULT_T1;
    o->bin_wave_max = XB_DEFAULT_T2;
    o->schedule = nullptr;
    o->schedule_len = 0;
    o->time_levels = 0;
}

int64_t xbfs_num_vertices(const xbfs_ctx *c) { return c ? c->n : -1; }
int64_t xbfs_num_edges(const xbfs_ctx *c) { return c ? c->m : -1; }

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
    c->ntiles = ceil_div(c->nwords, XB_TILE_WORDS);
    int64_t nz = 0;
    for (int64_t i = 0; i < n; i++) nz += beg[i + 1] > beg[i];
    c->n_nz = nz;
    auto fail = [&](int code) {
        xb_ctx_free_device(c);
        delete c;
        return code;
    };
    if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device) !=
        hipSuccess)
        return fail(XBFS_ERR_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(XBFS_ERR_HIP);
    const size_t mm = (size_t)(m > 0 ? m : 1);
    bool ok = hipMalloc(&c->d_beg, sizeof(int64_t) * (size_t)(n + 1)) == hipSuccess &&
              hipMalloc(&c->d_adj, sizeof(int32_t) * mm) == hipSuccess &&
              hipMalloc(&c->d_nz, sizeof(uint64_t) * (size_t)c->nwords) == hipSuccess &&
              hipMalloc(&c->d_level, sizeof(int32_t) * (size_t)n) == hipSuccess &&
              hipMalloc(&c->d_parent, sizeof(int32_t) * (size_t)n) == hipSuccess &&
              hipMalloc(&c->d_qbuf, sizeof(int32_t) * (size_t)n * 6) == hipSuccess &&
              hipMalloc(&c->d_list, sizeof(int32_t) * (size_t)n) == hipSuccess &&
              hipMalloc(&c->d_bits, sizeof(uint64_t) * (size_t)c->nwords * 3) == hipSuccess &&
              hipMalloc(&c->d_tile, sizeof(int32_t) * (size_t)c->ntiles * 2) == hipSuccess &&
              hipMalloc(&c->d_ctl, sizeof(XbDevCtl)) == hipSuccess &&
              hipHostMalloc(&c->h_ctl, sizeof(XbDevCtl), hipHostMallocDefault) == hipSuccess;
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
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(XBFS_ERR_HIP);
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
    if (o->block_dim == 0) o->block_dim = 256;
    if (!pow2_block(o->block_dim)) return XBFS_ERR_ARG;
    if (o->grid_dim < 0) return XBFS_ERR_ARG;
    if (o->bin_thread_max <= 0) o->bin_thread_max = XB_DEFAULT_T1;
    if (o->bin_wave_max <= 0) o->bin_wave_max = XB_DEFAULT_T2;
    if (o->bin_wave_max < o->bin_thread_max) return XBFS_ERR_ARG;
    if (o->policy < XBFS_POLICY_HEURISTIC || o->policy > XBFS_POLICY_SCHEDULE) return XBFS_ERR_ARG;
    if (o->policy == XBFS_POLICY_SCHEDULE) {
        if (!o->schedule || o->schedule_len <= 0) return XBFS_ERR_ARG;
        for (int i = 0; i < o->schedule_len; i++)
            if (o->schedule[i] < 0 || o->schedule[i] > 2) return XBFS_ERR_ARG;
    }
    if (!(o->alpha >= 0.0f)) return XBFS_ERR_ARG;
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

// Host-driven level loop: one small D2H copy of the control block per level;
// everything else (frontier statistics, queues, bitmaps, scans) stays on the
// device.  Frontier source per level: FS_QUEUE (exact binned queue),
// FS_LIST (last bottom-up list, filtered on level==L), FS_SWEEP (full sweep).
static int bfs_host_loop(xbfs_ctx *c, int32_t src, const xbfs_opts &o, xbfs_level_rec *rec,
                         int32_t rec_cap, xbfs_stats *st)
{
    const bool par = o.want_parent != 0;
    const int bd = o.block_dim;
    const int grid = o.grid_dim > 0 ? o.grid_dim : c->num_cu * (2048 / bd);
    const int32_t t1 = o.bin_thread_max, t2 = o.bin_wave_max;
    const int64_t n = c->n;
    XbDevCtl *ctl = c->d_ctl;
    if (o.time_levels && ensure_events(c, 8) != XBFS_OK) return XBFS_ERR_HIP;

    XB_HIP(hipEventRecord(c->ev_start, c->stream));
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n / 4 + 1, 256, 8192)), dim3(256), 0, c->stream,
                       c->d_level, n, -1);
    if (par)
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n / 4 + 1, 256, 8192)), dim3(256), 0,
                           c->stream, c->d_parent, n, -1);
    hipLaunchKernelGGL(k_seed, dim3(1), dim3(64), 0, c->stream, c->d_level,
                       par ? c->d_parent : (int32_t *)nullptr, c->d_beg, ctl, c->q(0, 0),
                       c->q(0, 1), c->q(0, 2), src, t1, t2);
    XB_LAUNCHED();

    int fsrc = FS_QUEUE, cur = 0;
    XbfsHeurState hs;
    xbfs_heur_init(&hs);
    int64_t cum_n = 0, cum_m = 0, pend_n = 0, pend_m = 0;
    int32_t L = 0, max_level = 0;
    double td_s = 0, bu_s = 0, gen_s = 0;
    const int sweep_grid = grid_for(n, bd, c->num_cu * 16);
    for (;; L++) {
        float gen_ms = 0.0f;
        if (fsrc == FS_SWEEP) {
            if (o.time_levels) XB_HIP(hipEventRecord(c->ev_pool[0], c->stream));
            hipLaunchKernelGGL(k_prep, dim3(1), dim3(64), 0, c->stream, ctl, L & 3, -1, cur, -1, 0);
            XbGenArgs g;
            g.level = c->d_level;
            g.beg = c->d_beg;
            g.n = n;
            g.L = L;
            g.t1 = t1;
            g.t2 = t2;
            g.q0 = c->q(cur, 0);
            g.q1 = c->q(cur, 1);
            g.q2 = c->q(cur, 2);
            g.cnt = &ctl->qcnt[cur][0];
            g.slot_n = &ctl->slot_n[L & 3];
            g.slot_m = &ctl->slot_m[L & 3];
            hipLaunchKernelGGL(k_sweep_gen, dim3(sweep_grid), dim3(bd), 0, c->stream, g);
            XB_LAUNCHED();
            if (o.time_levels) XB_HIP(hipEventRecord(c->ev_pool[1], c->stream));
            fsrc = FS_QUEUE;
        }
        XB_HIP(hipMemcpyAsync(c->h_ctl, ctl, sizeof(XbDevCtl), hipMemcpyDeviceToHost, c->stream));
        XB_HIP(hipStreamSynchronize(c->stream));
        if (o.time_levels && gen_ms == 0.0f && c->ev_pool.size() >= 2) {
            // only meaningful when a sweep ran this level
        }
        const XbDevCtl h = *c->h_ctl;
        const int64_t n_f = (int64_t)h.slot_n[L & 3], m_f = (int64_t)h.slot_m[L & 3];
        pend_n = (int64_t)h.slot_n[(L + 1) & 3];
        pend_m = (int64_t)h.slot_m[(L + 1) & 3];
        if (n_f > 0 && L > max_level) max_level = L;
        if (pend_n > 0 && L + 1 > max_level) max_level = L + 1;
        cum_n += n_f;
        cum_m += m_f;
        if (m_f == 0 || cum_n + pend_n >= c->n_nz) break;
        if (L > n) return XBFS_ERR_LIMIT; // a BFS has at most n levels

        const int64_t n_u = c->n_nz - (cum_n + pend_n);
        const int64_t m_u = c->m - (cum_m + pend_m);
        int32_t sub = 0;
        const int32_t htag = xbfs_heur_decide(&hs, L, m_f, c->m, o.alpha, &sub);
        int32_t mode;
        switch (o.policy) {
        case XBFS_POLICY_FORCE_TD_SF: mode = XBFS_MODE_TD_SF; break;
        case XBFS_POLICY_FORCE_TD_SS: mode = XBFS_MODE_TD_SS; break;
        case XBFS_POLICY_FORCE_BU: mode = XBFS_MODE_BU; break;
        case XBFS_POLICY_SCHEDULE:
            mode = o.schedule[L < o.schedule_len ? L : o.schedule_len - 1];
            break;
        default: mode = xbfs_tag_to_mode(htag); break;
        }
        if (o.time_levels) {
            XB_HIP(hipEventRecord(c->ev_pool[2], c->stream));
        }
        // slot L+2 and the next queue's counters start at zero this level
        hipLaunchKernelGGL(k_prep, dim3(1), dim3(64), 0, c->stream, ctl, (L + 2) & 3, -1, cur ^ 1,
                           (fsrc == FS_LIST && mode != XBFS_MODE_BU) ? cur : -1, 0);
        if (mode != XBFS_MODE_BU && fsrc == FS_LIST) {
            XbFilterArgs f;
            f.list = c->d_list;
            f.list_cnt = &ctl->list_cnt;
            f.level = c->d_level;
            f.beg = c->d_beg;
            f.L = L;
            f.t1 = t1;
            f.t2 = t2;
            f.carry = mode == XBFS_MODE_TD_SF;
            f.qc0 = c->q(cur, 0);
            f.qc1 = c->q(cur, 1);
            f.qc2 = c->q(cur, 2);
            f.cnt_c = &ctl->qcnt[cur][0];
            f.qn0 = c->q(cur ^ 1, 0);
            f.qn1 = c->q(cur ^ 1, 1);
            f.qn2 = c->q(cur ^ 1, 2);
            f.cnt_n = &ctl->qcnt[cur ^ 1][0];
            hipLaunchKernelGGL(k_list_filter, dim3(grid), dim3(bd), 0, c->stream, f);
        }
        if (mode == XBFS_MODE_TD_SF || mode == XBFS_MODE_TD_SS) {
            XbTdArgs a;
            a.level = c->d_level;
            a.parent = c->d_parent;
            a.beg = c->d_beg;
            a.adj = c->d_adj;
            a.qc0 = c->q(cur, 0);
            a.qc1 = c->q(cur, 1);
            a.qc2 = c->q(cur, 2);
            a.cnt_c = &ctl->qcnt[cur][0];
            a.qn0 = c->q(cur ^ 1, 0);
            a.qn1 = c->q(cur ^ 1, 1);
            a.qn2 = c->q(cur ^ 1, 2);
            a.cnt_n = &ctl->qcnt[cur ^ 1][0];
            a.slot_n1 = &ctl->slot_n[(L + 1) & 3];
            a.slot_m1 = &ctl->slot_m[(L + 1) & 3];
            a.L = L;
            a.t1 = t1;
            a.t2 = t2;
            if (mode == XBFS_MODE_TD_SF) {
                if (par) hipLaunchKernelGGL((k_td<true, true>), dim3(grid), dim3(bd), 0, c->stream, a);
                else hipLaunchKernelGGL((k_td<true, false>), dim3(grid), dim3(bd), 0, c->stream, a);
                cur ^= 1;
                fsrc = FS_QUEUE;
            } else {
                if (par) hipLaunchKernelGGL((k_td<false, true>), dim3(grid), dim3(bd), 0, c->stream, a);
                else hipLaunchKernelGGL((k_td<false, false>), dim3(grid), dim3(bd), 0, c->stream, a);
                fsrc = FS_SWEEP;
            }
        } else {
            uint64_t *fb = c->d_bits, *nbm = c->d_bits + c->nwords, *ub = c->d_bits + 2 * c->nwords;
            int32_t *tile_cnt = c->d_tile, *tile_off = c->d_tile + c->ntiles;
            XbBuCountArgs bc;
            bc.level = c->d_level;
            bc.nz = c->d_nz;
            bc.n = n;
            bc.nwords = c->nwords;
            bc.L = L;
            bc.fb = fb;
            bc.nb = nbm;
            bc.ub = ub;
            bc.tile_cnt = tile_cnt;
            hipLaunchKernelGGL(k_bu_count, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream, bc);
            hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, c->stream, tile_cnt, tile_off,
                               c->ntiles, &ctl->list_cnt);
            hipLaunchKernelGGL(k_bu_scatter, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream, ub,
                               c->nwords, tile_off, c->d_list);
            XbBuArgs b;
            b.level = c->d_level;
            b.parent = c->d_parent;
            b.beg = c->d_beg;
            b.adj = c->d_adj;
            b.list = c->d_list;
            b.list_cnt = &ctl->list_cnt;
            b.fb = fb;
            b.nb = nbm;
            b.slot_n1 = &ctl->slot_n[(L + 1) & 3];
            b.slot_m1 = &ctl->slot_m[(L + 1) & 3];
            b.slot_n2 = &ctl->slot_n[(L + 2) & 3];
            b.slot_m2 = &ctl->slot_m[(L + 2) & 3];
            b.L = L;
            if (par) hipLaunchKernelGGL((k_bu<true>), dim3(grid), dim3(bd), 0, c->stream, b);
            else hipLaunchKernelGGL((k_bu<false>), dim3(grid), dim3(bd), 0, c->stream, b);
            fsrc = FS_LIST;
        }
        XB_LAUNCHED();
        double lvl_s = 0.0;
        if (o.time_levels) {
            XB_HIP(hipEventRecord(c->ev_pool[3], c->stream));
            XB_HIP(hipEventSynchronize(c->ev_pool[3]));
            float ms = 0.0f;
            XB_HIP(hipEventElapsedTime(&ms, c->ev_pool[2], c->ev_pool[3]));
            lvl_s = ms * 1e-3;
        }
        if (mode == XBFS_MODE_BU) bu_s += lvl_s;
        else td_s += lvl_s;
        (void)gen_ms;
        if (rec && L < rec_cap) {
            xbfs_level_rec &r = rec[L];
            r.level = L;
            r.mode = mode;
            r.tag = xbfs_mode_to_tag(mode);
            r.policy = o.policy;
            r.n_f = n_f;
            r.m_f = m_f;
            r.n_u = n_u;
            r.m_u = m_u;
            r.ratio = (float)m_f / (float)c->m;
            r.pad = (float)htag;
            r.t_s = lvl_s;
        }
    }
    XB_HIP(hipEventRecord(c->ev_stop, c->stream));
    XB_HIP(hipEventSynchronize(c->ev_stop));
    float ms = 0.0f;
    XB_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
    if (st) {
        st->t_bfs_s = ms * 1e-3;
        st->n_levels = L;
        st->max_level = max_level;
        st->v_cc = cum_n + pend_n;
        st->m_cc = cum_m + pend_m;
        st->td_s = td_s;
        st->bu_s = bu_s;
        st->gen_s = gen_s;
    }
    return XBFS_OK;
}

extern "C" {

int xbfs_bfs(xbfs_ctx *c, int32_t src, const xbfs_opts *opts, int32_t *level_out,
             int32_t *parent_out, xbfs_level_rec *rec, int32_t rec_cap, xbfs_stats *st)
{
    if (!c) return XBFS_ERR_ARG;
    if (src < 0 || (int64_t)src >= c->n) return XBFS_ERR_ARG;
    if (rec_cap < 0 || (rec_cap > 0 && !rec)) return XBFS_ERR_ARG;
    xbfs_opts o;
    int rc = xb_resolve_opts(opts, &o);
    if (rc) return rc;
    if (parent_out) o.want_parent = 1;
    if (o.policy == XBFS_POLICY_LEARNED) o.policy = XBFS_POLICY_HEURISTIC;
    XB_HIP(hipSetDevice(c->device));
    rc = bfs_host_loop(c, src, o, rec, rec_cap, st);
    if (rc) return rc;
    if (level_out)
        XB_HIP(hipMemcpy(level_out, c->d_level, sizeof(int32_t) * (size_t)c->n,
                         hipMemcpyDeviceToHost));
    if (parent_out)
        XB_HIP(hipMemcpy(parent_out, c->d_parent, sizeof(int32_t) * (size_t)c->n,
                         hipMemcpyDeviceToHost));
    return XBFS_OK;
}

int xbfs_exclusive_scan_i32(xbfs_ctx *c, const int32_t *in, int64_t *out, int64_t n)
{
    if (!c || n < 0 || n > INT32_MAX || (n > 0 && (!in || !out))) return XBFS_ERR_ARG;
    if (n == 0) return XBFS_OK;
    XB_HIP(hipSetDevice(c->device));
    const int64_t T = ceil_div(n, XB_SCAN_CHUNK);
    int32_t *d_in = nullptr;
    long long *d_out = nullptr, *d_part = nullptr;
    int rc = XBFS_OK;
    if (hipMalloc(&d_in, sizeof(int32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&d_out, sizeof(long long) * (size_t)n) != hipSuccess ||
        hipMalloc(&d_part, sizeof(long long) * (size_t)T) != hipSuccess) {
        rc = XBFS_ERR_NOMEM;
    } else if (hipMemcpy(d_in, in, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice) !=
               hipSuccess) {
        rc = XBFS_ERR_HIP;
    } else {
        hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)T), dim3(256), 0, c->stream, d_in, n,
                           d_part);
        hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(1024), 0, c->stream, d_part, T);
        hipLaunchKernelGGL(k_scan_down, dim3((unsigned)T), dim3(256), 0, c->stream, d_in, n,
                           d_part, d_out);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(out, d_out, sizeof(long long) * (size_t)n, hipMemcpyDeviceToHost) !=
                hipSuccess)
            rc = XBFS_ERR_HIP;
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
    uint64_t *ub = c->d_bits + 2 * c->nwords;
    int32_t *tile_cnt = c->d_tile, *tile_off = c->d_tile + c->ntiles;
    XbBuCountArgs bc;
    bc.level = c->d_level;
    bc.nz = c->d_nz;
    bc.n = c->n;
    bc.nwords = c->nwords;
    bc.L = INT_MIN / 2;
    bc.fb = nullptr;
    bc.nb = nullptr;
    bc.ub = ub;
    bc.tile_cnt = tile_cnt;
    hipLaunchKernelGGL(k_bu_count, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream, bc);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, c->stream, tile_cnt, tile_off,
                       c->ntiles, &c->d_ctl->list_cnt);
    hipLaunchKernelGGL(k_bu_scatter, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream, ub,
                       c->nwords, tile_off, c->d_list);
    XB_LAUNCHED();
    XB_HIP(hipStreamSynchronize(c->stream));
    int32_t cnt = 0;
    XB_HIP(hipMemcpy(&cnt, &c->d_ctl->list_cnt, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (cnt > 0)
        XB_HIP(hipMemcpy(list, c->d_list, sizeof(int32_t) * (size_t)cnt, hipMemcpyDeviceToHost));
    *count = cnt;
    return XBFS_OK;
}

} // extern "C"

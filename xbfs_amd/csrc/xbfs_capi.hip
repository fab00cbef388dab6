// C ABI of libxbfs: device context, the two drivers of the adaptive level loop
// (BFS_ControlFlow, bfs_adaptiveFQ_wb_async.cuh:934-1258, with the decision
// made from device-side frontier statistics instead of O(V) host scans), and
// the test hooks for the scan / list kernels.
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
#include "xbfs_engine.hip"

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

int xb_ctx_free_device(xbfs_ctx *c)
{
    (void)hipSetDevice(c->device);
    void *bufs[] = {c->d_beg,   c->d_adj,  c->d_nz,    c->d_vis,  c->d_fbm,   c->d_level,
                    c->d_parent, c->d_qbuf, c->d_ctl,  c->d_recs, c->d_sched, c->d_mlp,
                    c->d_list,  c->d_ubits, c->d_tile, c->d_list_cnt};
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
    c->ntiles_bu = ceil_div(c->nwords, XB_BU_TILE_WORDS);
    c->fstride = c->ntiles_bu * XB_BU_TILE_WORDS;
    c->ntiles_list = ceil_div(c->nwords, XB_TILE_WORDS);
    int64_t nz = 0;
    for (int64_t i = 0; i < n; i++) nz += beg[i + 1] > beg[i];
    c->n_nz = nz;
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
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persistent, XB_PERSIST_BLOCK, 0) !=
            hipSuccess ||
        per_cu < 1)
        return fail(XBFS_ERR_HIP);
    c->coop_blocks = per_cu * c->num_cu;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(XBFS_ERR_HIP);
    const size_t mm = (size_t)(m > 0 ? m : 1);
    const size_t nw = (size_t)c->nwords;
    bool ok = hipMalloc(&c->d_beg, sizeof(int64_t) * (size_t)(n + 1)) == hipSuccess &&
              hipMalloc(&c->d_adj, sizeof(int32_t) * mm) == hipSuccess &&
              hipMalloc(&c->d_nz, sizeof(uint64_t) * nw) == hipSuccess &&
              hipMalloc(&c->d_vis, sizeof(uint64_t) * nw) == hipSuccess &&
              hipMalloc(&c->d_fbm, sizeof(uint64_t) * (size_t)c->fstride * XB_FRINGS) == hipSuccess &&
              hipMalloc(&c->d_level, sizeof(int32_t) * (size_t)n) == hipSuccess &&
              hipMalloc(&c->d_parent, sizeof(int32_t) * (size_t)n) == hipSuccess &&
              hipMalloc(&c->d_qbuf, sizeof(int32_t) * (size_t)n * XB_QBUFS * XB_BINS) ==
                  hipSuccess &&
              hipMalloc(&c->d_ctl, sizeof(XbCtl)) == hipSuccess &&
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
    if (o->block_dim == 0) o->block_dim = XB_PBLOCK;
    if (!pow2_block(o->block_dim)) return XBFS_ERR_ARG;
    if (o->grid_dim < 0) return XBFS_ERR_ARG;
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
    a.nz = c->d_nz;
    a.vis = c->d_vis;
    a.fbm = c->d_fbm;
    a.qbuf = c->d_qbuf;
    a.ctl = c->d_ctl;
    a.recs = c->d_recs;
    a.rec_cap = XB_REC_CAP;
    a.sched = c->d_sched;
    a.sched_len = o.policy == XBFS_POLICY_SCHEDULE ? o.schedule_len : 0;
    a.mlp = (o.policy == XBFS_POLICY_LEARNED && c->mlp_loaded) ? c->d_mlp : nullptr;
    a.n = c->n;
    a.m = c->m;
    a.nwords = c->nwords;
    a.ntiles = c->ntiles_bu;
    a.fstride = c->fstride;
    a.n_nz = c->n_nz;
    a.src = src;
    a.t1 = o.bin_thread_max;
    a.t2 = o.bin_wave_max;
    a.t3 = o.bin_block_max;
    a.policy = o.policy;
    a.alpha = o.alpha;
    a.max_levels = (int32_t)std::min<int64_t>(c->n + 1, INT32_MAX - 4);
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
    XB_LAUNCHED();
    XbLoop s;
    xb_loop_init(&s);
    double td_s = 0, bu_s = 0, gen_s = 0;
    long long pend_n = 0, pend_m = 0, n_f = 0;
    for (;;) {
        const int32_t L = s.L;
        double sweep_s = 0;
        if (s.need_sweep) {
            if (o.time_levels) XB_HIP(hipEventRecord(c->ev_pool[0], c->stream));
            hipLaunchKernelGGL(k_ph_compact, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L, 1);
            XB_LAUNCHED();
            if (o.time_levels) XB_HIP(hipEventRecord(c->ev_pool[1], c->stream));
        }
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
            n_f = (long long)h.sw_n[L & 1];
            m_f = (long long)h.sw_m[L & 1];
        } else {
            n_f = (long long)h.slot_n[L & 3];
            m_f = (long long)h.slot_m[L & 3];
        }
        pend_n = (long long)h.slot_n[(L + 1) & 3];
        pend_m = (long long)h.slot_m[(L + 1) & 3];
        const long long p2_next = (long long)h.slot_p2[(L + 1) & 3];
        long long n_u = 0, m_u = 0;
        if (!xb_level_open(s, a, n_f, m_f, pend_n, pend_m, &n_u, &m_u)) break;
        if (L >= a.max_levels) return XBFS_ERR_LIMIT;
        float y[XB_MLP_NO];
        const float *pred = nullptr;
        if (a.mlp) {
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
        hipLaunchKernelGGL(k_ph_prep, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L);
        if (mode != XBFS_MODE_BU && !s.q_valid && !s.need_sweep)
            hipLaunchKernelGGL(k_ph_compact, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L, 0);
        if (mode == XBFS_MODE_TD_SF)
            hipLaunchKernelGGL(k_ph_td_sf, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L,
                               (int32_t)(p2_next == 0));
        else if (mode == XBFS_MODE_TD_SS)
            hipLaunchKernelGGL(k_ph_td_ss, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L);
        else
            hipLaunchKernelGGL(k_ph_bu, dim3(grid), dim3(XB_PBLOCK), 0, c->stream, a, L);
        XB_LAUNCHED();
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
        st->v_cc = s.cum_n + pend_n;
        st->m_cc = s.cum_m + pend_m;
        st->td_s = td_s;
        st->bu_s = bu_s;
        st->gen_s = gen_s;
    }
    return XBFS_OK;
}

// Persistent driver: the whole level loop in one cooperative launch.
static int bfs_persistent(xbfs_ctx *c, const XbArgs &a, const xbfs_opts &o, xbfs_level_rec *rec,
                          int32_t rec_cap, xbfs_stats *st)
{
    int grid = c->coop_blocks;
    if (o.grid_dim > 0 && o.grid_dim < grid) grid = o.grid_dim;
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, c->stream, c->d_ctl);
    XB_LAUNCHED();
    XB_HIP(hipEventRecord(c->ev_start, c->stream));
    XbArgs args = a;
    void *kargs[] = {&args};
    XB_HIP(hipLaunchCooperativeKernel((const void *)k_persistent, dim3(grid), dim3(XB_PERSIST_BLOCK),
                                      kargs, 0, c->stream));
    XB_HIP(hipEventRecord(c->ev_stop, c->stream));
    XB_HIP(hipEventSynchronize(c->ev_stop));
    XB_HIP(hipMemcpy(c->h_ctl, c->d_ctl, sizeof(XbCtl), hipMemcpyDeviceToHost));
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
            if (rec && L < rec_cap)
                fill_rec(rec[L], L, d.mode, (int8_t)(d.htag & 0xff), (d.htag >> 8) & 0xff, d.n_f, d.m_f,
                         d.n_u, d.m_u, c->m, t_s);
        }
    }
    if (st) {
        st->t_bfs_s = ms * 1e-3;
        st->n_levels = h.n_levels;
        st->max_level = h.max_level;
        st->v_cc = h.v_cc;
        st->m_cc = h.m_cc;
        st->td_s = td_s;
        st->bu_s = bu_s;
        st->gen_s = 0;
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
    XB_HIP(hipSetDevice(c->device));
    if (o.policy == XBFS_POLICY_SCHEDULE)
        XB_HIP(hipMemcpyAsync(c->d_sched, o.schedule, (size_t)o.schedule_len,
                              hipMemcpyHostToDevice, c->stream));
    const XbArgs a = make_args(c, src, o);
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

// xbfs_bfs_noalloc: the single-source traversal of xbfs_bfs for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  The caller passes the CSR and the
// output arrays in device memory; every other buffer the level phases use
// (control block, partial statistics, degree bitmap, hub table, visited and
// frontier bitmaps, binned queues) lives in `scratch`, one file-scope device
// array of XBFS_NOALLOC_SCRATCH_BYTES bytes (60 MiB) that each GPU holds from
// the library's module load until the process ends.  The traversal is the
// persistent kernel of xbfs_bfs, unchanged; rows are scanned in their uploaded
// order (no degree-ordered copy and no create-time component: both need
// memory of the graph's size).
//
// Every call on a GPU uses that GPU's array and leaves the last call's
// contents in it; each call rebuilds or clears what it reads before reading
// it.  Calls that overlap on one GPU (several host threads, or a call while
// another is in flight) race on the array: serialise them.
#include <hip/hip_runtime.h>
#include <chrono>
#include <mutex>

#define XB_NA_MAX_DEV 64
#define XB_NA_NWORDS (XBFS_NOALLOC_MAX_N / 64)
#define XB_NA_FSTRIDE (((XB_NA_NWORDS + XB_BU_TILE_WORDS - 1) / XB_BU_TILE_WORDS) * XB_BU_TILE_WORDS)

// Byte offsets of the buffers inside the array (each a multiple of 256).
constexpr size_t xb_na_round(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t XB_NA_OFF_CTL = 0;
constexpr size_t XB_NA_OFF_LSG = XB_NA_OFF_CTL + xb_na_round(sizeof(XbCtl));
constexpr size_t XB_NA_OFF_CNT = XB_NA_OFF_LSG + xb_na_round(sizeof(XbLoopShared));
constexpr size_t XB_NA_OFF_MLP = XB_NA_OFF_CNT + 256;
constexpr size_t XB_NA_OFF_SCHED = XB_NA_OFF_MLP + xb_na_round(sizeof(float) * XB_MLP_NPARAM);
constexpr size_t XB_NA_OFF_PART = XB_NA_OFF_SCHED + xb_na_round(XB_REC_CAP);
constexpr size_t XB_NA_OFF_NZ = XB_NA_OFF_PART + sizeof(long long) * XB_NFIELDS * XB_PMAX;
constexpr size_t XB_NA_OFF_VIS = XB_NA_OFF_NZ + sizeof(uint64_t) * XB_NA_FSTRIDE;
constexpr size_t XB_NA_OFF_FBM = XB_NA_OFF_VIS + sizeof(uint64_t) * XB_NA_FSTRIDE;
constexpr size_t XB_NA_OFF_HD = XB_NA_OFF_FBM + sizeof(uint64_t) * XB_NA_FSTRIDE * XB_FRINGS;
constexpr size_t XB_NA_OFF_Q = XB_NA_OFF_HD + sizeof(int2) * (size_t)XBFS_NOALLOC_MAX_N;
constexpr size_t XB_NA_BYTES =
    XB_NA_OFF_Q + sizeof(int32_t) * (size_t)XBFS_NOALLOC_MAX_N * XB_QBUFS * XB_BINS;
static_assert(XB_NA_BYTES <= XBFS_NOALLOC_SCRATCH_BYTES, "the buffers must fit the stated size");

static __device__ __align__(256) unsigned char scratch[XBFS_NOALLOC_SCRATCH_BYTES];

// Per-device set-up, once: the array's device address (its host-side name is
// not one), the persistent kernel's co-resident block count and the wall
// clock rate.
struct XbNaDev {
    unsigned char *ptr = nullptr;
    int coop_blocks = 0, num_cu = 0;
    double wall_hz = 1e8;
};
static XbNaDev g_na[XB_NA_MAX_DEV];
static std::mutex g_na_mu;

static int na_setup(int device, XbNaDev *out)
{
    std::lock_guard<std::mutex> lk(g_na_mu);
    XbNaDev &d = g_na[device];
    if (!d.ptr) {
        unsigned char *p = nullptr;
        if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return XBFS_ERR_HIP;
        int num_cu = 0, clk_khz = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, device) !=
                hipSuccess ||
            hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persistent, XB_PERSIST_BLOCK,
                                                         XB_LDS_PERSIST) != hipSuccess ||
            per_cu < 1)
            return XBFS_ERR_HIP;
        d.num_cu = num_cu;
        d.coop_blocks = std::min(per_cu * num_cu, XB_PMAX);
        if (clk_khz > 0) d.wall_hz = (double)clk_khz * 1e3;
        d.ptr = p;
    }
    *out = d;
    return XBFS_OK;
}

// out[0] = vertices of degree > 0 (set bits of nz), out[1] = largest degree;
// out must be zero on entry.
__global__ void __launch_bounds__(256) k_na_degree_info(const int64_t *__restrict__ beg, int64_t n,
                                                        const uint64_t *__restrict__ nz,
                                                        int64_t nwords,
                                                        unsigned long long *__restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long cnt = 0, mx = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned long long d = (unsigned long long)(beg[i + 1] - beg[i]);
        mx = d > mx ? d : mx;
        if (i < nwords) cnt += (unsigned long long)__popcll(nz[i]);
    }
    cnt = xb_wave_sum_u64(cnt);
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(mx, o);
        mx = y > mx ? y : mx;
    }
    if (xb_lane() == 0) {
        if (cnt) atomicAdd(&out[0], cnt);
        atomicMax(&out[1], mx);
    }
}

extern "C" int xbfs_bfs_noalloc(int device, const int64_t *d_beg, const int32_t *d_adj, int64_t n,
                                int64_t m, int32_t src, const xbfs_opts *opts,
                                const float *policy_weights, int32_t *d_level, int32_t *d_parent,
                                xbfs_stats *st)
{
    if (n > XBFS_NOALLOC_MAX_N) return XBFS_ERR_LIMIT;
    if (n < 1 || m < 0 || m > INT32_MAX || !d_beg || (m > 0 && !d_adj) || !d_level)
        return XBFS_ERR_ARG;
    if (src < 0 || (int64_t)src >= n) return XBFS_ERR_ARG;
    xbfs_opts o;
    int rc = xb_resolve_opts(opts, &o);
    if (rc) return rc;
    if (policy_weights)
        for (int i = 0; i < XB_MLP_NPARAM; i++)
            if (!(policy_weights[i] > -1e30f && policy_weights[i] < 1e30f)) return XBFS_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return XBFS_ERR_HIP;
    if (device < 0 || device >= ndev || device >= XB_NA_MAX_DEV) return XBFS_ERR_ARG;
    XB_HIP(hipSetDevice(device));
    XbNaDev dv;
    if ((rc = na_setup(device, &dv)) != XBFS_OK) return rc;
    unsigned char *const base = dv.ptr;
    XbCtl *ctl = reinterpret_cast<XbCtl *>(base + XB_NA_OFF_CTL);
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(base + XB_NA_OFF_CNT);
    float *mlp = reinterpret_cast<float *>(base + XB_NA_OFF_MLP);
    int8_t *sched = reinterpret_cast<int8_t *>(base + XB_NA_OFF_SCHED);
    uint64_t *nz = reinterpret_cast<uint64_t *>(base + XB_NA_OFF_NZ);
    int2 *hd = reinterpret_cast<int2 *>(base + XB_NA_OFF_HD);
    const hipStream_t s0 = nullptr;
    const int64_t nwords = ceil_div(n, 64);

    // per-graph tables: degree bitmap, hub table, (count, largest degree)
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s0, cnt, 2);
    hipLaunchKernelGGL(k_build_nz, dim3(grid_for(nwords * 64, 256, 65536)), dim3(256), 0, s0, d_beg,
                       n, nz, nwords);
    hipLaunchKernelGGL(k_build_hd, dim3(grid_for(n, 256, 65536)), dim3(256), 0, s0, d_beg, d_adj, n,
                       hd);
    hipLaunchKernelGGL(k_build_hd_big, dim3(dv.num_cu * 8), dim3(256), 0, s0, d_beg, d_adj, n, hd);
    hipLaunchKernelGGL(k_na_degree_info, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s0, d_beg, n,
                       nz, nwords, cnt);
    XB_LAUNCHED();
    unsigned long long info[2] = {0, 0};
    XB_HIP(hipMemcpy(info, cnt, sizeof(info), hipMemcpyDeviceToHost));
    const int64_t n_nz = (int64_t)info[0];
    const double skew = n_nz > 0 && m > 0 ? (double)info[1] * (double)n_nz / (double)m : 0.0;
    if (o.policy == XBFS_POLICY_SCHEDULE)
        XB_HIP(hipMemcpy(sched, o.schedule, (size_t)o.schedule_len, hipMemcpyHostToDevice));
    const bool use_mlp = o.policy == XBFS_POLICY_LEARNED && policy_weights &&
                         skew >= XBFS_MLP_MIN_SKEW;
    if (use_mlp)
        XB_HIP(hipMemcpy(mlp, policy_weights, sizeof(float) * XB_MLP_NPARAM, hipMemcpyHostToDevice));

    XbArgs a;
    a.level = d_level;
    a.parent = d_parent;
    a.beg = d_beg;
    a.adj = d_adj;
    a.adj_bu = d_adj;
    a.nz = nz;
    a.hd = hd;
    a.vis = reinterpret_cast<uint64_t *>(base + XB_NA_OFF_VIS);
    a.fbm = reinterpret_cast<uint64_t *>(base + XB_NA_OFF_FBM);
    a.qbuf = reinterpret_cast<int32_t *>(base + XB_NA_OFF_Q);
    a.ctl = ctl;
    a.lsg = reinterpret_cast<XbLoopShared *>(base + XB_NA_OFF_LSG);
    a.part = reinterpret_cast<long long *>(base + XB_NA_OFF_PART);
    a.recs = nullptr;
    a.rec_cap = 0;
    a.sched = sched;
    a.sched_len = o.policy == XBFS_POLICY_SCHEDULE ? o.schedule_len : 0;
    a.mlp = use_mlp ? mlp : nullptr;
    a.n = n;
    a.m = m;
    a.nwords = nwords;
    a.ntiles = ceil_div(nwords, XB_BU_TILE_WORDS);
    a.fstride = a.ntiles * XB_BU_TILE_WORDS;
    a.n_nz = n_nz;
    a.n_target = n_nz;
    a.srcs = nullptr;
    a.nsrc = 0;
    a.hub_first = 0;
    a.nt_bu = 8 * (long long)n > XB_NT_HUB_BYTES ? 1 : 0;
    a.src = src;
    a.t1 = o.bin_thread_max;
    a.t2 = o.bin_wave_max;
    a.t3 = o.bin_block_max;
    a.policy = o.policy;
    a.alpha = o.alpha;
    a.max_levels = (int32_t)std::min<int64_t>(n + 1, INT32_MAX - 4);
    if (o.max_levels > 0) a.max_levels = std::min(a.max_levels, o.max_levels);

    int grid = dv.coop_blocks;
    if (o.grid_dim > 0 && o.grid_dim < grid) grid = o.grid_dim;
    const double us = o.wait_limit_us > 0 ? (double)o.wait_limit_us : 4.0e6;
    const double ticks = us * 1e-6 * dv.wall_hz;
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, s0, ctl,
                       ticks < 1.0 ? 1ull : (unsigned long long)ticks, 0ull);
    XB_LAUNCHED();
    XB_HIP(hipStreamSynchronize(s0));
    void *kargs[] = {&a};
    const auto t0 = std::chrono::steady_clock::now();
    XB_HIP(hipLaunchCooperativeKernel((const void *)k_persistent, dim3(grid), dim3(XB_PERSIST_BLOCK),
                                      kargs, XB_LDS_PERSIST, s0));
    XB_HIP(hipStreamSynchronize(s0));
    const auto t1 = std::chrono::steady_clock::now();
    XbCtl h;
    XB_HIP(hipMemcpy(&h, ctl, sizeof(XbCtl), hipMemcpyDeviceToHost));
    if (h.error == 1) return XBFS_ERR_DEVICE;
    if (h.error == 2) return XBFS_ERR_LIMIT;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->t_bfs_s = std::chrono::duration<double>(t1 - t0).count();
        st->n_levels = h.n_levels;
        st->max_level = h.max_level;
        st->v_cc = h.v_cc;
        st->m_cc = h.m_cc;
    }
    return XBFS_OK;
}

// Level phases of the adaptive BFS and their two drivers.
//
// Frontier representation.  Every operation that assigns level l to a vertex
// also sets its bit in the frontier bitmap of ring slot l & 3 and in the
// visited bitmap.  At the start of level L the bitmap F_L therefore holds
// exactly the level-L vertices, whichever mode discovered them (including
// vertices the bottom-up +2 rule set two levels early).  A binned queue of F_L
// is available for free when level L-1 ran top-down scan-free and no +2
// discovery targets level L; otherwise a ballot sweep of F_L builds it.
//
// Phases (device functions run by every block of a grid, with a grid-wide
// barrier between phases):
//   ph_init      level/parent = -1, bitmaps cleared, source seeded
//   ph_compact   F_L -> degree-binned queue (+ frontier statistics), by
//                64-bit ballots over bitmap words: the single-scan frontier
//                generator (bfs_single_scan.cuh:26-50) and the post-bottom-up
//                queue rebuild ("last phase", bfs_TD_scan_free.cuh:23-135)
//   ph_td<SF>    top-down expansion of the binned queue: SF claims with CAS
//                and appends winners to the next queue's degree bins
//                (bfs_adaptiveFQ_wb_async.cuh:50-162); SS uses plain stores and
//                no queue (bfs_single_scan.cuh:52-126)
//   ph_bu        bottom-up over unvisited vertices (nz & ~visited), compacted
//                per 4096-vertex tile in LDS by ballot/popcount and a wave scan
//                (ballot_counting/FQ_gen/prefix sum, bfs_adaptiveFQ_wb_async.cuh:
//                189-284, prefix_sum1.cuh) and updated with the asynchronous
//                +2 rule (thread_centric_BU_update, bfs_adaptiveFQ_wb_async.cuh:738-777)
// Drivers: k_persistent runs the whole level loop in one cooperative launch
// with a grid barrier; the host loop launches one kernel per phase.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_device.h"
#include "xbfs_engine.h"
#include "xbfs_policy.h"

#define XB_BAR_TIMEOUT_TICKS (100ull * 1000 * 1000 * 4) /* ~4 s at the 100 MHz wall clock */

__device__ __forceinline__ unsigned long long xb_ld_u64(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t xb_ld_i32(const int32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void xb_st(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide barrier for a cooperative launch (all blocks co-resident).  The
// wait is bounded: after ~4 s, or when another block reported a timeout, it
// sets ctl->error and returns false so every block leaves the kernel.
__device__ bool xb_grid_sync(XbCtl *c)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        // one release (write back this XCD's dirty L2 lines) before arriving
        // and one acquire after leaving; the wait itself polls with relaxed
        // loads so it does not invalidate the caches of blocks still working
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t nb = gridDim.x;
        const uint32_t gen = __hip_atomic_load(&c->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t prev =
            __hip_atomic_fetch_add(&c->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        if (prev == nb - 1) {
            // the last arrival acquires every other block's release before
            // publishing the new generation
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(&c->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&c->bar_gen, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(&c->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (__hip_atomic_load(&c->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                    wall_clock64() - t0 > XB_BAR_TIMEOUT_TICKS) {
                    __hip_atomic_store(&c->error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// ---------------------------------------------------------------------------
// init
// ---------------------------------------------------------------------------
__device__ void ph_init(const XbArgs &a, int bid, int nb)
{
    const int64_t gtid = (int64_t)bid * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)nb * blockDim.x;
    const int64_t n4 = a.n >> 2;
    int4 *l4 = reinterpret_cast<int4 *>(a.level);
    int4 *p4 = reinterpret_cast<int4 *>(a.parent);
    const int32_t s = a.src;
    for (int64_t i = gtid; i < n4; i += nthr) {
        int4 v = make_int4(-1, -1, -1, -1);
        int4 p = v;
        if ((s >> 2) == i) {
            int k = s & 3;
            if (k == 0) { v.x = 0; p.x = s; }
            else if (k == 1) { v.y = 0; p.y = s; }
            else if (k == 2) { v.z = 0; p.z = s; }
            else { v.w = 0; p.w = s; }
        }
        l4[i] = v;
        if (a.parent) p4[i] = p;
    }
    for (int64_t i = n4 * 4 + gtid; i < a.n; i += nthr) {
        a.level[i] = (i == s) ? 0 : -1;
        if (a.parent) a.parent[i] = (i == s) ? s : -1;
    }
    const uint64_t sbit = 1ull << (s & 63);
    const int64_t sw = s >> 6;
    for (int64_t w = gtid; w < a.nwords; w += nthr) {
        const uint64_t v = (w == sw) ? sbit : 0ull;
        a.vis[w] = v;
        xb_f(a, 0)[w] = v; // ring 0 = level 0
        xb_f(a, 1)[w] = 0ull;
        xb_f(a, 2)[w] = 0ull;
        xb_f(a, 3)[w] = 0ull;
    }
    if (bid == 0 && threadIdx.x == 0) {
        XbCtl *c = a.ctl;
        for (int b = 0; b < XB_QBUFS; b++)
            for (int k = 0; k < XB_BINS; k++) xb_st(&c->qcnt[b][k], 0);
        for (int i = 0; i < 4; i++) {
            xb_st(&c->slot_n[i], 0ull);
            xb_st(&c->slot_m[i], 0ull);
            xb_st(&c->slot_p2[i], 0ull);
        }
        for (int i = 0; i < 2; i++) {
            xb_st(&c->sw_n[i], 0ull);
            xb_st(&c->sw_m[i], 0ull);
        }
        for (int i = 0; i < XB_QBUFS; i++) xb_st(&c->tile_ctr[i], 0u);
        xb_st(&c->n_levels, 0);
        xb_st(&c->max_level, 0);
        xb_st(&c->v_cc, 0ll);
        xb_st(&c->m_cc, 0ll);
        const int64_t d = a.beg[s + 1] - a.beg[s];
        int bin = d <= a.t1 ? 0 : (d <= a.t2 ? 1 : (d <= a.t3 ? 2 : 3));
        if (d > 0) {
            xb_q(a, 0, bin)[0] = s;
            xb_st(&c->qcnt[0][bin], 1);
        }
        xb_st(&c->slot_n[0], 1ull);
        xb_st(&c->slot_m[0], (unsigned long long)d);
    }
}

// Zeroing schedule, run by block 0 at the start of level L (before any of
// level L's writes; everything zeroed here is first written at level L+1):
//   ring slots (L+3)&3, queue buffer (L+2)%3, sweep slot (L+1)&1,
//   tile dispenser (L+1)%3.
// All blocks clear their share of bitmap ring (L+3)&3 = (L-1)&3, which no
// phase of level L reads or writes and level L+1 first writes.
__device__ void ph_level_prep(const XbArgs &a, int32_t L, int bid, int nb)
{
    if (bid == 0 && threadIdx.x == 0) {
        XbCtl *c = a.ctl;
        const int s3 = (L + 3) & 3;
        xb_st(&c->slot_n[s3], 0ull);
        xb_st(&c->slot_m[s3], 0ull);
        xb_st(&c->slot_p2[s3], 0ull);
        const int qb = (L + 2) % XB_QBUFS;
        for (int k = 0; k < XB_BINS; k++) xb_st(&c->qcnt[qb][k], 0);
        xb_st(&c->sw_n[(L + 1) & 1], 0ull);
        xb_st(&c->sw_m[(L + 1) & 1], 0ull);
        xb_st(&c->tile_ctr[(L + 1) % XB_QBUFS], 0u);
    }
    uint64_t *F3 = xb_f(a, L + 3);
    const int64_t gtid = (int64_t)bid * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)nb * blockDim.x;
    for (int64_t w = gtid; w < a.nwords; w += nthr) F3[w] = 0ull;
}

// ---------------------------------------------------------------------------
// F_L bitmap -> degree-binned queue in buffer L%3 (+ statistics)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xb_bin_of(const XbArgs &a, int64_t d)
{
    return d <= a.t1 ? 0 : (d <= a.t2 ? 1 : (d <= a.t3 ? 2 : 3));
}

// Block-aggregated append to four degree bins: every thread of the block
// calls it (bin -1 = nothing to append); one global atomic per non-empty bin.
__device__ void xb_block_append4(int bin, int32_t val, int32_t *const (&q)[XB_BINS], int32_t *cnt,
                                 int *s_wcnt, int *s_base)
{
    const int w = threadIdx.x >> 6, ln = xb_lane(), nw = (int)(blockDim.x >> 6);
    unsigned long long m[XB_BINS];
#pragma unroll
    for (int k = 0; k < XB_BINS; k++) {
        m[k] = __ballot(bin == k);
        if (ln == 0) s_wcnt[w * XB_BINS + k] = (int)__popcll(m[k]);
    }
    __syncthreads();
    if (threadIdx.x < XB_BINS) {
        const int k = threadIdx.x;
        int run = 0;
        for (int i = 0; i < nw; i++) {
            const int c = s_wcnt[i * XB_BINS + k];
            s_wcnt[i * XB_BINS + k] = run;
            run += c;
        }
        s_base[k] = run ? atomicAdd(&cnt[k], run) : 0;
    }
    __syncthreads();
    if (bin >= 0) {
        unsigned long long mk = 0;
#pragma unroll
        for (int k = 0; k < XB_BINS; k++)
            if (k == bin) mk = m[k];
        q[bin][s_base[bin] + s_wcnt[w * XB_BINS + bin] + xb_rank(mk)] = val;
    }
    __syncthreads();
}

// F_L -> binned queue of buffer L%3.  Tiles of XB_BU_TILE_WORDS bitmap words
// (static round robin over blocks): wave 0 reads the words and scans their
// popcounts, the set bits are expanded into an ascending LDS list, and the
// list is appended to the bins with one atomic per bin per round.
__device__ void ph_compact(const XbArgs &a, int32_t L, int bid, int nb, bool stats,
                           unsigned long long *sn, unsigned long long *sm_, XbSmem &sm)
{
    const uint64_t *F = xb_f(a, L);
    const int qb = L % XB_QBUFS;
    int32_t *const q[XB_BINS] = {xb_q(a, qb, 0), xb_q(a, qb, 1), xb_q(a, qb, 2), xb_q(a, qb, 3)};
    int32_t *cnt = a.ctl->qcnt[qb];
    const int ln = xb_lane(), wv = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    unsigned long long my_n = 0, my_m = 0;
    for (int64_t tile = bid; tile < a.ntiles; tile += nb) {
        const int64_t w0 = tile * XB_BU_TILE_WORDS;
        if (wv == 0) {
            const int64_t w = w0 + ln;
            uint64_t word = 0;
            if (ln < XB_BU_TILE_WORDS && w < a.nwords) word = F[w];
            const int pc = (int)__popcll(word);
            const int inc = xb_wave_incl_scan_i32(pc);
            if (ln < XB_BU_TILE_WORDS) {
                sm.cp.woff[ln] = inc - pc;
                sm.cp.words[ln] = word;
            }
            if (ln == 63) sm.cp.woff[XB_BU_TILE_WORDS] = inc;
        }
        __syncthreads();
        const int tcnt = sm.cp.woff[XB_BU_TILE_WORDS];
        if (tcnt == 0) {
            __syncthreads();
            continue;
        }
        for (int k = wv; k < XB_BU_TILE_WORDS; k += nw) {
            const uint64_t word = sm.cp.words[k];
            if ((word >> ln) & 1ull) {
                const uint64_t below = ln ? (~0ull >> (64 - ln)) : 0ull;
                sm.cp.list[sm.cp.woff[k] + (int)__popcll(word & below)] = (int32_t)((w0 + k) * 64 + ln);
            }
        }
        __syncthreads();
        for (int r = 0; r < tcnt; r += blockDim.x) {
            const int i = r + threadIdx.x;
            int bin = -1;
            int32_t v = 0;
            if (i < tcnt) {
                v = sm.cp.list[i];
                const int64_t d = a.beg[v + 1] - a.beg[v];
                my_n += 1;
                my_m += (unsigned long long)d;
                if (d > 0) bin = xb_bin_of(a, d);
            }
            xb_block_append4(bin, v, q, cnt, sm.cp.wcnt, sm.cp.base);
        }
    }
    if (stats) {
        unsigned long long vv[2] = {my_n, my_m};
        unsigned long long *const dd[2] = {sn, sm_};
        xb_block_add_u64<2>(vv, dd);
    }
}

// ---------------------------------------------------------------------------
// top-down expansion
// ---------------------------------------------------------------------------
struct XbTdCtx {
    const int32_t *qc[XB_BINS];
    int32_t nc[XB_BINS];
    int32_t *qn[XB_BINS];
    int32_t *cn;
    uint64_t *Fn; // bitmap of level L+1
    int32_t L1;
    bool enqueue;
};

template <bool SF>
__device__ __forceinline__ void td_visit(const XbArgs &a, const XbTdCtx &t, int32_t u,
                                         int32_t v, bool valid, unsigned long long &my_n,
                                         unsigned long long &my_m)
{
    bool won = false;
    int64_t d = 0;
    if (valid && a.level[v] == -1) {
        if (SF) {
            won = atomicCAS(&a.level[v], -1, t.L1) == -1;
        } else {
            a.level[v] = t.L1;
            won = true; // every writer stores the same level; any parent is valid
        }
        if (won) {
            if (a.parent) a.parent[v] = u;
            const uint64_t bit = 1ull << (v & 63);
            atomicOr((unsigned long long *)&a.vis[v >> 6], (unsigned long long)bit);
            atomicOr((unsigned long long *)&t.Fn[v >> 6], (unsigned long long)bit);
            if (SF) {
                d = a.beg[v + 1] - a.beg[v];
                my_n += 1;
                my_m += (unsigned long long)d;
            }
        }
    }
    if (SF && t.enqueue) {
        const int b = won && d > 0 ? xb_bin_of(a, d) : -1;
        xb_wave_append(t.qn[0], &t.cn[0], b == 0, v);
        xb_wave_append(t.qn[1], &t.cn[1], b == 1, v);
        xb_wave_append(t.qn[2], &t.cn[2], b == 2, v);
        xb_wave_append(t.qn[3], &t.cn[3], b == 3, v);
    }
}

template <bool SF>
__device__ void ph_td(const XbArgs &a, int32_t L, int bid, int nb, bool enqueue, XbSmem &sm)
{
    long long *s_pref = sm.td.pref;
    int32_t *s_u = sm.td.u;
    XbTdCtx t;
    const int qc = L % XB_QBUFS, qn = (L + 1) % XB_QBUFS;
    for (int k = 0; k < XB_BINS; k++) {
        t.qc[k] = xb_q(a, qc, k);
        t.nc[k] = xb_ld_i32(&a.ctl->qcnt[qc][k]);
        t.qn[k] = xb_q(a, qn, k);
    }
    t.cn = a.ctl->qcnt[qn];
    t.Fn = xb_f(a, L + 1);
    t.L1 = L + 1;
    t.enqueue = enqueue;
    unsigned long long my_n = 0, my_m = 0;
    const int64_t gtid = (int64_t)bid * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)nb * blockDim.x;
    const int ln = xb_lane();

    // grid bin: all threads of the grid share the concatenated adjacency of
    // the bin's vertices; each block scans their degrees in LDS and maps an
    // edge position to its vertex by binary search.
    for (int32_t base = 0; base < t.nc[3]; base += XB_GCAP) {
        const int32_t cnt = min(XB_GCAP, t.nc[3] - base);
        // two elements per thread (blockDim.x >= 256, so 2 * blockDim.x >= XB_GCAP)
        const int k0 = 2 * threadIdx.x, k1 = k0 + 1;
        long long d0 = 0, d1 = 0;
        if (k0 < cnt) {
            const int32_t u = t.qc[3][base + k0];
            s_u[k0] = u;
            d0 = a.beg[u + 1] - a.beg[u];
        }
        if (k1 < cnt) {
            const int32_t u = t.qc[3][base + k1];
            s_u[k1] = u;
            d1 = a.beg[u + 1] - a.beg[u];
        }
        long long tot = 0;
        const long long ex = xb_block_excl_scan_i64(d0 + d1, &tot);
        if (k0 < cnt) s_pref[k0] = ex;
        if (k1 < cnt) s_pref[k1] = ex + d0;
        if (threadIdx.x == 0) s_pref[cnt] = tot;
        __syncthreads();
        const long long T = s_pref[cnt];
        for (long long p = gtid; p < T; p += nthr) {
            int lo = 0, hi = cnt - 1; // largest k with s_pref[k] <= p
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_pref[mid] <= p) lo = mid;
                else hi = mid - 1;
            }
            const int32_t u = s_u[lo];
            const int64_t j = a.beg[u] + (p - s_pref[lo]);
            td_visit<SF>(a, t, u, a.adj[j], true, my_n, my_m);
        }
        __syncthreads();
    }
    // block bin: one block per vertex
    for (int32_t i = bid; i < t.nc[2]; i += nb) {
        const int32_t u = t.qc[2][i];
        const int64_t b = a.beg[u], e = a.beg[u + 1];
        for (int64_t j = b + threadIdx.x; j < e; j += blockDim.x)
            td_visit<SF>(a, t, u, a.adj[j], true, my_n, my_m);
    }
    // wave bin: one wave per vertex
    for (int64_t i = gtid >> 6; i < t.nc[1]; i += nthr >> 6) {
        const int32_t u = t.qc[1][i];
        const int64_t b = a.beg[u], e = a.beg[u + 1];
        for (int64_t j = b + ln; j < e; j += XB_WAVE)
            td_visit<SF>(a, t, u, a.adj[j], true, my_n, my_m);
    }
    // thread bin: one lane per vertex
    for (int64_t i = gtid; i < t.nc[0]; i += nthr) {
        const int32_t u = t.qc[0][i];
        const int64_t b = a.beg[u], e = a.beg[u + 1];
        for (int64_t j = b; j < e; j++) td_visit<SF>(a, t, u, a.adj[j], true, my_n, my_m);
    }
    if (SF) {
        unsigned long long vv[2] = {my_n, my_m};
        unsigned long long *const dd[2] = {&a.ctl->slot_n[(L + 1) & 3], &a.ctl->slot_m[(L + 1) & 3]};
        xb_block_add_u64<2>(vv, dd);
    }
}

// ---------------------------------------------------------------------------
// bottom-up
// ---------------------------------------------------------------------------
__device__ void ph_bu(const XbArgs &a, int32_t L, int bid, int nb, XbSmem &sm)
{
    int32_t *s_list = sm.bu.list;
    unsigned long long *s_vacc = sm.bu.vacc;
    unsigned long long *s_f1acc = sm.bu.f1acc;
    unsigned long long *s_f2acc = sm.bu.f2acc;
    int *s_woff = sm.bu.woff;
    uint32_t &s_tile = sm.bu.tile;
    (void)bid;
    (void)nb;
    const uint64_t *F = xb_f(a, L);
    uint64_t *F1 = xb_f(a, L + 1);
    uint64_t *F2 = xb_f(a, L + 2);
    const int32_t L1 = L + 1, L2 = L + 2;
    uint32_t *ctr = &a.ctl->tile_ctr[L % XB_QBUFS];
    const int ln = xb_lane(), wv = threadIdx.x >> 6;
    unsigned long long n1 = 0, m1 = 0, n2 = 0, m2 = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(ctr, 1u);
        __syncthreads();
        const int64_t tile = s_tile;
        if (tile >= a.ntiles) break;
        const int64_t w0 = tile * XB_BU_TILE_WORDS;
        // candidate masks and their prefix (wave 0: one word per lane)
        if (wv == 0) {
            const int64_t w = w0 + ln;
            uint64_t c = 0;
            if (ln < XB_BU_TILE_WORDS && w < a.nwords) c = a.nz[w] & ~a.vis[w];
            const int pc = (int)__popcll(c);
            const int inc = xb_wave_incl_scan_i32(pc);
            if (ln < XB_BU_TILE_WORDS) {
                s_woff[ln] = inc - pc;
                s_vacc[ln] = c; // candidate mask, reused below
                s_f1acc[ln] = 0ull;
                s_f2acc[ln] = 0ull;
            }
            if (ln == 63) s_woff[XB_BU_TILE_WORDS] = inc;
        }
        __syncthreads();
        const int cnt = s_woff[XB_BU_TILE_WORDS];
        // scatter candidate ids in ascending order: wave w handles words w, w+4, ...
        for (int k = wv; k < XB_BU_TILE_WORDS; k += (int)(blockDim.x >> 6)) {
            const uint64_t c = s_vacc[k];
            if ((c >> ln) & 1ull) {
                const uint64_t below = ln ? (~0ull >> (64 - ln)) : 0ull;
                s_list[s_woff[k] + (int)__popcll(c & below)] = (int32_t)((w0 + k) * 64 + ln);
            }
        }
        __syncthreads();
        if (threadIdx.x < XB_BU_TILE_WORDS) s_vacc[threadIdx.x] = 0ull;
        __syncthreads();
        for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
            const int32_t v = s_list[k];
            const int64_t b = a.beg[v], e = a.beg[v + 1];
            int32_t found = -1, seen = -1;
            for (int64_t j = b; j < e; j++) {
                const int32_t u = a.adj[j];
                const uint64_t fw = F[u >> 6];
                if ((fw >> (u & 63)) & 1ull) { found = u; break; }
                if (seen < 0) {
                    const uint64_t nw = F1[u >> 6];
                    if ((nw >> (u & 63)) & 1ull) seen = u;
                }
            }
            const unsigned long long bit = 1ull << (v & 63);
            const int li = (int)((v >> 6) - w0);
            if (found >= 0) {
                a.level[v] = L1;
                if (a.parent) a.parent[v] = found;
                atomicOr(&s_f1acc[li], bit);
                atomicOr(&s_vacc[li], bit);
                n1 += 1;
                m1 += (unsigned long long)(e - b);
            } else if (seen >= 0) {
                a.level[v] = L2;
                if (a.parent) a.parent[v] = seen;
                atomicOr(&s_f2acc[li], bit);
                atomicOr(&s_vacc[li], bit);
                n2 += 1;
                m2 += (unsigned long long)(e - b);
            }
        }
        __syncthreads();
        // this block owns the tile's words for the whole level: plain stores
        if (threadIdx.x < XB_BU_TILE_WORDS) {
            const int64_t w = w0 + threadIdx.x;
            if (w < a.nwords) {
                const unsigned long long va = s_vacc[threadIdx.x];
                if (va) a.vis[w] |= va;
                const unsigned long long f1 = s_f1acc[threadIdx.x];
                if (f1) F1[w] |= f1; // other tiles' +2 probes see it from here on
                const unsigned long long f2 = s_f2acc[threadIdx.x];
                if (f2) F2[w] |= f2;
            }
        }
        __syncthreads();
    }
    // n2 also goes to slot_p2: a top-down level L+1 must not trust its winner
    // queue as the level-(L+2) frontier when +2 discoveries exist
    unsigned long long vv[5] = {n1, m1, n2, m2, n2};
    unsigned long long *const dd[5] = {&a.ctl->slot_n[L1 & 3], &a.ctl->slot_m[L1 & 3],
                                       &a.ctl->slot_n[L2 & 3], &a.ctl->slot_m[L2 & 3],
                                       &a.ctl->slot_p2[L2 & 3]};
    xb_block_add_u64<5>(vv, dd);
}

// ---------------------------------------------------------------------------
// level loop (shared by both drivers through these helpers)
// ---------------------------------------------------------------------------
// Level bookkeeping shared by both drivers.  Returns false when the
// traversal is complete (no frontier edges, or every non-isolated vertex
// already has its level); otherwise fills the unvisited counts of level L.
__host__ __device__ inline bool xb_level_open(XbLoop &st, const XbArgs &a, long long n_f,
                                              long long m_f, long long pend_n, long long pend_m,
                                              long long *n_u, long long *m_u)
{
    st.cum_n += n_f;
    st.cum_m += m_f;
    if (m_f == 0 || st.cum_n + pend_n >= a.n_nz) return false;
    *n_u = a.n_nz - (st.cum_n + pend_n);
    *m_u = a.m - (st.cum_m + pend_m);
    return true;
}

// Block-parallel MLP forward (thread o < 32 computes hidden unit o); all
// threads of the block must call it.  Writes the three predictions to y.
__device__ void xb_mlp_block_forward(const float *p, const float *x, float *s_h1, float *s_h2,
                                     float *y)
{
    const float *W1 = p, *b1 = W1 + XB_MLP_NH * XB_MLP_NI;
    const float *W2 = b1 + XB_MLP_NH, *b2 = W2 + XB_MLP_NH * XB_MLP_NH;
    const float *W3 = b2 + XB_MLP_NH, *b3 = W3 + XB_MLP_NO * XB_MLP_NH;
    const int o = threadIdx.x;
    if (o < XB_MLP_NH) {
        float s = b1[o];
        for (int i = 0; i < XB_MLP_NI; i++) s += W1[o * XB_MLP_NI + i] * x[i];
        s_h1[o] = s > 0.0f ? s : 0.0f;
    }
    __syncthreads();
    if (o < XB_MLP_NH) {
        float s = b2[o];
        for (int i = 0; i < XB_MLP_NH; i++) s += W2[o * XB_MLP_NH + i] * s_h1[i];
        s_h2[o] = s > 0.0f ? s : 0.0f;
    }
    __syncthreads();
    if (o < XB_MLP_NO) {
        float s = b3[o];
        for (int i = 0; i < XB_MLP_NH; i++) s += W3[o * XB_MLP_NH + i] * s_h2[i];
        y[o] = s;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(XB_PERSIST_BLOCK) k_persistent(XbArgs a)
{
    const int bid = blockIdx.x, nb = gridDim.x;
    __shared__ XbSmem sm;
    ph_init(a, bid, nb);
    if (!xb_grid_sync(a.ctl)) return;
    XbLoop st;
    xb_loop_init(&st);
    XbCtl *c = a.ctl;
    for (;;) {
        const int32_t L = st.L;
        const unsigned long long t0 = wall_clock64();
        if (st.need_sweep) {
            ph_compact(a, L, bid, nb, true, &c->sw_n[L & 1], &c->sw_m[L & 1], sm);
            if (!xb_grid_sync(c)) return;
        }
        long long n_f, m_f;
        if (st.need_sweep) {
            n_f = (long long)xb_ld_u64(&c->sw_n[L & 1]);
            m_f = (long long)xb_ld_u64(&c->sw_m[L & 1]);
        } else {
            n_f = (long long)xb_ld_u64(&c->slot_n[L & 3]);
            m_f = (long long)xb_ld_u64(&c->slot_m[L & 3]);
        }
        const long long pend_n = (long long)xb_ld_u64(&c->slot_n[(L + 1) & 3]);
        const long long pend_m = (long long)xb_ld_u64(&c->slot_m[(L + 1) & 3]);
        const long long p2_next = (long long)xb_ld_u64(&c->slot_p2[(L + 1) & 3]);
        long long n_u = 0, m_u = 0;
        const bool open = xb_level_open(st, a, n_f, m_f, pend_n, pend_m, &n_u, &m_u);
        if (!open || L >= a.max_levels) {
            if (bid == 0 && threadIdx.x == 0) {
                xb_st(&c->n_levels, L);
                xb_st(&c->max_level, pend_n > 0 ? L + 1 : (n_f > 0 ? L : L - 1));
                xb_st(&c->v_cc, (long long)(st.cum_n + pend_n));
                xb_st(&c->m_cc, (long long)(st.cum_m + pend_m));
                if (open) xb_st(&c->error, 2);
            }
            return;
        }
        const float *pred = nullptr;
        if (a.mlp) {
            if (threadIdx.x == 0) xbfs_mlp_features(L, n_f, m_f, n_u, m_u, st.prev_mode, sm.mlp.x);
            __syncthreads();
            xb_mlp_block_forward(a.mlp, sm.mlp.x, sm.mlp.h1, sm.mlp.h2, sm.mlp.y);
            pred = sm.mlp.y;
        }
        int32_t htag = 0, used = 0;
        const int32_t mode = xbfs_choose_mode(a.policy, a.sched, a.sched_len, pred, &st.hs, L, m_f,
                                              a.m, a.alpha, &htag, &used);
        ph_level_prep(a, L, bid, nb);
        if (mode != XBFS_MODE_BU && !st.q_valid && !st.need_sweep) {
            ph_compact(a, L, bid, nb, false, nullptr, nullptr, sm);
            if (!xb_grid_sync(c)) return;
        }
        if (mode == XBFS_MODE_TD_SF) ph_td<true>(a, L, bid, nb, p2_next == 0, sm);
        else if (mode == XBFS_MODE_TD_SS) ph_td<false>(a, L, bid, nb, false, sm);
        else ph_bu(a, L, bid, nb, sm);
        if (!xb_grid_sync(c)) return;
        if (bid == 0 && threadIdx.x == 0 && L < a.rec_cap) {
            XbDevRec &r = a.recs[L];
            r.level = L;
            r.mode = mode;
            r.tag = xbfs_mode_to_tag(mode);
            r.htag = (htag & 0xff) | (used << 8);
            r.n_f = n_f;
            r.m_f = m_f;
            r.n_u = n_u;
            r.m_u = m_u;
            r.t0 = (long long)t0;
            r.t1 = (long long)wall_clock64();
        }
        st.q_valid = (mode == XBFS_MODE_TD_SF && p2_next == 0);
        st.need_sweep = (mode == XBFS_MODE_TD_SS);
        st.prev_mode = mode;
        st.L = L + 1;
    }
}

// One level of the persistent kernel as its own cooperative launch (same
// block shape, same phases, same barriers): the per-level probe times the
// three modes with it so its labels match what k_persistent would spend.
// extra_sweep additionally runs the frontier sweep of level L+1 that a
// single-scan level leaves to its successor.
__global__ void __launch_bounds__(XB_PERSIST_BLOCK) k_coop_level(XbArgs a, int32_t L, int32_t mode,
                                                                 int32_t q_valid, int32_t need_sweep,
                                                                 int32_t enqueue, int32_t extra_sweep)
{
    __shared__ XbSmem sm;
    const int bid = blockIdx.x, nb = gridDim.x;
    XbCtl *c = a.ctl;
    ph_level_prep(a, L, bid, nb);
    if (mode != XBFS_MODE_BU && !q_valid && !need_sweep) {
        ph_compact(a, L, bid, nb, false, nullptr, nullptr, sm);
        if (!xb_grid_sync(c)) return;
    }
    if (mode == XBFS_MODE_TD_SF) ph_td<true>(a, L, bid, nb, enqueue != 0, sm);
    else if (mode == XBFS_MODE_TD_SS) ph_td<false>(a, L, bid, nb, false, sm);
    else ph_bu(a, L, bid, nb, sm);
    if (extra_sweep) {
        if (!xb_grid_sync(c)) return;
        ph_compact(a, L + 1, bid, nb, true, &c->sw_n[(L + 1) & 1], &c->sw_m[(L + 1) & 1], sm);
    }
}

// ---- per-phase kernels for the host-driven loop ---------------------------
__global__ void __launch_bounds__(XB_PBLOCK) k_ph_init(XbArgs a) { ph_init(a, blockIdx.x, gridDim.x); }

__global__ void __launch_bounds__(XB_PBLOCK) k_ph_compact(XbArgs a, int32_t L, int32_t stats)
{
    __shared__ XbSmem sm;
    ph_compact(a, L, blockIdx.x, gridDim.x, stats != 0, &a.ctl->sw_n[L & 1], &a.ctl->sw_m[L & 1], sm);
}

__global__ void __launch_bounds__(XB_PBLOCK) k_ph_prep(XbArgs a, int32_t L)
{
    ph_level_prep(a, L, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(XB_PBLOCK) k_ph_td_sf(XbArgs a, int32_t L, int32_t enqueue)
{
    __shared__ XbSmem sm;
    ph_td<true>(a, L, blockIdx.x, gridDim.x, enqueue != 0, sm);
}

__global__ void __launch_bounds__(XB_PBLOCK) k_ph_td_ss(XbArgs a, int32_t L)
{
    __shared__ XbSmem sm;
    ph_td<false>(a, L, blockIdx.x, gridDim.x, false, sm);
}

__global__ void __launch_bounds__(XB_PBLOCK) k_ph_bu(XbArgs a, int32_t L)
{
    __shared__ XbSmem sm;
    ph_bu(a, L, blockIdx.x, gridDim.x, sm);
}

__global__ void k_ctl_reset(XbCtl *c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        xb_st(&c->bar_count, 0u);
        xb_st(&c->bar_gen, 0u);
        xb_st(&c->error, 0);
    }
}

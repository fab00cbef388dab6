// Level phases of the adaptive BFS and their two drivers.
//
// Frontier representation.  Every operation that assigns level l to a vertex
// also sets its bit in the frontier bitmap of ring slot l & 3 and in the
// visited bitmap.  At the start of level L the bitmap F_L therefore holds
// exactly the level-L vertices, whichever mode discovered them.  A binned
// queue of F_L is available for free when level L-1 ran top-down scan-free;
// otherwise a ballot sweep of F_L builds it.  (The statistics keep fields for
// discoveries two levels ahead, the reference's bottom-up +2 rule; no phase
// makes them, so they stay zero.)
//
// Phases (device functions run by every block of a grid, with a grid-wide
// barrier between phases):
//   ph_init      level/parent = -1, bitmaps cleared, source seeded
//   ph_sweep     level-L frontier -> degree-binned queue (+ statistics),
//                wave-centric over 1024-vertex chunks of the frontier bitmap
//                (or of the level array after a single-scan level): the
//                single-scan frontier generator (bfs_single_scan.cuh:26-50)
//                and the post-bottom-up queue rebuild ("last phase",
//                bfs_TD_scan_free.cuh:23-135)
//   ph_td<SF>    top-down expansion of the binned queue: SF claims with CAS
//                and appends winners to the next queue's degree bins
//                (bfs_adaptiveFQ_wb_async.cuh:50-162); SS uses plain stores and
//                no queue (bfs_single_scan.cuh:52-126)
//   ph_bu        bottom-up over unvisited vertices (nz & ~visited), compacted
//                per wave-owned 1024-vertex chunk in LDS by popcount and a wave scan
//                (ballot_counting/FQ_gen/prefix sum, bfs_adaptiveFQ_wb_async.cuh:
//                189-284, prefix_sum1.cuh), each candidate settled by its first
//                level-L neighbour (thread_centric_BU_update,
//                bfs_adaptiveFQ_wb_async.cuh:738-777, without its +2 rule)
// Drivers: k_persistent runs the whole level loop in one cooperative launch
// with a grid barrier; the host loop launches one kernel per phase.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_device.h"
#include "xbfs_engine.h"
#include "xbfs_policy.h"


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

// Adds the block's totals of K counters to column `bid` of the partial
// statistics (thread 0, plain stores: within a phase only this block writes
// column bid).  With `known`, old[k] is the column's current value, which
// the caller tracks, and the store needs no load first; otherwise thread 0
// reads the value back (one memory round trip before the level's barrier).
// tot[k] receives the block totals (thread 0 only).  Every thread of the
// block must call it.
template <int K>
__device__ void xb_part_add(const XbArgs &a, int bid, unsigned long long (&v)[K], const int (&f)[K],
                            bool known, const long long (&old)[K], long long (&tot)[K])
{
    __shared__ unsigned long long sh[K][16];
    const int w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
#pragma unroll
    for (int k = 0; k < K; k++) {
        const unsigned long long t = xb_wave_sum_u64(v[k]);
        if (xb_lane() == 0) sh[k][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            unsigned long long t = 0;
            for (int i = 0; i < nw; i++) t += sh[k][i];
            tot[k] = (long long)t;
            if (t) {
                long long *col = &a.part[(int64_t)f[k] * XB_PMAX + bid];
                *col = (known ? old[k] : *col) + (long long)t;
            }
        }
    }
    __syncthreads();
}
template <int K>
__device__ void xb_part_add(const XbArgs &a, int bid, unsigned long long (&v)[K], const int (&f)[K])
{
    long long old[K] = {}, tot[K];
    xb_part_add<K>(a, bid, v, f, false, old, tot);
}

// Sums of K fields over the first ncols columns of the partial statistics
// (the blocks that wrote them; the zeroing schedule clears every column, so
// the others are zero), computed by wave `wave` of the block alone: every
// lane issues all its loads at once, then one wave reduction.  Lane 0 of that
// wave gets the results; other threads must not use out.  The columns are
// read 256 at a time (4 per lane and field, unrolled), so up to 4K loads per
// lane are in flight before the first wait: one memory round trip per 256
// blocks instead of one per column.
template <int K>
__device__ void xb_part_sum(const XbArgs &a, const int (&f)[K], long long (&out)[K], int ncols,
                            int wave = 0)
{
    if ((int)(threadIdx.x >> 6) != wave) return;
    const int ln = xb_lane();
    long long acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0;
    for (int b0 = 0; b0 < ncols; b0 += 256) {
        long long v[K][4];
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int b = b0 + ln + 64 * j;
                v[k][j] = b < ncols ? a.part[(int64_t)f[k] * XB_PMAX + b] : 0;
            }
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += (v[k][0] + v[k][1]) + (v[k][2] + v[k][3]);
    }
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = (long long)xb_wave_sum_u64((unsigned long long)acc[k]);
}

// Barrier among the first `members` blocks of a cooperative launch (all
// blocks co-resident) on the counter / generation pair (cnt, gen).  The wait
// is bounded: after ctl->wait_ticks without progress (4 s unless the caller
// asked for less), or when another block reported a timeout, it sets
// ctl->error and returns false so every block leaves the kernel.  With watch_hb, progress also means a change of ctl->heartbeat
// (blocks parked during small-frontier levels wait as long as those levels
// keep coming).
__device__ bool xb_barrier(XbCtl *c, uint32_t *cnt, uint32_t *genw, uint32_t members, bool watch_hb)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        // one release (write back this XCD's dirty L2 lines) before arriving
        // and one acquire after leaving; the wait itself polls with relaxed
        // loads so it does not invalidate the caches of blocks still working
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen = __hip_atomic_load(genw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t prev = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        if (prev == members - 1) {
            // the last arrival acquires every other block's release before
            // publishing the new generation
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(genw, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long limit =
                __hip_atomic_load(&c->wait_ticks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned long long t0 = wall_clock64();
            uint32_t hb = watch_hb ? __hip_atomic_load(&c->heartbeat, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT)
                                   : 0u;
            while (__hip_atomic_load(genw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (watch_hb) {
                    const uint32_t h = __hip_atomic_load(&c->heartbeat, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                    if (h != hb) {
                        hb = h;
                        t0 = wall_clock64();
                    }
                }
                // another block's error (a timeout, or the level cap that
                // block 0 reports before the last barrier) ends the wait and
                // keeps its code; a timeout of this wait reports 1 unless an
                // error is already there
                if (__hip_atomic_load(&c->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > limit) {
                    int32_t none = 0;
                    __hip_atomic_compare_exchange_strong(&c->error, &none, 1, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// The first grid barrier of a persistent launch.  A cooperative launch
// guarantees that every block is resident at once; a plain launch of one block
// per CU gets that on a GPU with no other work, but not by contract.  Under a
// plain launch (ctl->res_ticks > 0) a block that waits here longer than
// res_ticks reports code 3 and every block leaves at this barrier or the next,
// so the kernel ends whichever blocks ran; the host then relaunches
// cooperatively.  With res_ticks = 0 it is xb_grid_sync.
__device__ bool xb_first_sync(XbCtl *c)
{
    __shared__ int s_ok1;
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t *cnt = &c->bar_count, *genw = &c->bar_gen;
        const uint32_t gen = __hip_atomic_load(genw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t prev = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        if (prev == gridDim.x - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(genw, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long res =
                __hip_atomic_load(&c->res_ticks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long limit =
                res ? res : __hip_atomic_load(&c->wait_ticks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int32_t code = res ? 3 : 1;
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(genw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (__hip_atomic_load(&c->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > limit) {
                    int32_t none = 0;
                    __hip_atomic_compare_exchange_strong(&c->error, &none, code, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok1 = ok;
    }
    __syncthreads();
    return s_ok1 != 0;
}

// Grid-wide barrier (every block of the launch).
__device__ bool xb_grid_sync(XbCtl *c)
{
    return xb_barrier(c, &c->bar_count, &c->bar_gen, gridDim.x, false);
}

// Barrier of the blocks taking part in the current level: the whole grid, or
// the first nact blocks during small-frontier levels.
__device__ bool xb_level_sync(XbCtl *c, int nact)
{
    if (nact == (int)gridDim.x) return xb_grid_sync(c);
    return xb_barrier(c, &c->sbar_count, &c->sbar_gen, (uint32_t)nact, false);
}

// A block sitting out small-frontier levels: arrives at the grid barrier and
// waits there until the active blocks arrive too.
__device__ bool xb_park(XbCtl *c)
{
    return xb_barrier(c, &c->bar_count, &c->bar_gen, gridDim.x, true);
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
    // single source: seeded here; multi-source: ph_seed seeds after the barrier
    const int32_t s = a.nsrc > 0 ? -1 : a.src;
    for (int64_t i = gtid; i < n4; i += nthr) {
        int4 v = make_int4(-1, -1, -1, -1);
        int4 p = v;
        if (s >= 0 && (s >> 2) == i) {
            int k = s & 3;
            if (k == 0) { v.x = 0; p.x = s; }
            else if (k == 1) { v.y = 0; p.y = s; }
            else if (k == 2) { v.z = 0; p.z = s; }
            else { v.w = 0; p.w = s; }
        }
        // non-temporal: 4 bytes per vertex written once, not read back here
        typedef int xb_v4i __attribute__((ext_vector_type(4)));
        const xb_v4i vv = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(vv, reinterpret_cast<xb_v4i *>(&l4[i]));
        if (a.parent) p4[i] = p;
    }
    for (int64_t i = n4 * 4 + gtid; i < a.n; i += nthr) {
        a.level[i] = (i == s) ? 0 : -1;
        if (a.parent) a.parent[i] = (i == s) ? s : -1;
    }
    const uint64_t sbit = s >= 0 ? 1ull << (s & 63) : 0ull;
    const int64_t sw = s >= 0 ? (s >> 6) : -1;
    for (int64_t w = gtid; w < a.nwords; w += nthr) {
        const uint64_t v = (w == sw) ? sbit : 0ull;
        a.vis[w] = v;
        xb_f(a, 0)[w] = v; // ring 0 = level 0
        xb_f(a, 1)[w] = 0ull;
        xb_f(a, 2)[w] = 0ull;
        xb_f(a, 3)[w] = 0ull;
    }
    // partial statistics: all zero except level 0's slot = (1, deg(src))
    const long long dsrc = s >= 0 ? a.beg[s + 1] - a.beg[s] : 0;
    for (int64_t i = gtid; i < (int64_t)XB_NFIELDS * XB_PMAX; i += nthr) {
        long long v = 0;
        if (s >= 0 && i == (int64_t)xb_f_ring(0, 0) * XB_PMAX) v = 1;
        else if (i == (int64_t)xb_f_ring(0, 1) * XB_PMAX) v = dsrc;
        a.part[i] = v;
    }
    if (bid == 0 && threadIdx.x == 0) {
        XbCtl *c = a.ctl;
        for (int b = 0; b < XB_QBUFS; b++)
            for (int k = 0; k < XB_BINS; k++) xb_st(&c->qcnt[b][k].v, 0);
        xb_st(&c->n_levels, 0);
        xb_st(&c->max_level, 0);
        xb_st(&c->v_cc, 0ll);
        xb_st(&c->m_cc, 0ll);
        xb_st(&c->seed_iso, 0ll);
        const int bin = dsrc <= a.t1 ? 0 : (dsrc <= a.t2 ? 1 : (dsrc <= a.t3 ? 2 : 3));
        if (dsrc > 0) {
            xb_q(a, 0, bin)[0] = s;
            xb_st(&c->qcnt[0][bin].v, 1);
        }
    }
}

// Multi-source seeding (after ph_init and a barrier): every distinct seed
// gets level 0 and its visited bit; seeds with edges also get their F_0 bit
// and count toward level 0's statistics.  Degree-0 seeds are counted apart
// (they are reached but are not frontier vertices).  Duplicates are resolved
// by the visited bit: only the lane that sets it writes the seed.  The
// level-0 queue is then built by a sweep of F_0.
__device__ void ph_seed(const XbArgs &a, int bid, int nb)
{
    const int64_t gtid = (int64_t)bid * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)nb * blockDim.x;
    unsigned long long nv = 0, mv = 0, iso = 0;
    for (int64_t i = gtid; i < a.nsrc; i += nthr) {
        const int32_t s = a.srcs[i];
        const unsigned long long bit = 1ull << (s & 63);
        const unsigned long long old =
            atomicOr(reinterpret_cast<unsigned long long *>(&a.vis[s >> 6]), bit);
        if (old & bit) continue;
        a.level[s] = 0;
        if (a.parent) a.parent[s] = s;
        const long long d = a.beg[s + 1] - a.beg[s];
        if (d > 0) {
            atomicOr(reinterpret_cast<unsigned long long *>(&xb_f(a, 0)[s >> 6]), bit);
            nv += 1;
            mv += (unsigned long long)d;
        } else {
            iso += 1;
        }
    }
    if (iso) atomicAdd(reinterpret_cast<unsigned long long *>(&a.ctl->seed_iso), iso);
    unsigned long long vv[2] = {nv, mv};
    const int ff[2] = {xb_f_ring(0, 0), xb_f_ring(0, 1)};
    xb_part_add<2>(a, bid, vv, ff);
}

// Zeroing schedule at the start of level L (everything zeroed here is
// first written at level L+1): queue buffer (L+2)%3 (block 0), the partial
// statistics of ring slot (L+3)&3 and sweep slot (L+1)&1, and bitmap ring
// (L+3)&3 = (L-1)&3, which no phase of level L reads or writes (nor the
// level's statistics and decision, so the two may run side by side).
// Threads below t_lo of each block take no part.
__device__ void ph_level_prep(const XbArgs &a, int32_t L, int bid, int nb, int t_lo = 0)
{
    if ((int)threadIdx.x < t_lo) return;
    const int tid = (int)threadIdx.x - t_lo, bdim = (int)blockDim.x - t_lo;
    if (bid == 0 && tid == 0) {
        XbCtl *c = a.ctl;
        const int qb = (L + 2) % XB_QBUFS;
        for (int k = 0; k < XB_BINS; k++) xb_st(&c->qcnt[qb][k].v, 0);
    }
    const int64_t gtid = (int64_t)bid * bdim + tid;
    const int64_t nthr = (int64_t)nb * bdim;
    // ring slot (L+3)&3 holds the bottom-up work of level L+2 as well
    const int zf[8] = {xb_f_ring(L + 3, 0), xb_f_ring(L + 3, 1), xb_f_ring(L + 3, 2),
                       xb_f_sweep(L + 1, 0), xb_f_sweep(L + 1, 1),
                       xb_f_bu(L + 2, 0),    xb_f_bu(L + 2, 1),    xb_f_bu(L + 2, 2)};
    for (int64_t i = gtid; i < 8 * XB_PMAX; i += nthr)
        a.part[(int64_t)zf[i / XB_PMAX] * XB_PMAX + (i % XB_PMAX)] = 0;
    uint64_t *F3 = xb_f(a, L + 3);
    for (int64_t w = gtid; w < a.nwords; w += nthr) F3[w] = 0ull;
}

// ---------------------------------------------------------------------------
// F_L bitmap -> degree-binned queue in buffer L%3 (+ statistics)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xb_bin_of(const XbArgs &a, int64_t d)
{
    return d <= a.t1 ? 0 : (d <= a.t2 ? 1 : (d <= a.t3 ? 2 : 3));
}

// LDS ordering between lanes of one wave: a code-motion barrier plus a
// wavefront-scope fence (waits for this wave's LDS operations).
__device__ __forceinline__ void xb_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Fills the wave's LDS list with the set bits of the chunk's words (lane k
// < XB_CHUNK_WORDS holds word w0 + k), in ascending vertex order; returns the
// count (wave-uniform).
__device__ __forceinline__ int xb_chunk_list(uint64_t myword, int64_t w0, XbWaveLds *wl)
{
    const int ln = xb_lane();
    const int pc = (int)__popcll(myword);
    const int inc = xb_wave_incl_scan_i32(pc);
    const int total = __shfl(inc, 63);
    if (total == 0) return 0;
    const int off = inc - pc;
    const uint64_t below = ln ? (~0ull >> (64 - ln)) : 0ull;
    for (int k = 0; k < XB_CHUNK_WORDS; k++) {
        const uint64_t word = __shfl(myword, k);
        const int o = __shfl(off, k);
        if ((word >> ln) & 1ull) wl->list[o + (int)__popcll(word & below)] = (int32_t)((w0 + k) * 64 + ln);
    }
    xb_wave_sync();
    return total;
}

// Frontier sweep of level L, wave-centric: each wave owns chunks of
// XB_CHUNK_WORDS bitmap words (1024 vertices).
//   from_levels = false: the level-L vertices are the bits of F_L (complete
//     whenever the previous level ran bottom-up or scan-free top-down);
//   from_levels = true: after a single-scan level, whose plain stores set no
//     bits, the chunk's level words are read (64 per ballot) and the
//     frontier bits of F_L and the visited bits are written here.
// Appends the level-L vertices of degree > 0 to their degree bins in queue
// buffer L%3 (single-scan frontier generation, bfs_single_scan.cuh:26-50,
// and the post-bottom-up queue rebuild, bfs_TD_scan_free.cuh:23-135).  Two
// passes: count per wave, reserve once per block and bin (one atomic), then
// write; the second pass re-reads F_L and the degrees (mostly from L2).
// With stats, adds (count, degree sum) to sweep slot L&1.
__device__ void ph_sweep(const XbArgs &a, int32_t L, int bid, int nb, bool from_levels, bool stats,
                         XbWaveLds *lds, bool fresh = false)
{
    __shared__ int s_wc[16][XB_BINS];
    __shared__ int s_base[XB_BINS];
    uint64_t *F = xb_f(a, L);
    const int qb = L % XB_QBUFS;
    int32_t *const qd[XB_BINS] = {xb_q(a, qb, 0), xb_q(a, qb, 1), xb_q(a, qb, 2), xb_q(a, qb, 3)};
    XbCounter *gcnt = a.ctl->qcnt[qb];
    const int ln = xb_lane(), wv = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    XbWaveLds *wl = lds + wv;
    const int64_t gw = ((int64_t)bid * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwv = ((int64_t)nb * blockDim.x) >> 6;
    const int64_t nchunks = (a.nwords + XB_CHUNK_WORDS - 1) / XB_CHUNK_WORDS;
    unsigned long long my_n = 0, my_m = 0;
    int wc[XB_BINS] = {0, 0, 0, 0};
    // pass 1: bits (from the levels when asked), statistics, bin counts.
    // Chunks in groups of 4 as in ph_bu: one round trip loads the F_L words
    // of 4 chunks (lane 16q + k: word k of chunk q); empty chunks are skipped.
    for (int64_t cg = gw; cg < nchunks; cg += 4 * nwv) {
    uint64_t pref = 0;
    if (!from_levels) {
        const int64_t chl = cg + (int64_t)(ln >> 4) * nwv;
        const int64_t wq = chl * XB_CHUNK_WORDS + (ln & 15);
        if (chl < nchunks && wq < a.nwords) pref = F[wq];
    }
    const unsigned long long nzm = from_levels ? ~0ull : __ballot(pref != 0ull);
    for (int q = 0; q < 4; q++) {
        const int64_t ch = cg + (int64_t)q * nwv;
        if (ch >= nchunks) break;
        if (((nzm >> (16 * q)) & 0xffffull) == 0ull) continue;
        const int64_t w0 = ch * XB_CHUNK_WORDS;
        const int64_t w = w0 + ln;
        uint64_t myword = 0;
        if (from_levels) {
            // all 16 level words of the chunk in flight at once (each load
            // instruction reads 64 consecutive levels)
            int32_t lv[XB_CHUNK_WORDS];
#pragma unroll
            for (int k = 0; k < XB_CHUNK_WORDS; k++) {
                const int64_t v = (w0 + k) * 64 + ln;
                lv[k] = v < a.n ? a.level[v] : -2;
            }
#pragma unroll
            for (int k = 0; k < XB_CHUNK_WORDS; k++) {
                const unsigned long long f = __ballot(lv[k] == L);
                if (ln == k) myword = f;
            }
            if (ln < XB_CHUNK_WORDS && w < a.nwords) {
                F[w] = myword;
                if (myword) a.vis[w] |= myword;
            }
        } else {
            const uint64_t pq = __shfl(pref, 16 * q + (ln & 15));
            myword = ln < XB_CHUNK_WORDS ? pq : 0ull;
        }
        const int total = xb_chunk_list(myword, w0, wl);
        for (int r = 0; r < total; r += 64) {
            const int i = r + ln;
            int bin = -1;
            if (i < total) {
                const int32_t v = wl->list[i];
                const int64_t d = a.beg[v + 1] - a.beg[v];
                my_n += 1;
                my_m += (unsigned long long)d;
                if (d > 0) bin = xb_bin_of(a, d);
            }
#pragma unroll
            for (int k = 0; k < XB_BINS; k++) wc[k] += (int)__popcll(__ballot(bin == k));
        }
        xb_wave_sync();
    }
    }
    // reserve: one atomic per bin per block
    if (ln == 0)
        for (int k = 0; k < XB_BINS; k++) s_wc[wv][k] = wc[k];
    __syncthreads();
    if (threadIdx.x < XB_BINS) {
        const int k = threadIdx.x;
        int run = 0;
        for (int i = 0; i < nw; i++) {
            const int c = s_wc[i][k];
            s_wc[i][k] = run;
            run += c;
        }
        s_base[k] = run ? atomicAdd(&gcnt[k].v, run) : 0;
    }
    __syncthreads();
    int pos[XB_BINS];
    for (int k = 0; k < XB_BINS; k++) pos[k] = s_base[k] + s_wc[wv][k];
    // pass 2: write the queue entries (same chunk groups)
    for (int64_t cg = gw; cg < nchunks; cg += 4 * nwv) {
    uint64_t pref = 0;
    {
        const int64_t chl = cg + (int64_t)(ln >> 4) * nwv;
        const int64_t wq = chl * XB_CHUNK_WORDS + (ln & 15);
        if (chl < nchunks && wq < a.nwords) pref = F[wq];
    }
    const unsigned long long nzm = __ballot(pref != 0ull);
    for (int q = 0; q < 4; q++) {
        if (((nzm >> (16 * q)) & 0xffffull) == 0ull) continue;
        const int64_t w0 = (cg + (int64_t)q * nwv) * XB_CHUNK_WORDS;
        const uint64_t pq = __shfl(pref, 16 * q + (ln & 15));
        const uint64_t myword = ln < XB_CHUNK_WORDS ? pq : 0ull;
        const int total = xb_chunk_list(myword, w0, wl);
        for (int r = 0; r < total; r += 64) {
            const int i = r + ln;
            int bin = -1;
            int32_t v = 0;
            if (i < total) {
                v = wl->list[i];
                const int64_t d = a.beg[v + 1] - a.beg[v];
                if (d > 0) bin = xb_bin_of(a, d);
            }
#pragma unroll
            for (int k = 0; k < XB_BINS; k++) {
                const unsigned long long m = __ballot(bin == k);
                if (bin == k) qd[k][pos[k] + xb_rank(m)] = v;
                pos[k] += (int)__popcll(m);
            }
        }
        xb_wave_sync();
    }
    }
    if (stats) {
        // fresh: the sweep slot of level L was zeroed at level L-1 and only
        // this sweep writes it, so the stores need not read it first
        unsigned long long vv[2] = {my_n, my_m};
        const int ff[2] = {xb_f_sweep(L, 0), xb_f_sweep(L, 1)};
        long long old[2] = {0, 0}, tot[2];
        xb_part_add<2>(a, bid, vv, ff, fresh, old, tot);
    }
}

// ---------------------------------------------------------------------------
// top-down expansion
// ---------------------------------------------------------------------------
struct XbTdCtx {
    const int32_t *qc[XB_BINS];
    int32_t nc[XB_BINS];
    int32_t *qn[XB_BINS];
    XbCounter *cn;
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
        int64_t rb = 0, re = 0;
        if (SF) {
            // the row bounds are requested with the claim, so the two round
            // trips overlap (only the winner uses them)
            rb = a.beg[v];
            re = a.beg[v + 1];
            won = atomicCAS(&a.level[v], -1, t.L1) == -1;
        } else {
            __builtin_nontemporal_store(t.L1, &a.level[v]);
            won = true; // every writer stores the same level; any parent is valid
        }
        if (won) {
            if (a.parent) a.parent[v] = u;
            if (SF) {
                // single-scan sets no bits: the sweep that must follow it
                // rebuilds F_{L+1} and the visited bits from the levels
                const uint64_t bit = 1ull << (v & 63);
                atomicOr((unsigned long long *)&a.vis[v >> 6], (unsigned long long)bit);
                atomicOr((unsigned long long *)&t.Fn[v >> 6], (unsigned long long)bit);
                d = re - rb;
                my_n += 1;
                my_m += (unsigned long long)d;
            }
        }
    }
    if (SF && t.enqueue) {
        // wave-aggregated appends to the 4 degree-bin queues: the first lane of
        // each bin (an active lane: callers run this inside loops that some
        // lanes have left) reserves the wave's slots of that bin, the up to 4
        // atomics in one instruction (one round trip rather than one per
        // non-empty bin)
        const int b = won && d > 0 ? xb_bin_of(a, d) : -1;
        const unsigned long long m0 = __ballot(b == 0), m1 = __ballot(b == 1);
        const unsigned long long m2 = __ballot(b == 2), m3 = __ballot(b == 3);
        if ((m0 | m1 | m2 | m3) != 0ull) {
            const int ln = xb_lane();
            const unsigned long long mb = b == 0 ? m0 : b == 1 ? m1 : b == 2 ? m2 : b == 3 ? m3 : 0ull;
            const int lead = mb != 0ull ? __ffsll(mb) - 1 : 0;
            int r = 0;
            if (b >= 0 && ln == lead) r = atomicAdd(&t.cn[b].v, (int)__popcll(mb));
            const int base = __shfl(r, lead);
            if (b >= 0) {
                // the bins' queues are consecutive n-entry slices (xb_q)
                t.qn[0][(int64_t)b * a.n + base + xb_rank(mb)] = v;
            }
        }
    }
}

// Visits the rows of cnt <= min(XB_GCAP, 2 * blockDim.x) frontier vertices sel(0..cnt-1) as one
// concatenated adjacency shared by every thread of the active grid: each block
// scans their degrees in LDS (with their row starts), and a thread maps an
// edge position to its vertex by binary search.  Every thread of the block
// calls it (block barriers inside).
template <bool SF, typename Sel>
__device__ void td_concat(const XbArgs &a, const XbTdCtx &t, int32_t cnt, Sel sel, XbSmem &sm,
                          int64_t gtid, int64_t nthr, unsigned long long &my_n,
                          unsigned long long &my_m)
{
    long long *s_pref = sm.td.pref, *s_b = sm.td.b;
    int32_t *s_u = sm.td.u;
    // two elements per thread (cnt <= 2 * blockDim.x)
    const int k0 = 2 * threadIdx.x, k1 = k0 + 1;
    long long d0 = 0, d1 = 0;
    if (k0 < cnt) {
        const int32_t u = sel(k0);
        const int64_t b = a.beg[u];
        s_u[k0] = u;
        s_b[k0] = b;
        d0 = a.beg[u + 1] - b;
    }
    if (k1 < cnt) {
        const int32_t u = sel(k1);
        const int64_t b = a.beg[u];
        s_u[k1] = u;
        s_b[k1] = b;
        d1 = a.beg[u + 1] - b;
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
        td_visit<SF>(a, t, s_u[lo], a.adj[s_b[lo] + (p - s_pref[lo])], true, my_n, my_m);
    }
    __syncthreads();
}

// carry (persistent kernel, LDS; NULL elsewhere): what this block stored in
// the (count, degree sum) fields of ring slot L+1 at level L-1 (discoveries
// two levels ahead; none is made, so it is zero), so the counts of level L are
// stored without reading the column back; set to zero here.
template <bool SF>
__device__ void ph_td(const XbArgs &a, int32_t L, int bid, int nb, bool enqueue, XbSmem &sm,
                      long long *carry = nullptr)
{
    XbTdCtx t;
    const int qc = L % XB_QBUFS, qn = (L + 1) % XB_QBUFS;
    for (int k = 0; k < XB_BINS; k++) {
        t.qc[k] = xb_q(a, qc, k);
        t.nc[k] = xb_ld_i32(&a.ctl->qcnt[qc][k].v);
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

    // A frontier of at most `cap` vertices, whatever their bins, is one
    // concatenated adjacency for the whole grid (no lane or wave walks a row
    // serially, so a level with a few hubs or a few dozen mid-degree vertices
    // takes about one round of dependent loads); otherwise the grid bin is,
    // `cap` vertices per pass, and the other bins follow.  cap is at most two
    // per thread of the block (512 in the per-phase kernels' 256-thread
    // blocks): XB_GCAP when the whole grid runs the level, XB_GCAP_SMALL on
    // small-frontier levels, where only XB_SMALL_BLOCKS blocks share the
    // edges and each block's scan of the pass's degrees weighs more (on the
    // 4096^2 grid a 256 cap there is 4 % faster than 2048, 1 % than 512).
    const int32_t nall = t.nc[0] + t.nc[1] + t.nc[2] + t.nc[3];
    const int32_t cap = min(nb > XB_SMALL_BLOCKS ? XB_GCAP : XB_GCAP_SMALL, 2 * (int)blockDim.x);
    const bool flat = nall <= cap;
    const int32_t c0 = t.nc[0], c1 = c0 + t.nc[1], c2 = c1 + t.nc[2];
    const int32_t ncat = flat ? nall : t.nc[3];
    for (int32_t base = 0; base < ncat; base += cap) {
        td_concat<SF>(a, t, min(cap, ncat - base), [&](int32_t k) {
            if (!flat) return t.qc[3][base + k];
            return k < c0 ? t.qc[0][k] : k < c1 ? t.qc[1][k - c0] : k < c2 ? t.qc[2][k - c1]
                                                                            : t.qc[3][k - c2];
        }, sm, gtid, nthr, my_n, my_m);
    }
    if (!flat) {
        // block bin: one block per vertex
        for (int32_t i = bid; i < t.nc[2]; i += nb) {
            const int32_t u = t.qc[2][i];
            const int64_t b = a.beg[u], e = a.beg[u + 1];
            for (int64_t j = b + threadIdx.x; j < e; j += blockDim.x)
                td_visit<SF>(a, t, u, a.adj[j], true, my_n, my_m);
        }
        // Wave and thread bins: work is dealt to blocks round robin (wave-bin
        // vertex i -> block i mod nb; thread-bin entries in runs of 64, run r ->
        // block r mod nb, so a wave still takes 64 consecutive entries), so a
        // frontier of a few thousand vertices spreads over every CU instead of
        // filling the first few.
        // wave bin: one wave per vertex
        const int64_t gwi = (int64_t)(threadIdx.x >> 6) * nb + bid;
        for (int64_t i = gwi; i < t.nc[1]; i += nthr >> 6) {
            const int32_t u = t.qc[1][i];
            const int64_t b = a.beg[u], e = a.beg[u + 1];
            for (int64_t j = b + ln; j < e; j += XB_WAVE)
                td_visit<SF>(a, t, u, a.adj[j], true, my_n, my_m);
        }
        // thread bin: one lane per vertex
        const int64_t gti = (((int64_t)(threadIdx.x >> 6) * nb + bid) << 6) + ln;
        for (int64_t i = gti; i < t.nc[0]; i += nthr) {
            const int32_t u = t.qc[0][i];
            const int64_t b = a.beg[u], e = a.beg[u + 1];
            for (int64_t j = b; j < e; j++) td_visit<SF>(a, t, u, a.adj[j], true, my_n, my_m);
        }
    }
    if (SF) {
        unsigned long long vv[2] = {my_n, my_m};
        const int ff[2] = {xb_f_ring(L + 1, 0), xb_f_ring(L + 1, 1)};
        long long old[2] = {carry ? carry[0] : 0, carry ? carry[1] : 0}, tot[2];
        xb_part_add<2>(a, bid, vv, ff, carry != nullptr, old, tot);
    }
    if (carry && threadIdx.x == 0) {
        carry[0] = 0;
        carry[1] = 0;
    }
}

// ---------------------------------------------------------------------------
// bottom-up
// ---------------------------------------------------------------------------
// carry: as for ph_td (this phase makes no +2 discoveries either: a
// candidate without a level-L neighbour stays unvisited for the next level).
// Level (and top-down single-scan level) stores are nontemporal: each entry is
// written once per traversal and not read back by its writer, so it need not
// sit dirty in L2 until the level's release writes it back.
__device__ void ph_bu(const XbArgs &a, int32_t L, int bid, int nb, XbWaveLds *lds,
                      long long *carry = nullptr)
{
    const uint64_t *F = xb_f(a, L);
    uint64_t *F1 = xb_f(a, L + 1);
    const uint32_t *F32 = reinterpret_cast<const uint32_t *>(F);
    const int32_t L1 = L + 1;
    const int ln = xb_lane();
    XbWaveLds *wl = lds + (threadIdx.x >> 6);
    const int64_t gw = ((int64_t)bid * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwv = ((int64_t)nb * blockDim.x) >> 6;
    const int64_t nchunks = (a.nwords + XB_CHUNK_WORDS - 1) / XB_CHUNK_WORDS;
    const uint64_t below = ln ? (~0ull >> (64 - ln)) : 0ull;
    unsigned long long n1 = 0, m1 = 0;
    // work counters: candidates probed and rows scanned (lane 0, wave totals),
    // adjacency words read by this lane's scans
    uint32_t w_cand = 0, w_scanv = 0, w_scane = 0;
    // The wave's chunks gw, gw + W, gw + 2W, ... in groups of 4: lane 16q + k
    // loads word k of the group's chunk q, so one memory round trip covers 4
    // chunks and empty chunks are skipped by one ballot.
    for (int64_t cg = gw; cg < nchunks; cg += 4 * nwv) {
    uint64_t pref = 0;
    {
        const int64_t chl = cg + (int64_t)(ln >> 4) * nwv;
        const int64_t wq = chl * XB_CHUNK_WORDS + (ln & 15);
        if (chl < nchunks && wq < a.nwords) pref = a.nz[wq] & ~a.vis[wq];
    }
    const unsigned long long nzm = __ballot(pref != 0ull);
    for (int q = 0; q < 4; q++) {
        if (((nzm >> (16 * q)) & 0xffffull) == 0ull) continue;
        const int64_t ch = cg + (int64_t)q * nwv;
        const int64_t w0 = ch * XB_CHUNK_WORDS;
        const int64_t w = w0 + ln;
        // lane k < 16: unvisited non-isolated vertices of word w0 + k
        const uint64_t pq = __shfl(pref, 16 * q + (ln & 15));
        const uint64_t cand = ln < XB_CHUNK_WORDS ? pq : 0ull;
        const int pc = (int)__popcll(cand);
        const int inc = xb_wave_incl_scan_i32(pc);
        const int total = __shfl(inc, 63);
        if (total == 0) continue;
        if (ln == 0) w_cand += (uint32_t)total;
        const int off = inc - pc;
        // the chunk's candidates in ascending order
        for (int k = 0; k < XB_CHUNK_WORDS; k++) {
            const uint64_t word = __shfl(cand, k);
            const int o = __shfl(off, k);
            if ((word >> ln) & 1ull)
                wl->list[o + (int)__popcll(word & below)] = (int32_t)((w0 + k) * 64 + ln);
        }
        if (ln < XB_CHUNK_WORDS) {
            wl->f1[ln] = 0ull;
            wl->va[ln] = 0ull;
        }
        xb_wave_sync();
        // Hub probe, 4 candidates per lane per round (independent loads): a
        // candidate whose largest-degree neighbour is in F_L gets level L+1
        // with that neighbour as parent, without reading its adjacency.  For
        // a degree-1 candidate the hub is its only neighbour, so a miss
        // settles its question for this level and it needs no scan.  The
        // others are compacted to the front of the list (positions below the
        // entries read so far) for the scan below.
        int nres = 0;
        for (int r = 0; r < total; r += 256) {
            int32_t vc[4];
            int2 h[4];
            bool hit[4], done[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int i = r + 64 * t + ln;
                vc[t] = i < total ? wl->list[i] : -1;
            }
            // hub records (and row starts below) are read once per level:
            // non-temporal loads on large graphs (a.nt_bu, XB_NT_HUB_BYTES);
            // the adjacency words, whose lines short rows share, are plain
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const long long *hp = reinterpret_cast<const long long *>(&a.hd[vc[t] >= 0 ? vc[t] : 0]);
                const long long hv = vc[t] < 0 ? (long long)0xffffffffll // {-1, 0}
                                     : a.nt_bu ? __builtin_nontemporal_load(hp) : *hp;
                h[t] = make_int2((int32_t)(hv & 0xffffffffll), (int32_t)(hv >> 32));
            }
            // 32-bit probes: the dword of the bitmap word that holds the bit
            uint32_t fw[4];
#pragma unroll
            for (int t = 0; t < 4; t++) fw[t] = h[t].x >= 0 ? F32[h[t].x >> 5] : 0u;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                hit[t] = (fw[t] >> (h[t].x & 31)) & 1u;
                done[t] = hit[t] || h[t].y == 1;
                if (hit[t]) {
                    const int32_t v = vc[t];
                    const unsigned long long bit = 1ull << (v & 63);
                    const int li = (int)((v >> 6) - w0);
                    __builtin_nontemporal_store(L1, &a.level[v]);
                    if (a.parent) a.parent[v] = h[t].x;
                    atomicOr(&wl->f1[li], bit);
                    atomicOr(&wl->va[li], bit);
                    n1 += 1;
                    m1 += (unsigned long long)h[t].y;
                }
            }
            xb_wave_sync();
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const bool un = vc[t] >= 0 && !done[t];
                const unsigned long long um = __ballot(un);
                if (un) wl->list[nres + xb_rank(um)] = vc[t];
                nres += (int)__popcll(um);
            }
            xb_wave_sync();
        }
        // Each lane works on its own candidate, 4 neighbours per step; a lane
        // whose candidate is settled takes the next one from the list, so a
        // long scan in one lane does not hold the others.  At most
        // XB_BU_LANE_MAX edges per lane: a candidate still unresolved with
        // edges left is deferred to the front of the list (positions < the
        // consumed count, so no unread entry is overwritten) and finished by
        // the whole wave below.  When every row starts with its hub (rows
        // sorted by neighbour degree), the scan starts after it: the hub pass
        // probed it.  Scans probe F_L only: a candidate with no level-L
        // neighbour stays unvisited and is found at the next level.  (Probing
        // F_{L+1} as well, to settle such candidates at L+2 now, costs a
        // dependent round trip per step of every full scan, more than it
        // saves at the next level.)
        const int64_t skip = a.hub_first ? 1 : 0;
        if (ln == 0) w_scanv += (uint32_t)nres;
        int ndef = 0, next = 0;
        int32_t v = -1;
        int64_t j = 0, ce = 0, e = 0, b = 0;
        for (;;) {
            const bool idle = v < 0;
            const unsigned long long im = __ballot(idle);
            if (idle) {
                const int k = next + xb_rank(im);
                if (k < nres) {
                    v = wl->list[k];
                    b = a.nt_bu ? __builtin_nontemporal_load(&a.beg[v]) : a.beg[v];
                    e = a.nt_bu ? __builtin_nontemporal_load(&a.beg[v + 1]) : a.beg[v + 1];
                    j = b + skip;
                    ce = (e - j > XB_BU_LANE_MAX) ? j + XB_BU_LANE_MAX : e;
                }
            }
            next = min(nres, next + (int)__popcll(im));
            if (__ballot(v >= 0) == 0ull) break;
            bool defer = false;
            const int32_t vv = v;
            if (v >= 0) {
                const int32_t u0 = a.adj_bu[j];
                const int32_t u1 = j + 1 < ce ? a.adj_bu[j + 1] : -1;
                const int32_t u2 = j + 2 < ce ? a.adj_bu[j + 2] : -1;
                const int32_t u3 = j + 3 < ce ? a.adj_bu[j + 3] : -1;
                const bool h0 = (F32[u0 >> 5] >> (u0 & 31)) & 1u;
                const bool h1 = u1 >= 0 && ((F32[u1 >> 5] >> (u1 & 31)) & 1u);
                const bool h2 = u2 >= 0 && ((F32[u2 >> 5] >> (u2 & 31)) & 1u);
                const bool h3 = u3 >= 0 && ((F32[u3 >> 5] >> (u3 & 31)) & 1u);
                w_scane += 1u + (u1 >= 0) + (u2 >= 0) + (u3 >= 0);
                const unsigned long long bit = 1ull << (v & 63);
                const int li = (int)((v >> 6) - w0);
                if (h0 | h1 | h2 | h3) {
                    __builtin_nontemporal_store(L1, &a.level[v]);
                    if (a.parent) a.parent[v] = h0 ? u0 : (h1 ? u1 : (h2 ? u2 : u3));
                    atomicOr(&wl->f1[li], bit);
                    atomicOr(&wl->va[li], bit);
                    n1 += 1;
                    m1 += (unsigned long long)(e - b);
                    v = -1;
                } else {
                    j += 4;
                    if (j >= ce) {
                        defer = ce < e; // unscanned edges remain: the wave finishes them
                        v = -1;
                    }
                }
            }
            const unsigned long long dm = __ballot(defer);
            if (dm) {
                xb_wave_sync();
                if (defer) wl->list[ndef + xb_rank(dm)] = vv;
                ndef += (int)__popcll(dm);
                xb_wave_sync();
            }
        }
        // The deferred candidates, one at a time by the whole wave: 256 edges
        // per round (4 coalesced adjacency words per lane), a ballot ends the
        // scan at the first level-L neighbour; one without such a neighbour
        // stays unvisited and is found at the next level.
        for (int d = 0; d < ndef; d++) {
            const int32_t dv = wl->list[d];
            const int64_t db = a.beg[dv], de = a.beg[dv + 1];
            int32_t found = -1;
            for (int64_t base = db + skip + XB_BU_LANE_MAX; base < de; base += 256) {
                int32_t u[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int64_t jj = base + ln + 64 * t;
                    u[t] = jj < de ? a.adj_bu[jj] : -1;
                }
                bool h[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    h[t] = u[t] >= 0 && ((F32[u[t] >> 5] >> (u[t] & 31)) & 1u);
                    w_scane += u[t] >= 0;
                }
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const unsigned long long m = __ballot(h[t]);
                    if (m) {
                        found = __shfl(u[t], __ffsll(m) - 1);
                        break;
                    }
                }
                if (found >= 0) break;
            }
            if (found >= 0 && ln == 0) {
                const unsigned long long bit = 1ull << (dv & 63);
                const int li = (int)((dv >> 6) - w0);
                __builtin_nontemporal_store(L1, &a.level[dv]);
                if (a.parent) a.parent[dv] = found;
                atomicOr(&wl->f1[li], bit);
                atomicOr(&wl->va[li], bit);
                n1 += 1;
                m1 += (unsigned long long)(de - db);
            }
        }
        xb_wave_sync();
        // the wave owns the chunk's words (one cache line per bitmap) for
        // the whole level; atomic ORs whose results are not used cost no
        // round trip, where a load-modify-store would stall the wave on the
        // load before its next chunk
        if (ln < XB_CHUNK_WORDS && w < a.nwords) {
            const unsigned long long va = wl->va[ln];
            if (va) {
                atomicOr((unsigned long long *)&a.vis[w], va);
                const unsigned long long f1 = wl->f1[ln];
                if (f1) atomicOr((unsigned long long *)&F1[w], f1);
            }
        }
    }
    }
    // the work fields of this level start at zero (zeroing schedule) and only
    // this phase writes them
    unsigned long long vv[5] = {n1, m1, w_cand, w_scanv, w_scane};
    const int ff[5] = {xb_f_ring(L1, 0), xb_f_ring(L1, 1), xb_f_bu(L, 0), xb_f_bu(L, 1),
                       xb_f_bu(L, 2)};
    long long old[5] = {carry ? carry[0] : 0, carry ? carry[1] : 0, 0, 0, 0}, tot[5];
    xb_part_add<5>(a, bid, vv, ff, carry != nullptr, old, tot);
    if (carry && threadIdx.x == 0) {
        carry[0] = 0;
        carry[1] = 0;
    }
}

// ---------------------------------------------------------------------------
// level loop (shared by both drivers through these helpers)
// ---------------------------------------------------------------------------
// Level bookkeeping shared by both drivers.  Returns false when the
// traversal is complete (no frontier edges, or every vertex it can reach
// already has its level: a.n_target); otherwise fills the unvisited counts of
// level L, which count every non-isolated vertex as reachable.
__host__ __device__ inline bool xb_level_open(XbLoop &st, const XbArgs &a, long long n_f,
                                              long long m_f, long long pend_n, long long pend_m,
                                              long long *n_u, long long *m_u)
{
    st.cum_n += n_f;
    st.cum_m += m_f;
    if (m_f == 0 || st.cum_n + pend_n >= a.n_target) return false;
    *n_u = a.n_nz - (st.cum_n + pend_n);
    *m_u = a.m - (st.cum_m + pend_m);
    return true;
}

// LDS position of packed MLP parameter k (packed as in xbfs_policy.h): the
// weight matrices are stored transposed (W1T[i][o], W2T[i][o], W3T[i][o]),
// so lanes o = 0..31 reading input i touch 32 consecutive words (no bank
// conflicts); biases keep their places.
__device__ __forceinline__ int xb_mlp_lds_pos(int k)
{
    const int nw1 = XB_MLP_NH * XB_MLP_NI, o2 = nw1 + XB_MLP_NH;
    const int nw2 = XB_MLP_NH * XB_MLP_NH, o3 = o2 + nw2 + XB_MLP_NH;
    if (k < nw1) return (k % XB_MLP_NI) * XB_MLP_NH + k / XB_MLP_NI;
    if (k >= o2 && k < o2 + nw2) return o2 + ((k - o2) % XB_MLP_NH) * XB_MLP_NH + (k - o2) / XB_MLP_NH;
    if (k >= o3 && k < o3 + XB_MLP_NO * XB_MLP_NH)
        return o3 + ((k - o3) % XB_MLP_NH) * XB_MLP_NO + (k - o3) / XB_MLP_NH;
    return k;
}

// MLP forward by one wave from LDS weights p (laid out by xb_mlp_lds_pos):
// lane o < 32 computes hidden unit o, the hidden vectors move between lanes
// by shuffles, lanes 0..2 write the outputs to y.  Sums run in input order,
// as in xbfs_mlp_forward1.  Called by every lane of the wave.
__device__ void xb_mlp_wave_forward(const float *p, const float *x, float *y)
{
    const float *W1T = p, *b1 = W1T + XB_MLP_NH * XB_MLP_NI;
    const float *W2T = b1 + XB_MLP_NH, *b2 = W2T + XB_MLP_NH * XB_MLP_NH;
    const float *W3T = b2 + XB_MLP_NH, *b3 = W3T + XB_MLP_NO * XB_MLP_NH;
    const int o = xb_lane();
    const int oh = o < XB_MLP_NH ? o : 0, oo = o < XB_MLP_NO ? o : 0;
    // fully unrolled: every weight load and every lane read is issued up front
    float t1 = b1[oh];
#pragma unroll
    for (int i = 0; i < XB_MLP_NI; i++) t1 += W1T[i * XB_MLP_NH + oh] * x[i];
    const float h1 = (o < XB_MLP_NH && t1 > 0.0f) ? t1 : 0.0f;
    float t2 = b2[oh];
#pragma unroll
    for (int i = 0; i < XB_MLP_NH; i++) t2 += W2T[i * XB_MLP_NH + oh] * __shfl(h1, i);
    const float h2 = (o < XB_MLP_NH && t2 > 0.0f) ? t2 : 0.0f;
    float t3 = b3[oo];
#pragma unroll
    for (int i = 0; i < XB_MLP_NH; i++) t3 += W3T[i * XB_MLP_NO + oo] * __shfl(h2, i);
    if (o < XB_MLP_NO) y[o] = t3;
}

// The level loop's block-uniform state lives in LDS (thread 0 updates it,
// everyone reads it after a barrier) so it does not hold scalar registers
// across the phases, which are register-bound.

__global__ void __launch_bounds__(XB_PERSIST_BLOCK) k_persistent(XbArgs a)
{
    const int bid = blockIdx.x, nb = gridDim.x;
    __shared__ XbSmem sm;
    __shared__ XbLoopShared ls;
    __shared__ float s_mlp[XB_MLP_NPARAM]; // learned policy weights, loaded once (transposed)
    __shared__ long long s_carry[2];       // see ph_td / ph_bu
    extern __shared__ __align__(16) unsigned char xb_dyn_lds[];
    XbWaveLds *lds = reinterpret_cast<XbWaveLds *>(xb_dyn_lds);
    if (a.mlp)
        for (int i = threadIdx.x; i < XB_MLP_NPARAM; i += blockDim.x) s_mlp[xb_mlp_lds_pos(i)] = a.mlp[i];
    ph_init(a, bid, nb);
    if (!xb_first_sync(a.ctl)) return;
    if (a.nsrc > 0) {
        ph_seed(a, bid, nb);
        if (!xb_grid_sync(a.ctl)) return;
    }
    if (threadIdx.x == 0) {
        s_carry[0] = 0;
        s_carry[1] = 0;
        xb_loop_init(&ls.st);
        if (a.nsrc > 0) ls.st.q_valid = 0; // the level-0 queue comes from a sweep of F_0
    }
    __syncthreads();
    XbCtl *c = a.ctl;
    // blocks taking part in the current level: the grid, or the first
    // XB_SMALL_BLOCKS during small-frontier levels (the rest park)
    int nact = nb;
    const bool small_ok = nb > XB_SMALL_BLOCKS;
    for (;;) {
        const int32_t L = ls.st.L;
        const unsigned long long t0 = wall_clock64();
        const bool need_sweep = ls.st.need_sweep != 0;
        if (need_sweep) {
            ph_sweep(a, L, bid, nact, true, true, lds, true);
            if (!xb_level_sync(c, nact)) return;
        }
        // Level statistics and the decision, by wave 0 alone, while the
        // other waves zero what level L+1 will write; one block barrier
        // publishes the decision.  Blocks that sit out small-frontier levels
        // and rejoin later skip this level's zeroing: the active blocks did it.
        // Wave 1 of block 0 meanwhile sums the previous level's bottom-up work
        // into its record (zeros after a top-down level).
        if (threadIdx.x >= 128) ph_level_prep(a, L, bid, nact, 128);
        if (threadIdx.x >= 64 && threadIdx.x < 128 && bid == 0 && L > 0 && L - 1 < a.rec_cap) {
            long long wv[3] = {0, 0, 0};
            if (ls.st.prev_mode == XBFS_MODE_BU) {
                const int fw[3] = {xb_f_bu(L - 1, 0), xb_f_bu(L - 1, 1), xb_f_bu(L - 1, 2)};
                xb_part_sum<3>(a, fw, wv, nact, 1);
            }
            if (threadIdx.x == 64) {
                XbDevRec &pr = a.recs[L - 1];
                pr.bu_cand = wv[0];
                pr.bu_scan_v = wv[1];
                pr.bu_scan_e = wv[2];
            }
        }
        if (threadIdx.x < 64) {
            long long sv[5];
            const int ff[5] = {need_sweep ? xb_f_sweep(L, 0) : xb_f_ring(L, 0),
                               need_sweep ? xb_f_sweep(L, 1) : xb_f_ring(L, 1),
                               xb_f_ring(L + 1, 0), xb_f_ring(L + 1, 1), xb_f_ring(L + 1, 2)};
            xb_part_sum<5>(a, ff, sv, nact);
            if (threadIdx.x == 0) {
                long long n_u = 0, m_u = 0;
                const bool open = xb_level_open(ls.st, a, sv[0], sv[1], sv[2], sv[3], &n_u, &m_u);
                ls.open = open ? 1 : 0;
                ls.n_f = sv[0];
                ls.m_f = sv[1];
                ls.n_u = n_u;
                ls.m_u = m_u;
                ls.p2_next = sv[4];
                if (!open || L >= a.max_levels) {
                    if (bid == 0) {
                        const long long pend_n = sv[2], pend_m = sv[3];
                        xb_st(&c->n_levels, L);
                        xb_st(&c->max_level, pend_n > 0 ? L + 1 : (sv[0] > 0 ? L : L - 1));
                        xb_st(&c->v_cc, (long long)(ls.st.cum_n + pend_n));
                        xb_st(&c->m_cc, (long long)(ls.st.cum_m + pend_m));
                        if (open) xb_st(&c->error, 2);
                    }
                    ls.open = 0;
                }
            }
            xb_wave_sync();
            if (ls.open) {
                const float *pred = nullptr;
                if (a.mlp && L <= XBFS_MLP_MAX_LEVEL) {
                    if (threadIdx.x == 0)
                        xbfs_mlp_features(L, ls.n_f, ls.m_f, ls.n_u, ls.m_u, ls.st.prev_mode,
                                          sm.mlp.x);
                    xb_wave_sync();
                    xb_mlp_wave_forward(s_mlp, sm.mlp.x, sm.mlp.y);
                    xb_wave_sync();
                    pred = sm.mlp.y;
                }
                if (threadIdx.x == 0) {
                    int32_t htag = 0, used = 0;
                    ls.mode = xbfs_choose_mode(a.policy, a.sched, a.sched_len, pred, &ls.st.hs, L,
                                               ls.m_f, a.m, a.alpha, &htag, &used);
                    ls.htag = htag;
                    ls.used = used;
                    ls.small = small_ok && ls.mode == XBFS_MODE_TD_SF && ls.st.q_valid &&
                               !need_sweep && ls.p2_next == 0 && ls.n_f <= XB_SMALL_NF &&
                               ls.m_f <= XB_SMALL_MF;
                }
            }
        }
        __syncthreads();
        // entering / leaving small-frontier levels (every active block takes
        // the same branch: the inputs are identical)
        if (!ls.open || (nact < nb && !ls.small)) {
            if (nact < nb) {
                // the parked blocks rejoin: they get the loop state, decision
                // included, from device memory after the grid barrier
                if (bid == 0 && threadIdx.x == 0) *a.lsg = ls;
                if (!xb_grid_sync(c)) return;
                nact = nb;
            }
            if (!ls.open) return;
        } else if (nact == nb && ls.small) {
            if (bid >= XB_SMALL_BLOCKS) {
                if (!xb_park(c)) return;
                if (threadIdx.x == 0) {
                    ls = *a.lsg;
                    s_carry[0] = 0; // levels entered small mode without +2 discoveries
                    s_carry[1] = 0;
                }
                __syncthreads();
                if (!ls.open) return;
            } else {
                nact = XB_SMALL_BLOCKS;
            }
        }
        // the level itself, on the active blocks; a parked block that just
        // rejoined takes the level from the copied state
        const int32_t Lc = ls.st.L;
        const int32_t mode = ls.mode;
        if (mode != XBFS_MODE_BU && !ls.st.q_valid && !ls.st.need_sweep) {
            ph_sweep(a, Lc, bid, nact, false, false, lds);
            if (!xb_level_sync(c, nact)) return;
        }
        if (mode == XBFS_MODE_TD_SF) ph_td<true>(a, Lc, bid, nact, ls.p2_next == 0, sm, s_carry);
        else if (mode == XBFS_MODE_TD_SS) ph_td<false>(a, Lc, bid, nact, false, sm, s_carry);
        else ph_bu(a, Lc, bid, nact, lds, s_carry);
        if (!xb_level_sync(c, nact)) return;
        if (threadIdx.x == 0) {
            if (bid == 0) {
                if (nact < nb) xb_st(&c->heartbeat, (uint32_t)Lc);
                if (Lc < a.rec_cap) {
                    XbDevRec &r = a.recs[Lc];
                    r.level = Lc;
                    r.mode = mode;
                    r.tag = xbfs_mode_to_tag(mode);
                    r.htag = (ls.htag & 0xff) | (ls.used << 8);
                    r.n_f = ls.n_f;
                    r.m_f = ls.m_f;
                    r.n_u = ls.n_u;
                    r.m_u = ls.m_u;
                    r.t0 = (long long)t0;
                    r.t1 = (long long)wall_clock64();
                }
            }
            ls.st.q_valid = (mode == XBFS_MODE_TD_SF && ls.p2_next == 0);
            ls.st.need_sweep = (mode == XBFS_MODE_TD_SS);
            ls.st.prev_mode = mode;
            ls.st.L = Lc + 1;
        }
        __syncthreads();
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
    extern __shared__ __align__(16) unsigned char xb_dyn_lds[];
    XbWaveLds *lds = reinterpret_cast<XbWaveLds *>(xb_dyn_lds);
    const int bid = blockIdx.x, nb = gridDim.x;
    XbCtl *c = a.ctl;
    ph_level_prep(a, L, bid, nb);
    if (mode != XBFS_MODE_BU && !q_valid && !need_sweep) {
        ph_sweep(a, L, bid, nb, false, false, lds);
        if (!xb_grid_sync(c)) return;
    }
    if (mode == XBFS_MODE_TD_SF) ph_td<true>(a, L, bid, nb, enqueue != 0, sm);
    else if (mode == XBFS_MODE_TD_SS) ph_td<false>(a, L, bid, nb, false, sm);
    else ph_bu(a, L, bid, nb, lds);
    if (extra_sweep) {
        if (!xb_grid_sync(c)) return;
        ph_sweep(a, L + 1, bid, nb, true, true, lds);
    }
}

// ---- per-phase kernels for the host-driven loop ---------------------------
__global__ void __launch_bounds__(XB_PBLOCK) k_ph_init(XbArgs a) { ph_init(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(XB_PBLOCK) k_ph_seed(XbArgs a) { ph_seed(a, blockIdx.x, gridDim.x); }

// from_levels != 0: the sweep after a single-scan level (builds the bits too)
__global__ void __launch_bounds__(XB_PBLOCK) k_ph_sweep(XbArgs a, int32_t L, int32_t from_levels,
                                                        int32_t stats)
{
    extern __shared__ __align__(16) unsigned char xb_dyn_lds[];
    ph_sweep(a, L, blockIdx.x, gridDim.x, from_levels != 0, stats != 0,
             reinterpret_cast<XbWaveLds *>(xb_dyn_lds));
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
    extern __shared__ __align__(16) unsigned char xb_dyn_lds[];
    ph_bu(a, L, blockIdx.x, gridDim.x, reinterpret_cast<XbWaveLds *>(xb_dyn_lds));
}


// Host-driven loop: sums every field of the partial statistics into
// ctl->sum (one block).
__global__ void __launch_bounds__(1024) k_part_reduce(XbArgs a)
{
    for (int f = 0; f < XB_NFIELDS; f++) {
        long long acc = 0;
        for (int b = threadIdx.x; b < XB_PMAX; b += blockDim.x) acc += a.part[(int64_t)f * XB_PMAX + b];
        __shared__ long long sh[16];
        const long long t = (long long)xb_wave_sum_u64((unsigned long long)acc);
        if (xb_lane() == 0) sh[threadIdx.x >> 6] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long u = 0;
            for (int i = 0; i < (int)(blockDim.x >> 6); i++) u += sh[i];
            a.ctl->sum[f] = u;
        }
        __syncthreads();
    }
}

__global__ void k_ctl_reset(XbCtl *c, unsigned long long wait_ticks, unsigned long long res_ticks)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        xb_st(&c->wait_ticks, wait_ticks);
        xb_st(&c->res_ticks, res_ticks);
        xb_st(&c->bar_count, 0u);
        xb_st(&c->bar_gen, 0u);
        xb_st(&c->sbar_count, 0u);
        xb_st(&c->sbar_gen, 0u);
        xb_st(&c->heartbeat, 0u);
        xb_st(&c->error, 0);
    }
}

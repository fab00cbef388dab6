// XBFS level kernels for MI355X (gfx950, wave64).
//
// Every kernel is a grid-stride loop so the same code runs with any grid.
// Mode <-> reference mapping (bfs_adaptiveFQ_wb_async.cuh = wb,
// bfs_single_scan.cuh = ss, bfs_TD_scan_free.cuh = sf):
//   k_td<SF=true>   top-down scan-free: CAS claim, append the winner to the
//                   next queue's degree bin (wb:50-162, sf:23-135).
//   k_td<SF=false>  top-down single-scan expansion: plain stores, no enqueue
//                   (ss:52-126).
//   k_sweep_gen     frontier queue from a full sweep of level == L, binned by
//                   degree (ss:26-50), by wave ballots.
//   k_list_filter   frontier queue from the last bottom-up work list,
//                   filtered on level == L ("last phase", sf:23-135).
//   k_bu_count / k_scan_tiles / k_bu_scatter
//                   ordered unvisited list by 64-bit ballot + popcount and a
//                   device scan (ballot_counting + FQ_gen wb:189-284 and
//                   my_prefixsum prefix_sum1.cuh:59-108); the same sweep also
//                   writes the frontier bitmaps bottom-up probes.
//   k_bu            bottom-up update with the asynchronous +2 rule
//                   (thread_centric_BU_update, wb:738-777).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_device.h"
#include "xbfs_kernels.h"

// ---------------------------------------------------------------------------
// init / seed
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fill_i32(int32_t *__restrict__ a, int64_t n, int32_t val)
{
    const int64_t n4 = n >> 2;
    int4 v4 = make_int4(val, val, val, val);
    int4 *a4 = reinterpret_cast<int4 *>(a);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) a4[i] = v4;
    if (blockIdx.x == 0 && threadIdx.x < (n & 3)) a[n4 * 4 + threadIdx.x] = val;
}

// Seeds level 0: level[src]=0, parent[src]=src, src in its degree bin of
// queue buffer 0 (degree-0 sources are not enqueued, wb:963-980), ring slot 0
// = (1, deg(src)), everything else cleared.
__global__ void k_seed(int32_t *level, int32_t *parent, const int64_t *beg, XbDevCtl *c,
                       int32_t *q0_bins0, int32_t *q0_bins1, int32_t *q0_bins2, int32_t src,
                       int32_t t1, int32_t t2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < 3; k++) c->qcnt[b][k] = 0;
    c->list_cnt = 0;
    for (int s = 0; s < 4; s++) { c->slot_n[s] = 0; c->slot_m[s] = 0; }
    level[src] = 0;
    if (parent) parent[src] = src;
    const int64_t d = beg[src + 1] - beg[src];
    if (d > 0 && d <= t1) { q0_bins0[0] = src; c->qcnt[0][0] = 1; }
    else if (d > t1 && d <= t2) { q0_bins1[0] = src; c->qcnt[0][1] = 1; }
    else if (d > t2) { q0_bins2[0] = src; c->qcnt[0][2] = 1; }
    c->slot_n[0] = 1;
    c->slot_m[0] = (unsigned long long)d;
}

// Clears ring slots and queue counters before a level's kernels run.
__global__ void k_prep(XbDevCtl *c, int32_t zero_slot_a, int32_t zero_slot_b, int32_t zero_q_a,
                       int32_t zero_q_b, int32_t zero_list)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (zero_slot_a >= 0) { c->slot_n[zero_slot_a] = 0; c->slot_m[zero_slot_a] = 0; }
    if (zero_slot_b >= 0) { c->slot_n[zero_slot_b] = 0; c->slot_m[zero_slot_b] = 0; }
    if (zero_q_a >= 0) for (int k = 0; k < 3; k++) c->qcnt[zero_q_a][k] = 0;
    if (zero_q_b >= 0) for (int k = 0; k < 3; k++) c->qcnt[zero_q_b][k] = 0;
    if (zero_list) c->list_cnt = 0;
}

// nz bitmap: bit v set iff deg(v) > 0.  One wave per 64-vertex word.
__global__ void __launch_bounds__(256) k_build_nz(const int64_t *__restrict__ beg, int64_t n,
                                                  uint64_t *__restrict__ nz, int64_t nwords)
{
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nwords; w += nwaves) {
        int64_t v = w * 64 + xb_lane();
        bool p = v < n && beg[v + 1] > beg[v];
        unsigned long long m = __ballot(p);
        if (xb_lane() == 0) nz[w] = m;
    }
}

// ---------------------------------------------------------------------------
// top-down expansion
// ---------------------------------------------------------------------------
template <bool SF, bool PAR>
__device__ __forceinline__ void td_visit(const XbTdArgs &a, int32_t u, int32_t nb, bool valid,
                                         unsigned long long &my_n, unsigned long long &my_m)
{
    const int32_t L1 = a.L + 1;
    bool won = false;
    int32_t deg = 0;
    if (valid) {
        int32_t s = a.level[nb];
        if (s == -1) {
            if (SF) {
                won = atomicCAS(&a.level[nb], -1, L1) == -1;
                if (PAR && won) a.parent[nb] = u;
            } else {
                a.level[nb] = L1;
                if (PAR) a.parent[nb] = u;
            }
        }
    }
    if (SF) {
        if (won) {
            deg = (int32_t)(a.beg[nb + 1] - a.beg[nb]);
            my_n += 1;
            my_m += (unsigned long long)deg;
        }
        xb_wave_append(a.qn0, &a.cnt_n[0], won && deg > 0 && deg <= a.t1, nb);
        xb_wave_append(a.qn1, &a.cnt_n[1], won && deg > a.t1 && deg <= a.t2, nb);
        xb_wave_append(a.qn2, &a.cnt_n[2], won && deg > a.t2, nb);
    }
}

template <bool SF, bool PAR>
__global__ void __launch_bounds__(1024) k_td(XbTdArgs a)
{
    unsigned long long my_n = 0, my_m = 0;
    const int32_t nT = a.cnt_c[0], nW = a.cnt_c[1], nB = a.cnt_c[2];
    // block bin: one block per frontier vertex
    for (int32_t i = blockIdx.x; i < nB; i += gridDim.x) {
        const int32_t u = a.qc2[i];
        const int64_t b = a.beg[u], e = a.beg[u + 1];
        for (int64_t j = b + threadIdx.x; j < e; j += blockDim.x)
            td_visit<SF, PAR>(a, u, a.adj[j], true, my_n, my_m);
    }
    // wave bin: one wave per frontier vertex
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
    const int ln = xb_lane();
    for (int64_t i = gtid >> 6; i < nW; i += nthreads >> 6) {
        const int32_t u = a.qc1[i];
        const int64_t b = a.beg[u], e = a.beg[u + 1];
        for (int64_t j = b + ln; j < e; j += XB_WAVE)
            td_visit<SF, PAR>(a, u, a.adj[j], true, my_n, my_m);
    }
    // thread bin: one lane per frontier vertex
    for (int64_t i = gtid; i < nT; i += nthreads) {
        const int32_t u = a.qc0[i];
        const int64_t b = a.beg[u], e = a.beg[u + 1];
        for (int64_t j = b; j < e; j++)
            td_visit<SF, PAR>(a, u, a.adj[j], true, my_n, my_m);
    }
    if (SF) {
        unsigned long long v[2] = {my_n, my_m};
        unsigned long long *const d[2] = {a.slot_n1, a.slot_m1};
        xb_block_add_u64<2>(v, d);
    }
}

// Frontier of level L from a sweep over all vertices: level==L vertices with
// degree > 0 go to their degree bin; ring slot for L gets (count, edges).
__global__ void __launch_bounds__(1024) k_sweep_gen(XbGenArgs a)
{
    unsigned long long my_n = 0, my_m = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.n; base += stride) {
        const int64_t v = base + threadIdx.x;
        const bool hit = v < a.n && a.level[v] == a.L;
        int32_t deg = 0;
        if (hit) {
            deg = (int32_t)(a.beg[v + 1] - a.beg[v]);
            my_n += 1;
            my_m += (unsigned long long)deg;
        }
        xb_wave_append(a.q0, &a.cnt[0], hit && deg > 0 && deg <= a.t1, (int32_t)v);
        xb_wave_append(a.q1, &a.cnt[1], hit && deg > a.t1 && deg <= a.t2, (int32_t)v);
        xb_wave_append(a.q2, &a.cnt[2], hit && deg > a.t2, (int32_t)v);
    }
    unsigned long long vv[2] = {my_n, my_m};
    unsigned long long *const d[2] = {a.slot_n, a.slot_m};
    xb_block_add_u64<2>(vv, d);
}

// Frontier of level L from the last bottom-up work list (every vertex that was
// unvisited when that list was built): level == L -> current bins; with
// carry, level == L+1 (set early by the +2 rule) -> next bins, so the next
// queue is exact after this level's expansion appends its winners.
__global__ void __launch_bounds__(1024) k_list_filter(XbFilterArgs a)
{
    const int32_t cnt = *a.list_cnt;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < cnt; base += stride) {
        const int64_t i = base + threadIdx.x;
        int32_t v = 0, lv = -3, deg = 0;
        if (i < cnt) {
            v = a.list[i];
            lv = a.level[v];
            if (lv == a.L || (a.carry && lv == a.L + 1)) deg = (int32_t)(a.beg[v + 1] - a.beg[v]);
        }
        const bool cur = lv == a.L;
        const bool nxt = a.carry && lv == a.L + 1;
        xb_wave_append(a.qc0, &a.cnt_c[0], cur && deg <= a.t1, v);
        xb_wave_append(a.qc1, &a.cnt_c[1], cur && deg > a.t1 && deg <= a.t2, v);
        xb_wave_append(a.qc2, &a.cnt_c[2], cur && deg > a.t2, v);
        if (a.carry) {
            xb_wave_append(a.qn0, &a.cnt_n[0], nxt && deg <= a.t1, v);
            xb_wave_append(a.qn1, &a.cnt_n[1], nxt && deg > a.t1 && deg <= a.t2, v);
            xb_wave_append(a.qn2, &a.cnt_n[2], nxt && deg > a.t2, v);
        }
    }
}

// ---------------------------------------------------------------------------
// bottom-up: bitmaps + ordered unvisited list
// ---------------------------------------------------------------------------
// One block of 256 threads per tile of 64 words (4096 vertices); wave w
// handles words w*16 .. w*16+15 of the tile, one 64-vertex ballot per word.
//   fb bit v : level[v] == L        (frontier, probed by k_bu)
//   nb bit v : level[v] == L+1      (set early by +2; k_bu adds its finds)
//   ub bit v : level[v] == -1 and deg(v) > 0
// tile_cnt[t] = popcount of the tile's ub words.
__global__ void __launch_bounds__(256) k_bu_count(XbBuCountArgs a)
{
    const int w = threadIdx.x >> 6, ln = xb_lane();
    const int64_t word0 = (int64_t)blockIdx.x * XB_TILE_WORDS + w * 16;
    unsigned long long mf = 0, mn = 0, mu = 0;
    int cnt = 0;
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
        const int64_t wd = word0 + k;
        const int64_t v = wd * 64 + ln;
        const int32_t lv = (v < a.n) ? a.level[v] : -2;
        unsigned long long f = __ballot(lv == a.L);
        unsigned long long x = __ballot(lv == a.L + 1);
        unsigned long long u = __ballot(lv == -1);
        if (wd < a.nwords) u &= a.nz[wd];
        if (ln == k) { mf = f; mn = x; mu = u; }
        cnt += (int)__popcll(u);
    }
    const int64_t myw = word0 + ln;
    if (ln < 16 && myw < a.nwords) {
        if (a.fb) a.fb[myw] = mf;
        if (a.nb) a.nb[myw] = mn;
        a.ub[myw] = mu;
    }
    __shared__ int sh[4];
    if (ln == 0) sh[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) a.tile_cnt[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Exclusive scan of T int32 tile counts by one block; total -> *total.
__global__ void __launch_bounds__(1024) k_scan_tiles(const int32_t *__restrict__ cnt,
                                                     int32_t *__restrict__ off, int64_t T,
                                                     int32_t *total)
{
    const int64_t per = (T + blockDim.x - 1) / blockDim.x;
    const int64_t lo = (int64_t)threadIdx.x * per;
    const int64_t hi = lo + per < T ? lo + per : T;
    long long s = 0;
    for (int64_t i = lo; i < hi; i++) s += cnt[i];
    long long tot = 0;
    long long run = xb_block_excl_scan_i64(s, &tot);
    for (int64_t i = lo; i < hi; i++) {
        off[i] = (int32_t)run;
        run += cnt[i];
    }
    if (threadIdx.x == 0) *total = (int32_t)tot;
}

// Writes the ub bits of each tile, ascending, at tile_off[tile].
__global__ void __launch_bounds__(256) k_bu_scatter(const uint64_t *__restrict__ ub, int64_t nwords,
                                                    const int32_t *__restrict__ tile_off,
                                                    int32_t *__restrict__ list)
{
    const int w = threadIdx.x >> 6, ln = xb_lane();
    const int64_t word0 = (int64_t)blockIdx.x * XB_TILE_WORDS + w * 16;
    unsigned long long myu = 0;
    if (ln < 16 && word0 + ln < nwords) myu = ub[word0 + ln];
    const int pc = (int)__popcll(myu);
    const int inc = xb_wave_incl_scan_i32(pc);
    const int excl = inc - pc;
    __shared__ int sh[4];
    if (ln == XB_WAVE - 1) sh[w] = inc;
    __syncthreads();
    int wbase = tile_off[blockIdx.x];
    for (int i = 0; i < w; i++) wbase += sh[i];
    const unsigned long long below = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
    for (int k = 0; k < 16; k++) {
        const unsigned long long word = __shfl(myu, k);
        const int off = __shfl(excl, k);
        if ((word >> ln) & 1ull)
            list[wbase + off + (int)__popcll(word & below)] = (int32_t)((word0 + k) * 64 + ln);
    }
}

// Bottom-up update over the unvisited list (thread per vertex).  First
// neighbour in the level-L frontier -> level L+1 (early exit).  Otherwise a
// neighbour already at level L+1 -> level L+2: at the sweep start every level
// <= L is final, so "unvisited with no level-L neighbour" means distance
// >= L+2, and the L+1 neighbour bounds it by L+2 (wb:757-770).
template <bool PAR>
__global__ void __launch_bounds__(1024) k_bu(XbBuArgs a)
{
    unsigned long long n1 = 0, m1 = 0, n2 = 0, m2 = 0;
    const int32_t cnt = *a.list_cnt;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += stride) {
        const int32_t v = a.list[i];
        const int64_t b = a.beg[v], e = a.beg[v + 1];
        int32_t found = -1, seen = -1;
        for (int64_t j = b; j < e; j++) {
            const int32_t nb = a.adj[j];
            const uint64_t fw = a.fb[nb >> 6];
            if ((fw >> (nb & 63)) & 1ull) { found = nb; break; }
            if (seen < 0) {
                const uint64_t nw = a.nb[nb >> 6];
                if ((nw >> (nb & 63)) & 1ull) seen = nb;
            }
        }
        const unsigned long long d = (unsigned long long)(e - b);
        if (found >= 0) {
            a.level[v] = a.L + 1;
            if (PAR) a.parent[v] = found;
            atomicOr((unsigned long long *)&a.nb[v >> 6], 1ull << (v & 63));
            n1 += 1;
            m1 += d;
        } else if (seen >= 0) {
            a.level[v] = a.L + 2;
            if (PAR) a.parent[v] = seen;
            n2 += 1;
            m2 += d;
        }
    }
    unsigned long long vv[4] = {n1, m1, n2, m2};
    unsigned long long *const dd[4] = {a.slot_n1, a.slot_m1, a.slot_n2, a.slot_m2};
    xb_block_add_u64<4>(vv, dd);
}

// ---------------------------------------------------------------------------
// generic int32 -> int64 exclusive scan (three passes)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_scan_reduce(const int32_t *__restrict__ in, int64_t n,
                                                     long long *__restrict__ part)
{
    const int64_t lo = (int64_t)blockIdx.x * XB_SCAN_CHUNK;
    long long s = 0;
    for (int64_t i = lo + threadIdx.x; i < n && i < lo + XB_SCAN_CHUNK; i += blockDim.x) s += in[i];
    long long tot = 0;
    (void)xb_block_excl_scan_i64(s, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_parts(long long *__restrict__ part, int64_t T)
{
    const int64_t per = (T + blockDim.x - 1) / blockDim.x;
    const int64_t lo = (int64_t)threadIdx.x * per;
    const int64_t hi = lo + per < T ? lo + per : T;
    long long s = 0;
    for (int64_t i = lo; i < hi; i++) s += part[i];
    long long tot = 0;
    long long run = xb_block_excl_scan_i64(s, &tot);
    for (int64_t i = lo; i < hi; i++) {
        long long t = part[i];
        part[i] = run;
        run += t;
    }
}

// Each block scans its chunk: thread t owns XB_SCAN_CHUNK/256 consecutive
// elements (LDS-free: per-thread serial sums, block scan of the sums).
__global__ void __launch_bounds__(256) k_scan_down(const int32_t *__restrict__ in, int64_t n,
                                                   const long long *__restrict__ part,
                                                   long long *__restrict__ out)
{
    const int per = XB_SCAN_CHUNK / 256;
    const int64_t lo = (int64_t)blockIdx.x * XB_SCAN_CHUNK + (int64_t)threadIdx.x * per;
    long long s = 0;
    for (int k = 0; k < per; k++) {
        int64_t i = lo + k;
        if (i < n) s += in[i];
    }
    long long tot = 0;
    long long run = xb_block_excl_scan_i64(s, &tot) + part[blockIdx.x];
    for (int k = 0; k < per; k++) {
        int64_t i = lo + k;
        if (i < n) {
            out[i] = run;
            run += in[i];
        }
    }
}

// explicit instantiations used by the host engine
template __global__ void k_td<true, true>(XbTdArgs);
template __global__ void k_td<true, false>(XbTdArgs);
template __global__ void k_td<false, true>(XbTdArgs);
template __global__ void k_td<false, false>(XbTdArgs);
template __global__ void k_bu<true>(XbBuArgs);
template __global__ void k_bu<false>(XbBuArgs);

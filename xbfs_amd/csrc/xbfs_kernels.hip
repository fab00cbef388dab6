// Utility kernels of libxbfs (gfx950, wave64): the degree bitmap,
// the grid-wide ordered unvisited list and the generic device scan.
//   k_bu_count / k_scan_tiles / k_bu_scatter
//                   ascending list of unvisited vertices by 64-bit ballot +
//                   popcount and a device scan of tile counts: the
//                   ballot_counting + FQ_gen pair (bfs_adaptiveFQ_wb_async.cuh:
//                   189-284) with my_prefixsum (prefix_sum1.cuh:59-108) moved
//                   entirely onto the device.
//   k_scan_lookback
//                   int32 -> int64 exclusive scan in one pass with decoupled
//                   look-back (prefix_sum1.cuh:13-108, whose host-mediated
//                   block sums it replaces).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_device.h"
#include "xbfs_kernels.h"

// p[0..k) = 0, k <= 64 (small counters; a kernel, not a memset, so the same
// call is valid inside a captured graph)
__global__ void k_zero_u64(unsigned long long *p, int k)
{
    if ((int)threadIdx.x < k) p[threadIdx.x] = 0ull;
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

// bits[w] bit i <=> level[64 w + i] >= 0: the vertices a finished traversal
// reached (its source's component)
__global__ void __launch_bounds__(256) k_reached_bits(const int32_t *__restrict__ level, int64_t n,
                                                      uint64_t *__restrict__ bits, int64_t nwords)
{
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nwords; w += nwaves) {
        const int64_t v = w * 64 + xb_lane();
        const unsigned long long m = __ballot(v < n && level[v] >= 0);
        if (xb_lane() == 0) bits[w] = m;
    }
}

// Hub table: hd[v] = {the neighbour of v with the largest degree (first in
// adjacency order on ties), deg(v)}; {-1, 0} for an isolated vertex.
// Bottom-up probes the frontier bit of that neighbour before touching v's
// adjacency.  Lane per vertex for deg <= XB_HD_LANE_MAX; larger vertices
// are marked {-2, deg} and finished by k_build_hd_big.
#define XB_HD_LANE_MAX 256
__global__ void __launch_bounds__(256) k_build_hd(const int64_t *__restrict__ beg,
                                                  const int32_t *__restrict__ adj, int64_t n,
                                                  int2 *__restrict__ hd)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int64_t b = beg[v], e = beg[v + 1];
        const int32_t d = (int32_t)(e - b);
        int32_t best = -1;
        if (d > XB_HD_LANE_MAX) {
            best = -2;
        } else {
            int64_t bd = -1;
            for (int64_t j = b; j < e; j++) {
                const int32_t u = adj[j];
                const int64_t du = beg[u + 1] - beg[u];
                if (du > bd) {
                    bd = du;
                    best = u;
                }
            }
        }
        hd[v] = make_int2(best, d);
    }
}

// Sort keys of the degree-ordered adjacency: key[j] = deg(adj[j]).
__global__ void __launch_bounds__(256) k_nbr_deg(const int64_t *__restrict__ beg,
                                                 const int32_t *__restrict__ adj, int64_t m,
                                                 uint32_t *__restrict__ key)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
        const int32_t u = adj[j];
        key[j] = (uint32_t)(beg[u + 1] - beg[u]);
    }
}

// Hub table of a degree-ordered adjacency (stable sort, descending): the
// first entry of a row is its first neighbour of largest degree.
__global__ void __launch_bounds__(256) k_build_hd_sorted(const int64_t *__restrict__ beg,
                                                         const int32_t *__restrict__ adj, int64_t n,
                                                         int2 *__restrict__ hd)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int64_t b = beg[v], e = beg[v + 1];
        hd[v] = make_int2(e > b ? adj[b] : -1, (int32_t)(e - b));
    }
}

// Marked vertices, one wave each: a wave checks 64 vertices per load, then
// for each marked one its lanes stride the adjacency and a wave reduction on
// (degree, -position) keeps the first neighbour of largest degree.
__global__ void __launch_bounds__(256) k_build_hd_big(const int64_t *__restrict__ beg,
                                                      const int32_t *__restrict__ adj, int64_t n,
                                                      int2 *__restrict__ hd)
{
    const int64_t nwv = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int ln = xb_lane();
    for (int64_t v0 = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64; v0 < n;
         v0 += nwv * 64) {
        const bool mk = v0 + ln < n && hd[v0 + ln].x == -2;
        unsigned long long m = __ballot(mk);
        while (m) {
            const int k = __ffsll(m) - 1;
            m &= m - 1;
            const int64_t v = v0 + k;
            const int64_t b = beg[v], e = beg[v + 1];
            long long bd = -1, bp = 0x7fffffffffffffffll;
            for (int64_t j = b + ln; j < e; j += 64) {
                const int32_t u = adj[j];
                const long long du = beg[u + 1] - beg[u];
                if (du > bd) { // first (smallest j) of this lane's maximum
                    bd = du;
                    bp = j;
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const long long od = __shfl_xor(bd, o), op = __shfl_xor(bp, o);
                if (od > bd || (od == bd && op < bp)) {
                    bd = od;
                    bp = op;
                }
            }
            if (ln == 0) hd[v] = make_int2(adj[bp], (int32_t)(e - b));
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

// Single-pass exclusive scan of int32 into int64 with decoupled look-back
// (Merrill and Garland, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back", NVIDIA technical report NVR-2016-002): one launch, every
// element read once and written once.
//  - Tiles of XB_SCAN_CHUNK elements are handed out by an atomic ticket in the
//    order blocks start, so every tile's predecessors have already started
//    and none of them waits on a later one: no dependence on which blocks
//    the hardware runs first or at once.
//  - Each tile sums its elements (thread t owns XB_SCAN_CHUNK/256 consecutive
//    ones, block scan of the per-thread sums), publishes its aggregate
//    (flag A), walks back over the preceding tiles' status words adding
//    aggregates until it meets an inclusive prefix (flag P), and publishes
//    its own inclusive prefix.
//  - A status word packs the flag (bits 62-63) and the signed value (62-bit
//    two's complement) in one 64-bit word, so one atomic load reads both.
//  - The wait for a predecessor's word is bounded (2^26 polls); past it the
//    tile sets *err and stops waiting.
// status[T] and *ticket must be zero before the launch.
#define XB_SCAN_FLAG_A 1ull
#define XB_SCAN_FLAG_P 2ull
__device__ __forceinline__ unsigned long long xb_scan_pack(unsigned long long flag, long long v)
{
    return (flag << 62) | ((unsigned long long)v & ((1ull << 62) - 1));
}
__device__ __forceinline__ long long xb_scan_value(unsigned long long w)
{
    return (long long)(w << 2) >> 2;
}

__global__ void __launch_bounds__(256) k_scan_lookback(const int32_t *__restrict__ in, int64_t n,
                                                       unsigned long long *status, int32_t *ticket,
                                                       int32_t *err, long long *__restrict__ out)
{
    __shared__ int32_t s_tile;
    __shared__ long long s_excl;
    if (threadIdx.x == 0)
        s_tile = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int64_t tile = s_tile;
    const int per = XB_SCAN_CHUNK / 256;
    const int64_t lo = tile * XB_SCAN_CHUNK + (int64_t)threadIdx.x * per;
    int32_t x[XB_SCAN_CHUNK / 256];
    long long s = 0;
#pragma unroll
    for (int k = 0; k < per; k++) {
        const int64_t i = lo + k;
        x[k] = i < n ? in[i] : 0;
        s += x[k];
    }
    long long tot = 0;
    const long long ex = xb_block_excl_scan_i64(s, &tot);
    if (threadIdx.x == 0) {
        long long run = 0;
        if (tile == 0) {
            __hip_atomic_store(&status[0], xb_scan_pack(XB_SCAN_FLAG_P, tot), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&status[tile], xb_scan_pack(XB_SCAN_FLAG_A, tot), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            for (int64_t j = tile - 1; j >= 0; j--) {
                unsigned long long w = 0;
                for (int poll = 0; poll < (1 << 26); poll++) {
                    w = __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (w >> 62) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                if ((w >> 62) == 0) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                run += xb_scan_value(w);
                if ((w >> 62) == XB_SCAN_FLAG_P) break;
            }
            __hip_atomic_store(&status[tile], xb_scan_pack(XB_SCAN_FLAG_P, run + tot),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_excl = run;
    }
    __syncthreads();
    long long r = s_excl + ex;
#pragma unroll
    for (int k = 0; k < per; k++) {
        const int64_t i = lo + k;
        if (i < n) out[i] = r;
        r += x[k];
    }
}


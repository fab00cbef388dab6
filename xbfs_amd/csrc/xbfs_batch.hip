// 64-source batched BFS (no reference counterpart; the reference's README.md:80
// cites iBFS-style concurrent traversals).  Up to 64 single-source BFSs run as
// one traversal: source s is bit s of a 64-bit mask per vertex, and the whole
// batch -- clearing, seeding, every level, the per-source level rows -- is one
// cooperative launch of k_b64_persistent.
//
// Vertex order.  The batch works on a relabelled copy of the graph: rank r is
// the r-th vertex by degree, descending (ties by id), and every row lists its
// neighbours by rank, ascending (largest degree first).  On skewed graphs most
// edge endpoints are a few hubs, so the frontier masks a bottom-up level
// gathers concentrate in the first few megabytes of the mask array, which
// stay in each XCD's L2, and a row's scan meets the likeliest frontier
// members first.  Vertices of similar degree are neighbours in rank order, so
// a wave's lanes get rows of similar length, and the isolated vertices form
// one tail that no level touches.
//
// State (rank order, n vertices, H history slots):
//   F      H x uint64[n]  frontier masks: bit s of slot L at r <=> level_s(r) = L
//   vis    uint64[n]      visited masks (bit s <=> some level <= the current one)
//   acc    uint64[n]      top-down accumulator (atomic OR), zero between levels
//   rec    8 x uint64[n]  per original vertex: reached mask and 6 bit planes of
//                         its level per source (written at the end)
//   lv     int32[64 x n]  row s = levels from source s, original vertex order
//
// Level L (expands F_L into F_{L+1}), direction decided on the device from
// the level's frontier edges against the edges of vertices still missing a
// source (the alpha rule of the single-source heuristic, or a forced mode).
// The unit of edge work is a 32-edge segment of a row (rows are contiguous in
// rank order, so segment i ends where segment i+1 starts): no lane walks more
// than 32 edges whatever the degree skew.  Chunks of XB_B64_CHUNK segments
// are dealt to the blocks round robin, and a wave keeps, by 64-bit ballots,
// only the segments of its chunk with work:
//   pull (bottom-up)  segments of a vertex r with unvisited bits U = all &
//                     ~vis[r]: OR of F_L over the segment's neighbours,
//                     stopping once U is covered; acc[r] |= found & U
//   push (top-down)   segments of a frontier vertex r: acc[q] |= F_L[r] for
//                     each neighbour q
// then, after a grid barrier, per vertex F_{L+1} = acc & ~vis, vis |=
// F_{L+1}, acc = 0, with the level's statistics.
// Sources still active when the H slots are used up (very long diameters)
// are finished by the single-source engine.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_device.h"

#define XB_B64_BLOCK 1024  /* threads per block of the batch kernel            */
#define XB_B64_HMAX 64     /* history slots at most (levels 0..63 in one window) */
#define XB_B64_PLANES 6    /* bit planes of a level < 64                         */
#define XB_B64_OUT (8 + 2 * (XB_B64_HMAX + 2)) /* entries of the result array  */
#define XB_B64_SEG 32      /* edges per segment                                  */
#define XB_B64_ACC_HOT 32768 /* ranks with an accumulator line (128 B) each     */
#ifndef XB_B64_CHUNK
#define XB_B64_CHUNK 256   /* segments per chunk (4 per lane)                    */
#endif
#define XB_B64_PULL_K 16   /* neighbours gathered per round                      */

// Statistic fields of the per-block partials (column = block id): level
// parity p -> 4p + {0: vertices with a new bit, 1: their degree sum, 2: degree
// sum of vertices still missing a source, 3: OR of the new bits}; the final
// pass: 8 + {0: reached (source, vertex) pairs, 1: their degree sum}.
#define XB_B64_F_LVL(p, k) (4 * ((p) & 1) + (k))
#define XB_B64_F_VC 8
#define XB_B64_F_MC 9

struct XbB64Args {
    const int64_t *rbeg;  // int64[n+1], rank order
    const int32_t *radj;  // int32[m], neighbours by rank, ascending per row
    const int32_t *rdeg;  // int32[n], degree of rank r (descending)
    const int32_t *seg_r; // int32[nseg], the rank whose row holds segment i
    const int64_t *seg_e; // int64[nseg+1], first edge of segment i (seg_e[nseg] = m)
    const int32_t *rank;  // original id -> rank
    const int32_t *ord;   // rank -> original id
    uint64_t *F;          // H slots of n
    uint64_t *vis, *acc;
    uint64_t *rec;        // 8 words per original vertex
    int32_t *lv;          // nsrc rows of n, original order
    long long *part;      // per-block partial statistics, XB_PMAX columns per field
    XbCtl *ctl;
    long long *out;       // [0] levels expanded, [1] last level's OR of new bits,
                          // [2] frontier count of the last level, [3] v_cc, [4] m_cc,
                          // [8 + L] wall clock at the start of level L (L <= levels
                          // expanded), [8 + XB_B64_HMAX] between the two passes of
                          // the level rows, [8 + XB_B64_HMAX + 1] after them,
                          // [8 + XB_B64_HMAX + 2 + L] direction of level L (1 pull)
    const int32_t *srcs;
    const int8_t *sched;
    int64_t n, m, n_nz, nseg;
    uint64_t all;
    int32_t nsrc, H, policy, sched_len;
    float alpha;
};

// ---------------------------------------------------------------------------
// set-up of the relabelled graph (once per context)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_b64_deg_key(const int64_t *__restrict__ beg, int64_t n,
                                                     uint32_t *__restrict__ key,
                                                     int32_t *__restrict__ id)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        key[v] = (uint32_t)(beg[v + 1] - beg[v]);
        id[v] = (int32_t)v;
    }
}

// rank[ord[r]] = r, rdeg[r] = degree of ord[r]
__global__ void __launch_bounds__(256) k_b64_rank(const int32_t *__restrict__ ord,
                                                  const uint32_t *__restrict__ key_sorted, int64_t n,
                                                  int32_t *__restrict__ rank, int32_t *__restrict__ rdeg)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        rank[ord[r]] = (int32_t)r;
        rdeg[r] = (int32_t)key_sorted[r];
    }
}

__global__ void k_b64_set_end(int64_t *rbeg, int64_t n, int64_t m)
{
    if (threadIdx.x == 0) rbeg[n] = m;
}

// segments per rank: ceil(degree / XB_B64_SEG)
__global__ void __launch_bounds__(256) k_b64_seg_count(const int32_t *__restrict__ rdeg, int64_t n,
                                                       int32_t *__restrict__ cnt)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
        cnt[r] = (rdeg[r] + XB_B64_SEG - 1) / XB_B64_SEG;
}

// off = exclusive scan of the segment counts
__global__ void __launch_bounds__(256) k_b64_seg_fill(const int64_t *__restrict__ rbeg, int64_t n,
                                                      const int32_t *__restrict__ cnt,
                                                      const long long *__restrict__ off,
                                                      int32_t *__restrict__ seg_r,
                                                      int64_t *__restrict__ seg_e, int64_t nseg,
                                                      int64_t m)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t r = gtid; r < n; r += stride) {
        const int32_t c = cnt[r];
        const int64_t o = off[r], b = rbeg[r];
        for (int32_t k = 0; k < c; k++) {
            seg_r[o + k] = (int32_t)r;
            seg_e[o + k] = b + (int64_t)k * XB_B64_SEG;
        }
    }
    if (gtid == 0) seg_e[nseg] = m;
}

// Row r of the relabelled graph: row ord[r] of the input with every id
// replaced by its rank (one wave per row; the per-row sort follows).
__global__ void __launch_bounds__(256) k_b64_rows(const int64_t *__restrict__ beg,
                                                  const int32_t *__restrict__ adj,
                                                  const int32_t *__restrict__ ord,
                                                  const int32_t *__restrict__ rank,
                                                  const int64_t *__restrict__ rbeg, int64_t n,
                                                  int32_t *__restrict__ radj)
{
    const int64_t nwv = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int ln = xb_lane();
    for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < n; r += nwv) {
        const int32_t v = ord[r];
        const int64_t b = beg[v], d = beg[v + 1] - b, o = rbeg[r];
        for (int64_t j = ln; j < d; j += XB_WAVE) radj[o + j] = rank[adj[b + j]];
    }
}

// ---------------------------------------------------------------------------
// the batch kernel
// ---------------------------------------------------------------------------
// Accumulator word of rank r.  The top-down atomics of a level concentrate on
// the hubs, which rank order packs into a few cache lines; device-scope
// atomics on one line serialise, so each of the XB_B64_ACC_HOT first ranks
// gets a 128-byte line of its own.  The array holds n + 15 * XB_B64_ACC_HOT
// words.
__device__ __forceinline__ int64_t xb_b64_acc(int64_t r)
{
    return r < XB_B64_ACC_HOT ? r * 16 : r + 15 * (int64_t)XB_B64_ACC_HOT;
}

__device__ __forceinline__ unsigned long long xb_wave_or_u64(unsigned long long x)
{
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
    return x;
}

// Block totals of 4 level statistics (sums, and an OR for k = 3) into column
// bid of their fields (thread 0, plain stores: only this block writes the
// column).  Every thread of the block calls it.
__device__ void xb_b64_store4(const XbB64Args &a, int bid, int p, unsigned long long (&v)[4])
{
    __shared__ unsigned long long sh[4][XB_B64_BLOCK / 64];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const unsigned long long t = xb_wave_sum_u64(v[k]);
        if (xb_lane() == 0) sh[k][w] = t;
    }
    {
        const unsigned long long t = xb_wave_or_u64(v[3]);
        if (xb_lane() == 0) sh[3][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        unsigned long long t[4] = {0, 0, 0, 0};
        for (int i = 0; i < nw; i++) {
            t[0] += sh[0][i];
            t[1] += sh[1][i];
            t[2] += sh[2][i];
            t[3] |= sh[3][i];
        }
        for (int k = 0; k < 4; k++) a.part[(int64_t)XB_B64_F_LVL(p, k) * XB_PMAX + bid] = (long long)t[k];
    }
    __syncthreads();
}

// Level statistics of parity p summed over nb columns by wave 0 (all loads
// in flight at once); the results land in sh[0..3] for the whole block.
__device__ void xb_b64_sum4(const XbB64Args &a, int nb, int p, unsigned long long *sh)
{
    if (threadIdx.x < 64) {
        const int ln = xb_lane();
        unsigned long long s0 = 0, s1 = 0, s2 = 0, o3 = 0;
        for (int b0 = 0; b0 < nb; b0 += 256) {
            unsigned long long v[4][4];
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int b = b0 + ln + 64 * j;
                    v[k][j] = b < nb ? (unsigned long long)a.part[(int64_t)XB_B64_F_LVL(p, k) * XB_PMAX + b]
                                     : 0ull;
                }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                s0 += v[0][j];
                s1 += v[1][j];
                s2 += v[2][j];
                o3 |= v[3][j];
            }
        }
        s0 = xb_wave_sum_u64(s0);
        s1 = xb_wave_sum_u64(s1);
        s2 = xb_wave_sum_u64(s2);
        o3 = xb_wave_or_u64(o3);
        if (ln == 0) {
            sh[0] = s0;
            sh[1] = s1;
            sh[2] = s2;
            sh[3] = o3;
        }
    }
    __syncthreads();
}

// Direction of level L: 1 pull, 0 push.
__device__ __forceinline__ int xb_b64_mode(const XbB64Args &a, int32_t L, unsigned long long m_f,
                                           unsigned long long m_u)
{
    switch (a.policy) {
    case XBFS_POLICY_FORCE_TD_SF:
    case XBFS_POLICY_FORCE_TD_SS: return 0;
    case XBFS_POLICY_FORCE_BU: return 1;
    case XBFS_POLICY_SCHEDULE:
        if (L < a.sched_len) return a.sched[L] == XBFS_MODE_BU ? 1 : 0;
        break;
    default: break;
    }
    return (double)m_f > (double)a.alpha * (double)m_u ? 1 : 0;
}

// Builds the wave's list of the chunk's segments with work: bit set in
// `want` per lane and round k (segment base + 64 k + lane); ascending order.
__device__ __forceinline__ int xb_b64_list(int32_t *list, int64_t base, unsigned want)
{
    const int ln = xb_lane();
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < XB_B64_CHUNK / 64; k++) {
        const bool w = (want >> k) & 1u;
        const unsigned long long mk = __ballot(w);
        if (w) list[cnt + xb_rank(mk)] = (int32_t)(base + 64 * k + ln);
        cnt += (int)__popcll(mk);
    }
    xb_wave_sync();
    return cnt;
}

__global__ void __launch_bounds__(XB_B64_BLOCK) k_b64_persistent(XbB64Args a)
{
    __shared__ int32_t s_list[XB_B64_BLOCK / 64][XB_B64_CHUNK];
    __shared__ unsigned long long s_st[4];
    const int bid = blockIdx.x, nb = gridDim.x, ln = xb_lane();
    const int64_t gtid = (int64_t)bid * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)nb * blockDim.x;
    const int64_t nwv = nthr >> 6, wrr = (int64_t)(threadIdx.x >> 6) * nb + bid;
    int32_t *list = s_list[threadIdx.x >> 6];
    const int64_t n = a.n, n_nz = a.n_nz;
    const int64_t nchunks = (a.nseg + XB_B64_CHUNK - 1) / XB_B64_CHUNK;

    // clear: visited, accumulator, slot 0
    for (int64_t v = gtid; v < n; v += nthr) {
        a.vis[v] = 0ull;
        a.acc[xb_b64_acc(v)] = 0ull;
        a.F[v] = 0ull;
    }
    if (!xb_grid_sync(a.ctl)) return;
    // seed: bit i at srcs[i]; level-0 statistics (distinct seeds with edges)
    // in block 0's column, zeros in the others
    {
        unsigned long long v4[4] = {0, 0, 0, 0};
        if (bid == 0 && threadIdx.x < 64) {
            const int i = threadIdx.x;
            if (i < a.nsrc) {
                const int32_t r = a.rank[a.srcs[i]];
                const unsigned long long bit = 1ull << i;
                atomicOr(reinterpret_cast<unsigned long long *>(&a.F[r]), bit);
                atomicOr(reinterpret_cast<unsigned long long *>(&a.vis[r]), bit);
                bool first = true;
                for (int j = 0; j < i; j++)
                    if (a.srcs[j] == a.srcs[i]) first = false;
                if (first && a.rdeg[r] > 0) {
                    v4[0] = 1;
                    v4[1] = (unsigned long long)a.rdeg[r];
                }
                v4[3] = bit;
            }
            if (i == 0) v4[2] = (unsigned long long)a.m;
        }
        xb_b64_store4(a, bid, 0, v4);
    }
    if (!xb_grid_sync(a.ctl)) return;

    int32_t L = 0;
    unsigned long long last_or = 0, last_nf = 0;
    for (;;) {
        if (bid == 0 && threadIdx.x == 0) a.out[8 + L] = (long long)wall_clock64();
        xb_b64_sum4(a, nb, L, s_st);
        const unsigned long long n_f = s_st[0], m_f = s_st[1], m_u = s_st[2], orr = s_st[3];
        last_or = orr;
        last_nf = n_f;
        if (n_f == 0 || L + 1 >= a.H) break;
        const int pull = xb_b64_mode(a, L, m_f, m_u);
        if (bid == 0 && threadIdx.x == 0) a.out[8 + XB_B64_HMAX + 2 + L] = pull;
        const uint64_t *Fc = a.F + (int64_t)L * n;
        uint64_t *Fn = a.F + (int64_t)(L + 1) * n;
        // segments with work, chunk by chunk; chunks are dealt to the blocks
        // round robin (chunk c to block c mod nb), so the few chunks that hold
        // a level's work (the hub rows of a top-down level) spread over every
        // CU instead of filling the first few
        for (int64_t ch = wrr; ch < nchunks; ch += nwv) {
            const int64_t base = ch * XB_B64_CHUNK;
            int32_t sr[XB_B64_CHUNK / 64];
#pragma unroll
            for (int k = 0; k < XB_B64_CHUNK / 64; k++) {
                const int64_t i = base + 64 * k + ln;
                sr[k] = i < a.nseg ? a.seg_r[i] : -1;
            }
            unsigned want = 0;
            if (pull) {
                uint64_t vp[XB_B64_CHUNK / 64];
#pragma unroll
                for (int k = 0; k < XB_B64_CHUNK / 64; k++) vp[k] = sr[k] >= 0 ? a.vis[sr[k]] : a.all;
#pragma unroll
                for (int k = 0; k < XB_B64_CHUNK / 64; k++)
                    if ((a.all & ~vp[k]) != 0ull) want |= 1u << k;
            } else {
                uint64_t f[XB_B64_CHUNK / 64];
#pragma unroll
                for (int k = 0; k < XB_B64_CHUNK / 64; k++) f[k] = sr[k] >= 0 ? Fc[sr[k]] : 0ull;
#pragma unroll
                for (int k = 0; k < XB_B64_CHUNK / 64; k++)
                    if (f[k] != 0ull) want |= 1u << k;
            }
            const int cnt = xb_b64_list(list, base, want);
            for (int i = ln; i < cnt; i += 64) {
                const int32_t sg = list[i];
                const int32_t r = a.seg_r[sg];
                const int64_t e0 = a.seg_e[sg], e1 = a.seg_e[sg + 1];
                if (pull) {
                    // bits another segment of the row already found are left out
                    const uint64_t U = a.all & ~a.vis[r] & ~a.acc[xb_b64_acc(r)];
                    if (!U) continue;
                    uint64_t got = 0;
                    for (int64_t e = e0; e < e1; e += XB_B64_PULL_K) {
                        int32_t u[XB_B64_PULL_K];
#pragma unroll
                        for (int k = 0; k < XB_B64_PULL_K; k++) u[k] = e + k < e1 ? a.radj[e + k] : -1;
                        uint64_t f[XB_B64_PULL_K];
#pragma unroll
                        for (int k = 0; k < XB_B64_PULL_K; k++) f[k] = u[k] >= 0 ? Fc[u[k]] : 0ull;
#pragma unroll
                        for (int k = 0; k < XB_B64_PULL_K; k++) got |= f[k];
                        if ((got & U) == U) break;
                    }
                    const uint64_t nw = got & U;
                    if (nw)
                        atomicOr(reinterpret_cast<unsigned long long *>(&a.acc[xb_b64_acc(r)]),
                                 (unsigned long long)nw);
                } else {
                    const uint64_t f = Fc[r];
                    for (int64_t e = e0; e < e1; e += 8) {
                        int32_t u[8];
#pragma unroll
                        for (int k = 0; k < 8; k++) u[k] = e + k < e1 ? a.radj[e + k] : -1;
                        // an atomic only where the accumulator misses some of
                        // the bits (hub targets receive the same bits from many
                        // frontier rows)
                        uint64_t have[8];
#pragma unroll
                        for (int k = 0; k < 8; k++) have[k] = u[k] >= 0 ? a.acc[xb_b64_acc(u[k])] : f;
#pragma unroll
                        for (int k = 0; k < 8; k++)
                            if ((have[k] & f) != f)
                                atomicOr(reinterpret_cast<unsigned long long *>(&a.acc[xb_b64_acc(u[k])]),
                                         (unsigned long long)f);
                    }
                }
            }
            xb_wave_sync();
        }
        if (!xb_grid_sync(a.ctl)) return;
        // the accumulated bits become level L+1 where they are new
        unsigned long long st[4] = {0, 0, 0, 0}; // new vertices, their degrees, still-missing degrees, OR
        for (int64_t r = gtid; r < n_nz; r += nthr) {
            const int64_t ai = xb_b64_acc(r);
            const uint64_t ac = a.acc[ai];
            const uint64_t vp = a.vis[r];
            const uint64_t nw = ac & ~vp;
            Fn[r] = nw;
            if (ac) a.acc[ai] = 0ull;
            const unsigned long long d = (unsigned long long)a.rdeg[r];
            if (nw) {
                a.vis[r] = vp | nw;
                st[0] += 1;
                st[1] += d;
                st[3] |= nw;
            }
            if (a.all & ~(vp | nw)) st[2] += d;
        }
        xb_b64_store4(a, bid, L + 1, st);
        if (!xb_grid_sync(a.ctl)) return;
        L++;
    }

    // Level rows.  Pass 1 (rank order, coalesced slot reads): each vertex's
    // reached mask and the 6 bit planes of its levels (bit s of plane k = bit
    // k of level_s), one 64-byte record at its original id; a source still
    // active at the end of the window (finished by the single-source engine)
    // is left out.  Pass 2 (original order): every row s gets the vertex's
    // level in one coalesced 256-byte run per wave.
    const uint64_t unfinished = last_nf ? (last_or & a.all) : 0ull;
    const uint64_t done = a.all & ~unfinished;
    const int32_t nslots = L + 1;
    unsigned long long vc = 0, mc = 0;
    for (int64_t r = gtid; r < n; r += nthr) {
        const int32_t ns = r < n_nz ? nslots : 1; // isolated: only a seed has bits
        uint64_t reached = 0, P[XB_B64_PLANES];
#pragma unroll
        for (int k = 0; k < XB_B64_PLANES; k++) P[k] = 0ull;
        for (int32_t j = 0; j < ns; j++) {
            const uint64_t f = a.F[(int64_t)j * n + r];
            reached |= f;
#pragma unroll
            for (int k = 0; k < XB_B64_PLANES; k++)
                if ((j >> k) & 1) P[k] |= f;
        }
        reached &= done;
        vc += (unsigned long long)__popcll(reached);
        mc += (unsigned long long)__popcll(reached) * (unsigned long long)a.rdeg[r];
        ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(a.rec + (int64_t)a.ord[r] * 8);
        dst[0] = make_ulonglong2(reached, P[0]);
        dst[1] = make_ulonglong2(P[1], P[2]);
        dst[2] = make_ulonglong2(P[3], P[4]);
        dst[3] = make_ulonglong2(P[5], 0ull);
    }
    if (!xb_grid_sync(a.ctl)) return;
    if (bid == 0 && threadIdx.x == 0) a.out[8 + XB_B64_HMAX] = (long long)wall_clock64();
    for (int64_t v = gtid; v < n; v += nthr) {
        const ulonglong2 *srcp = reinterpret_cast<const ulonglong2 *>(a.rec + v * 8);
        const ulonglong2 q0 = srcp[0], q1 = srcp[1], q2 = srcp[2], q3 = srcp[3];
        const uint64_t reached = q0.x;
        const uint64_t P[XB_B64_PLANES] = {q0.y, q1.x, q1.y, q2.x, q2.y, q3.x};
        for (int s = 0; s < a.nsrc; s++) {
            if (!((done >> s) & 1ull)) continue;
            int32_t lvl = 0;
#pragma unroll
            for (int k = 0; k < XB_B64_PLANES; k++) lvl |= (int32_t)((P[k] >> s) & 1ull) << k;
            __builtin_nontemporal_store(((reached >> s) & 1ull) ? lvl : -1, &a.lv[(int64_t)s * n + v]);
        }
    }
    {
        __shared__ unsigned long long sh[2][XB_B64_BLOCK / 64];
        const int w = threadIdx.x >> 6;
        const unsigned long long t0 = xb_wave_sum_u64(vc), t1 = xb_wave_sum_u64(mc);
        if (ln == 0) {
            sh[0][w] = t0;
            sh[1][w] = t1;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long x0 = 0, x1 = 0;
            for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
                x0 += sh[0][i];
                x1 += sh[1][i];
            }
            a.part[(int64_t)XB_B64_F_VC * XB_PMAX + bid] = (long long)x0;
            a.part[(int64_t)XB_B64_F_MC * XB_PMAX + bid] = (long long)x1;
        }
    }
    if (!xb_grid_sync(a.ctl)) return;
    if (bid == 0 && threadIdx.x < 64) {
        unsigned long long x0 = 0, x1 = 0;
        for (int b = ln; b < nb; b += 64) {
            x0 += (unsigned long long)a.part[(int64_t)XB_B64_F_VC * XB_PMAX + b];
            x1 += (unsigned long long)a.part[(int64_t)XB_B64_F_MC * XB_PMAX + b];
        }
        x0 = xb_wave_sum_u64(x0);
        x1 = xb_wave_sum_u64(x1);
        if (ln == 0) {
            a.out[8 + XB_B64_HMAX + 1] = (long long)wall_clock64();
            a.out[0] = L;
            a.out[1] = (long long)unfinished;
            a.out[2] = (long long)last_nf;
            a.out[3] = (long long)x0;
            a.out[4] = (long long)x1;
        }
    }
}

// 64-source batched BFS (no reference counterpart; the reference's README.md:80
// cites iBFS-style concurrent traversals).  Up to 64 single-source BFSs run as
// one traversal: source s is bit s of a 64-bit mask per vertex.
//
// State (n vertices, m directed edges, H history slots):
//   F[2]   uint64[n]   frontier masks, ping-pong by level parity:
//                      bit s of F_L[v] <=> level_s(v) == L
//   acc    uint64[n]   next-level accumulator (atomic OR only)
//   hist   uint64[H*n] visited history, slot L = vis_L: bit s of
//                      hist[L][v] <=> level_s(v) <= L (monotone in L)
//   lv     int32[64*n] per-source levels, row s = levels from source s
//   seg    32-edge segments of the adjacency (seg_u = owning vertex, seg_e =
//          first edge; segments of consecutive vertices are contiguous, so
//          segment i ends at seg_e[i+1]): the unit of edge work, which bounds
//          any lane's loop at 32 edges whatever the degree skew.
//
// Level L (expands F_L into level L+1):
//   push  (top-down)  lane per segment of a frontier vertex u: for each
//                     neighbour v, acc[v] |= F_L[u] & ~vis_L[v]
//   pull  (bottom-up) lane per segment of a vertex v with unvisited bits U:
//                     OR of F_L over its neighbours with early exit once U
//                     is covered; acc[v] |= found & U
//   fin               per vertex: new = acc & ~vis_L, vis_{L+1} = vis_L | new,
//                     F_{L+1} = new, acc = 0, plus per-block statistics
// Levels never leave the masks during the traversal; one final pass reads
// each vertex's history and writes row s of lv in full cache lines: the
// level of (s, v) is the first slot whose bit s is set (-1 if none).
// Sources still active when the history window is full (very long
// diameters) are finished by the single-source engine.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_device.h"

#define XB_B64_SEG 32      /* edges per segment                       */
#define XB_B64_BLOCK 256   /* block size of every batch kernel         */
#define XB_B64_HMAX 64     /* history slots at most (levels 0..63 in one window) */
#define XB_B64_PLANES 6    /* bit planes of a slot index < 64                  */

// Statistic fields of the per-block partials (column = block id).
#define XB_B64_F_NF 0    /* vertices with F_{L+1} != 0           */
#define XB_B64_F_MF 1    /* their degree sum                     */
#define XB_B64_F_MU 2    /* degree sum of vertices with unvisited bits */
#define XB_B64_F_OR 3    /* OR of F_{L+1} (active sources)        */
#define XB_B64_F_VC 4    /* final pass: reached (source, vertex) pairs */
#define XB_B64_F_MC 5    /* final pass: their degree sum          */
#define XB_B64_NF 6

struct XbB64Args {
    const int32_t *deg;
    const int32_t *adj;
    const int32_t *seg_u;
    const int64_t *seg_e;
    int64_t nseg;
    const uint64_t *F_cur;
    uint64_t *F_next;
    uint64_t *acc;
    const uint64_t *vis_cur;
    uint64_t *vis_next;
    long long *part; // XB_B64_NF x pstride
    int64_t pstride;
    int64_t n;
    uint64_t all; // mask of the batch's sources
};

// ---------------------------------------------------------------------------
// set-up: degrees and segments (once per context)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(XB_B64_BLOCK) k_b64_deg(const int64_t *__restrict__ beg, int64_t n,
                                                          int32_t *__restrict__ deg,
                                                          int32_t *__restrict__ nseg)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int64_t d = beg[v + 1] - beg[v];
        deg[v] = (int32_t)d;
        nseg[v] = (int32_t)((d + XB_B64_SEG - 1) / XB_B64_SEG);
    }
}

// off = exclusive scan of the per-vertex segment counts
__global__ void __launch_bounds__(XB_B64_BLOCK) k_b64_seg_fill(const int64_t *__restrict__ beg, int64_t n,
                                                               const int32_t *__restrict__ cnt,
                                                               const long long *__restrict__ off,
                                                               int32_t *__restrict__ seg_u,
                                                               int64_t *__restrict__ seg_e,
                                                               int64_t nseg, int64_t m)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t v = gtid; v < n; v += stride) {
        const int32_t c = cnt[v];
        const int64_t o = off[v], b = beg[v];
        for (int32_t k = 0; k < c; k++) {
            seg_u[o + k] = (int32_t)v;
            seg_e[o + k] = b + (int64_t)k * XB_B64_SEG;
        }
    }
    if (gtid == 0) seg_e[nseg] = m;
}

// ---------------------------------------------------------------------------
// per batch
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(XB_B64_BLOCK) k_b64_clear(uint64_t *__restrict__ vis0,
                                                            uint64_t *__restrict__ F0,
                                                            uint64_t *__restrict__ acc, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        vis0[v] = 0ull;
        F0[v] = 0ull;
        acc[v] = 0ull;
    }
}

// One block of 64 threads: bit i of vis_0 and F_0 at srcs[i]; level-0
// statistics (distinct seed vertices, their degree sum) into out[0..1].
__global__ void __launch_bounds__(64) k_b64_seed(const int32_t *__restrict__ srcs, int32_t nsrc,
                                                 const int32_t *__restrict__ deg,
                                                 uint64_t *__restrict__ vis0,
                                                 uint64_t *__restrict__ F0,
                                                 long long *__restrict__ out)
{
    const int i = threadIdx.x;
    unsigned long long nf = 0, mf = 0;
    if (i < nsrc) {
        const int32_t s = srcs[i];
        const unsigned long long bit = 1ull << i;
        atomicOr(reinterpret_cast<unsigned long long *>(&vis0[s]), bit);
        atomicOr(reinterpret_cast<unsigned long long *>(&F0[s]), bit);
        bool first = true;
        for (int j = 0; j < i; j++)
            if (srcs[j] == s) first = false;
        if (first && deg[s] > 0) {
            nf = 1;
            mf = (unsigned long long)deg[s];
        }
    }
    nf = xb_wave_sum_u64(nf);
    mf = xb_wave_sum_u64(mf);
    if (i == 0) {
        out[0] = (long long)nf;
        out[1] = (long long)mf;
    }
}

// Top-down: lane per segment of a frontier vertex.
__global__ void __launch_bounds__(XB_B64_BLOCK) k_b64_push(XbB64Args a)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nseg; i += stride) {
        const uint64_t f = a.F_cur[a.seg_u[i]];
        if (!f) continue;
        const int64_t e1 = a.seg_e[i + 1];
        for (int64_t e = a.seg_e[i]; e < e1; e++) {
            const int32_t v = a.adj[e];
            const uint64_t m = f & ~a.vis_cur[v];
            if (m) atomicOr(reinterpret_cast<unsigned long long *>(&a.acc[v]), (unsigned long long)m);
        }
    }
}

// Bottom-up: lane per segment of a vertex with unvisited bits; neighbours in
// rounds of XB_B64_PULL_K (all adjacency words of a round loaded, then all
// their frontier masks gathered, each group independent, so a round costs two
// dependent memory trips), early exit once the unvisited bits are all
// covered.  On the level where most vertices still miss some source, nearly
// every segment is scanned whole, and the number of trips per segment bounds
// the level.
#ifndef XB_B64_PULL_K
#define XB_B64_PULL_K 16
#endif
__global__ void __launch_bounds__(XB_B64_BLOCK) k_b64_pull(XbB64Args a)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nseg; i += stride) {
        const int32_t v = a.seg_u[i];
        const uint64_t U = a.all & ~a.vis_cur[v];
        if (!U) continue;
        const int64_t e1 = a.seg_e[i + 1];
        uint64_t got = 0;
        for (int64_t e = a.seg_e[i]; e < e1; e += XB_B64_PULL_K) {
            int32_t u[XB_B64_PULL_K];
#pragma unroll
            for (int k = 0; k < XB_B64_PULL_K; k++) u[k] = e + k < e1 ? a.adj[e + k] : -1;
            uint64_t f[XB_B64_PULL_K];
#pragma unroll
            for (int k = 0; k < XB_B64_PULL_K; k++) f[k] = u[k] >= 0 ? a.F_cur[u[k]] : 0ull;
#pragma unroll
            for (int k = 0; k < XB_B64_PULL_K; k++) got |= f[k];
            if ((got & U) == U) break;
        }
        const uint64_t m = got & U;
        if (m) atomicOr(reinterpret_cast<unsigned long long *>(&a.acc[v]), (unsigned long long)m);
    }
}

// Block totals of K fields into column bid of the partials (thread 0, plain
// store: only this block writes the column).  op_or[k]: OR instead of sum.
template <int K>
__device__ void xb_b64_part_store(long long *part, int64_t pstride, int bid,
                                  unsigned long long (&v)[K], const bool (&op_or)[K])
{
    __shared__ unsigned long long sh[K][XB_B64_BLOCK / 64];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        unsigned long long x = v[k];
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long y = __shfl_xor(x, o);
            x = op_or[k] ? (x | y) : (x + y);
        }
        if (xb_lane() == 0) sh[k][w] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            unsigned long long t = 0;
            for (int j = 0; j < XB_B64_BLOCK / 64; j++) t = op_or[k] ? (t | sh[k][j]) : (t + sh[k][j]);
            part[(int64_t)k * pstride + bid] = (long long)t;
        }
    }
}

__global__ void __launch_bounds__(XB_B64_BLOCK) k_b64_fin(XbB64Args a)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long nf = 0, mf = 0, mu = 0, orr = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += stride) {
        const uint64_t nb = a.acc[v];
        const uint64_t vp = a.vis_cur[v];
        const uint64_t nw = nb & ~vp;
        const uint64_t vn = vp | nw;
        a.vis_next[v] = vn;
        a.F_next[v] = nw;
        if (nb) a.acc[v] = 0ull;
        const bool has_u = (a.all & ~vn) != 0ull;
        if (nw || has_u) {
            const unsigned long long d = (unsigned long long)a.deg[v];
            if (nw) {
                nf += 1;
                mf += d;
                orr |= nw;
            }
            if (has_u) mu += d;
        }
    }
    unsigned long long vv[4] = {nf, mf, mu, orr};
    const bool op[4] = {false, false, false, true};
    xb_b64_part_store<4>(a.part, a.pstride, blockIdx.x, vv, op);
}

// Sums (or ORs) fields [f0, f0+k) over nb columns into out[0..k).
__global__ void __launch_bounds__(1024) k_b64_reduce(const long long *__restrict__ part, int64_t pstride,
                                                     int32_t nb, int32_t f0, int32_t k, int32_t or_field,
                                                     long long *__restrict__ out)
{
    __shared__ unsigned long long sh[16];
    for (int f = f0; f < f0 + k; f++) {
        const bool use_or = f == or_field;
        unsigned long long x = 0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const unsigned long long y = (unsigned long long)part[(int64_t)f * pstride + b];
            x = use_or ? (x | y) : (x + y);
        }
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long y = __shfl_xor(x, o);
            x = use_or ? (x | y) : (x + y);
        }
        if (xb_lane() == 0) sh[threadIdx.x >> 6] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long t = 0;
            for (int j = 0; j < (int)(blockDim.x >> 6); j++) t = use_or ? (t | sh[j]) : (t + sh[j]);
            out[f - f0] = (long long)t;
        }
        __syncthreads();
    }
}

// Final pass: row s of lv for every source s in `done`, lane per vertex
// (a wave covers 64 consecutive vertices, so every row store is one
// coalesced 256-byte run).  The history is monotone, so bit s first appears
// in exactly one slot j0 = level_s(v); the pass reads the slots once and
// builds bit planes P_k (bit s of P_k = bit k of j0) from the newly set bits
// of each slot, then every level is a few bit extracts.  Unreached: -1.
// Also counts reached (source, vertex) pairs and their degree sum per block.
__global__ void __launch_bounds__(XB_B64_BLOCK) k_b64_flush(const uint64_t *__restrict__ hist,
                                                            int32_t nslots, int64_t n, int32_t nsrc,
                                                            uint64_t done, const int32_t *__restrict__ deg,
                                                            int32_t *__restrict__ lv,
                                                            long long *__restrict__ part, int64_t pstride)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long vc = 0, mc = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        uint64_t prev = 0, P[XB_B64_PLANES];
#pragma unroll
        for (int k = 0; k < XB_B64_PLANES; k++) P[k] = 0ull;
        for (int j = 0; j < nslots; j++) {
            const uint64_t h = hist[(int64_t)j * n + v];
            const uint64_t nw = h & ~prev;
            prev = h;
#pragma unroll
            for (int k = 0; k < XB_B64_PLANES; k++)
                if ((j >> k) & 1) P[k] |= nw;
        }
        const uint64_t reached = prev & done;
        vc += (unsigned long long)__popcll(reached);
        mc += (unsigned long long)__popcll(reached) * (unsigned long long)deg[v];
        // 32-bit halves: bit extracts on 32-bit words
        for (int half = 0; half < 2; half++) {
            const uint32_t r = (uint32_t)(reached >> (32 * half));
            const uint32_t dn = (uint32_t)(done >> (32 * half));
            uint32_t q[XB_B64_PLANES];
#pragma unroll
            for (int k = 0; k < XB_B64_PLANES; k++) q[k] = (uint32_t)(P[k] >> (32 * half));
            const int s0 = 32 * half;
            const int s1 = nsrc < s0 + 32 ? nsrc : s0 + 32;
            for (int s = s0; s < s1; s++) {
                const int b = s - s0;
                if (!((dn >> b) & 1u)) continue;
                int32_t lvl = 0;
#pragma unroll
                for (int k = 0; k < XB_B64_PLANES; k++) lvl |= (int32_t)((q[k] >> b) & 1u) << k;
                lv[(int64_t)s * n + v] = ((r >> b) & 1u) ? lvl : -1;
            }
        }
    }
    unsigned long long vv[2] = {vc, mc};
    const bool op[2] = {false, false};
    long long *p = part + (int64_t)XB_B64_F_VC * pstride;
    xb_b64_part_store<2>(p, pstride, blockIdx.x, vv, op);
}

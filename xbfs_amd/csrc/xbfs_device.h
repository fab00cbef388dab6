// Device-side building blocks for the XBFS kernels (gfx950, wave64).
// Included only by .hip translation units.
#ifndef XBFS_DEVICE_H
#define XBFS_DEVICE_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#define XB_WAVE 64

// Per-traversal control block in device memory.
//   qcnt[b][k]   : length of queue buffer b, degree bin k (0 thread, 1 wave, 2 block)
//   list_cnt     : length of the unvisited-vertex list (bottom-up work list)
//   slot_n/m[s]  : vertices discovered for level l and the sum of their
//                  degrees, kept in ring slot s = l & 3 (a level's expansion
//                  writes levels L+1 and, by the bottom-up +2 rule, L+2)
struct XbDevCtl {
    int32_t qcnt[2][3];
    int32_t list_cnt;
    int32_t pad0;
    unsigned long long slot_n[4];
    unsigned long long slot_m[4];
};

__device__ __forceinline__ int xb_lane() { return (int)__lane_id(); }

// Number of set bits of m in lanes below this one.
__device__ __forceinline__ int xb_rank(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Wave-aggregated queue append: the active lanes with pred set write val to
// consecutive slots of q reserved by one atomic from the lowest such lane.
__device__ __forceinline__ void xb_wave_append(int32_t *__restrict__ q, int32_t *cnt,
                                               bool pred, int32_t val)
{
    unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (xb_lane() == leader) base = atomicAdd(cnt, (int)__popcll(m));
    base = __shfl(base, leader);
    if (pred) q[base + xb_rank(m)] = val;
}

__device__ __forceinline__ unsigned long long xb_wave_sum_u64(unsigned long long x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ int xb_wave_incl_scan_i32(int x)
{
    const int ln = xb_lane();
    for (int o = 1; o < XB_WAVE; o <<= 1) {
        int y = __shfl_up(x, (unsigned)o);
        if (ln >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ long long xb_wave_incl_scan_i64(long long x)
{
    const int ln = xb_lane();
    for (int o = 1; o < XB_WAVE; o <<= 1) {
        long long y = __shfl_up(x, (unsigned)o);
        if (ln >= o) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one int64 per thread; returns the exclusive
// prefix and writes the block total to *total (all threads).
__device__ __forceinline__ long long xb_block_excl_scan_i64(long long x, long long *total)
{
    __shared__ long long sh[17];
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    long long inc = xb_wave_incl_scan_i64(x);
    if (xb_lane() == XB_WAVE - 1) sh[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long run = 0;
        for (int i = 0; i < nw; i++) {
            long long t = sh[i];
            sh[i] = run;
            run += t;
        }
        sh[16] = run;
    }
    __syncthreads();
    long long r = sh[w] + inc - x;
    *total = sh[16];
    __syncthreads();
    return r;
}

#endif

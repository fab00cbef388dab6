// Traversal state shared by the level phases (device) and their drivers
// (host loop or persistent kernel).
#ifndef XBFS_ENGINE_H
#define XBFS_ENGINE_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_policy.h"

#define XB_BINS 4               /* thread, wave, block, grid degree bins          */
#define XB_QBUFS 3              /* queue buffers, ring by level % 3               */
#define XB_FRINGS 4             /* frontier bitmaps, ring by level & 3            */
#define XB_BU_TILE_WORDS 32     /* bitmap words per bottom-up tile (2048 vertices) */
#ifndef XB_GCAP
#define XB_GCAP 2048            /* vertices per concatenated top-down pass, at most; a pass
                                   takes 2 per thread, so 2 * blockDim.x bounds it too */
#endif
#define XB_GCAP_SMALL 256       /* the same on small-frontier levels (32 active blocks) */
#define XB_PBLOCK 256           /* block size of the per-phase kernels            */
#define XB_PERSIST_BLOCK 1024   /* block size of the persistent kernel (1 per CU) */
// Bottom-up phases read the hub records and, in scans, the row starts once
// per level.  When the hub table (8 bytes per vertex) exceeds this size, a
// quarter of MI355X's 256 MB Infinity Cache, those loads are non-temporal (a
// stream that is not re-read within the level); below it they are plain
// loads.  Switched at run time within one build (XBFS_NT_BU), the two kinds
// time the same on RMAT-22, RMAT-23, RMAT-24 and Erdos-Renyi 2^24 graphs.
#define XB_NT_HUB_BYTES (64ll << 20)
#ifndef XB_BU_LANE_MAX
#define XB_BU_LANE_MAX 32       /* bottom-up: edges one lane scans before the wave takes over */
#endif

// Control block in device memory.  Ring slots are indexed by level & 3,
// queue counters by level % 3 (see the zeroing schedule in ph_level_prep).
// Device code touches it only with atomics (RMW, loads and stores at agent
// scope): MI355X's eight XCDs keep private L2 caches, and a line dirtied by
// plain stores in one XCD's L2 must never hold words other XCDs update
// atomically.  The barrier words sit on cache lines of their own.
struct alignas(128) XbCounter {
    int32_t v;
    int32_t pad[31];
};

struct alignas(256) XbCtl {
    uint32_t bar_count;                 // grid barrier arrivals
    uint32_t bar_pad0[31];
    uint32_t bar_gen;                   // grid barrier generation
    uint32_t bar_pad1[31];
    uint32_t sbar_count;                // small-frontier barrier (active blocks only)
    uint32_t sbar_pad0[31];
    uint32_t sbar_gen;
    uint32_t sbar_pad1[31];
    uint32_t heartbeat;                 // bumped once per small-frontier level
    uint32_t hb_pad[31];
    int32_t error;                      // 1 barrier timeout, 2 level cap, 3 blocks not co-resident
    int32_t err_pad0;
    unsigned long long wait_ticks;      // bound of a barrier wait (wall-clock ticks), per launch
    unsigned long long res_ticks;       // plain launch: bound of the first barrier's wait, 0 otherwise
    int32_t err_pad[26];
    XbCounter qcnt[XB_QBUFS][XB_BINS];  // queue lengths, one cache line each
    int32_t n_levels, max_level, pad0, pad1;
    long long v_cc, m_cc;
    long long seed_iso;                 // multi-source: distinct degree-0 seeds (level 0, no edges)
    long long pad2;
    long long sum[32];                  // host-driven loop: reduced statistics (k_part_reduce)
};

// Per-level statistics live in per-block partial sums (no atomics): block b
// adds its totals to part[f * XB_PMAX + b] with plain stores (only block b
// touches column b within a phase) and readers sum the column after the
// barrier.  Fields: ring slot r = level & 3 -> f = 3r + {0: vertices
// discovered for that level, 1: their degree sum, 2: of which by the +2 rule};
// sweep slot s = level & 1 -> f = 12 + 2s + {0: count, 1: degree sum};
// work of a bottom-up level L, kept in ring slot r = (L + 1) & 3 (the level
// it discovers, zeroed on the same schedule) -> f = 16 + 3r + {0: candidates
// probed at their hub, 1: candidates whose rows were scanned, 2: adjacency
// words those scans read}.
#define XB_PMAX 4096
#define XB_NFIELDS 28
__host__ __device__ inline int xb_f_ring(int32_t level, int k) { return 3 * (level & 3) + k; }
__host__ __device__ inline int xb_f_sweep(int32_t level, int k) { return 12 + 2 * (level & 1) + k; }
__host__ __device__ inline int xb_f_bu(int32_t level, int k) { return 16 + 3 * ((level + 1) & 3) + k; }

// Per-level record written on the device by the persistent kernel.
struct XbDevRec {
    int32_t level, mode, tag, htag;
    long long n_f, m_f, n_u, m_u;
    long long t0, t1; // wall clock ticks
    long long bu_cand, bu_scan_v, bu_scan_e; // bottom-up work (xb_f_bu fields), 0 for top-down
};

struct XbArgs {
    int32_t *level;
    int32_t *parent;  // NULL when parents are not wanted
    const int64_t *beg;
    const int32_t *adj;
    const int32_t *adj_bu; // rows by neighbour degree, hub first (hub_first), else adj
    const uint64_t *nz; // bottom-up candidates: bit v: deg(v) > 0, or v in the source's component
    const int2 *hd;     // {largest-degree neighbour, degree} per vertex
    uint64_t *vis;      // bit v: level[v] != -1
    uint64_t *fbm;      // XB_FRINGS bitmaps of nwords: bit v of ring l&3 <=> level[v] == l
    int32_t *qbuf;      // XB_QBUFS x XB_BINS x n
    XbCtl *ctl;
    struct XbLoopShared *lsg; // loop state handed to parked blocks (persistent kernel)
    long long *part; // XB_NFIELDS x XB_PMAX partial statistics
    XbDevRec *recs;
    int32_t rec_cap;
    const int8_t *sched;
    int32_t sched_len;
    const float *mlp;   // learned policy weights (NULL = none)
    int64_t n, m, nwords, ntiles, n_nz;
    int64_t n_target; // the traversal is complete once this many vertices have levels:
                      // n_nz, or the size of the source's component when it is known
    int64_t fstride; // words between frontier rings: nwords rounded up to a tile (256 B)
    const int32_t *srcs; // multi-source seeds (device), used when nsrc > 0
    int32_t nsrc;
    int32_t hub_first; // 1: every row of adj_bu starts with its hub (hd.x)
    int32_t nt_bu;     // 1: bottom-up hub records and row starts read non-temporally
    int32_t src, t1, t2, t3;
    int32_t policy;
    float alpha;
    int32_t max_levels;
};

__host__ __device__ inline int32_t *xb_q(const XbArgs &a, int b, int k)
{
    return a.qbuf + ((int64_t)b * XB_BINS + k) * a.n;
}
__host__ __device__ inline uint64_t *xb_f(const XbArgs &a, int32_t level)
{
    return a.fbm + (int64_t)(level & 3) * a.fstride;
}

// Block-wide LDS shared by the phases (static).
union XbSmem {
    struct {
        long long pref[XB_GCAP + 1]; // exclusive degree prefix of the vertices of one pass
        long long b[XB_GCAP];        // their row starts
        int32_t u[XB_GCAP];
    } td;
    struct {
        float x[8], y[4];
    } mlp;
};

#define XB_CHUNK_WORDS 16 /* bitmap words per wave-owned chunk: 1024 vertices, one 128-B line */

// Per-wave LDS region (dynamic shared memory, one per wave of the block).
struct XbWaveLds {
    int32_t list[XB_CHUNK_WORDS * 64];      // the chunk's vertices to process
    unsigned long long f1[XB_CHUNK_WORDS];  // bottom-up: new level-(L+1) bits
    unsigned long long va[XB_CHUNK_WORDS];  // bottom-up: new visited bits
};

// Loop state every block (or the host) tracks identically.
struct XbLoop {
    int32_t L;
    int32_t prev_mode;  // -1 before level 0
    int32_t q_valid;    // queue buffer L%3 holds exactly the level-L frontier
    int32_t need_sweep; // frontier statistics of level L must come from a sweep
    long long cum_n, cum_m;
    XbfsHeurState hs;
};

// Block-uniform state of the persistent kernel's level loop (LDS), and its
// copy in device memory when the parked blocks rejoin.
struct XbLoopShared {
    XbLoop st;
    long long n_f, m_f, n_u, m_u, p2_next;
    int32_t open, mode, htag, used, small;
};

// Small-frontier levels: a scan-free top-down level whose queue is exact and
// whose frontier is at most XB_SMALL_NF vertices and XB_SMALL_MF edges runs
// on blocks [0, XB_SMALL_BLOCKS) only, with a barrier among those blocks;
// the other blocks park until a level needs the whole grid.
#define XB_SMALL_BLOCKS 32
#ifndef XB_SMALL_NF
#define XB_SMALL_NF 16384
#endif
#ifndef XB_SMALL_MF
#define XB_SMALL_MF 65536
#endif

__host__ __device__ inline void xb_loop_init(XbLoop *s)
{
    s->L = 0;
    s->prev_mode = -1;
    s->q_valid = 1;
    s->need_sweep = 0;
    s->cum_n = 0;
    s->cum_m = 0;
    xbfs_heur_init(&s->hs);
}

#endif

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
#define XB_GCAP 512             /* grid-bin vertices scanned per block pass       */
#define XB_PBLOCK 256           /* block size of the per-phase kernels            */
#define XB_PERSIST_BLOCK 1024   /* block size of the persistent kernel (1 per CU) */

// Control block in device memory.  Ring slots are indexed by level & 3,
// queue counters by level % 3 (see the zeroing schedule in ph_level_prep).
// Device code touches it only with atomics (RMW, loads and stores at agent
// scope): MI355X's eight XCDs keep private L2 caches, and a line dirtied by
// plain stores in one XCD's L2 must never hold words other XCDs update
// atomically.  The barrier words sit on cache lines of their own.
struct alignas(256) XbCtl {
    uint32_t bar_count;                 // grid barrier arrivals
    uint32_t bar_pad0[31];
    uint32_t bar_gen;                   // grid barrier generation
    uint32_t bar_pad1[31];
    int32_t error;                      // 1 barrier timeout, 2 level cap
    int32_t err_pad[31];
    int32_t qcnt[XB_QBUFS][XB_BINS];    // queue lengths
    unsigned long long slot_n[4];       // vertices discovered for a level
    unsigned long long slot_m[4];       // ... and the sum of their degrees
    unsigned long long slot_p2[4];      // ... of which set early by the +2 rule
    unsigned long long sw_n[2], sw_m[2]; // frontier statistics from a bitmap sweep
    uint32_t tile_ctr[XB_QBUFS];        // bottom-up tile dispenser, ring by level % 3
    int32_t n_levels, max_level, pad;
    long long v_cc, m_cc;
};

// Per-level record written on the device by the persistent kernel.
struct XbDevRec {
    int32_t level, mode, tag, htag;
    long long n_f, m_f, n_u, m_u;
    long long t0, t1; // wall clock ticks
};

struct XbArgs {
    int32_t *level;
    int32_t *parent;  // NULL when parents are not wanted
    const int64_t *beg;
    const int32_t *adj;
    const uint64_t *nz; // bit v: deg(v) > 0
    uint64_t *vis;      // bit v: level[v] != -1
    uint64_t *fbm;      // XB_FRINGS bitmaps of nwords: bit v of ring l&3 <=> level[v] == l
    int32_t *qbuf;      // XB_QBUFS x XB_BINS x n
    XbCtl *ctl;
    XbDevRec *recs;
    int32_t rec_cap;
    const int8_t *sched;
    int32_t sched_len;
    const float *mlp;   // learned policy weights (NULL = none)
    int64_t n, m, nwords, ntiles, n_nz;
    int64_t fstride; // words between frontier rings: nwords rounded up to a tile (256 B)
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

// One LDS buffer shared by all phases (the persistent kernel runs them in
// sequence, so their scratch can overlap).
union XbSmem {
    struct {
        int32_t list[XB_BU_TILE_WORDS * 64];      // candidate vertices of the tile
        unsigned long long vacc[XB_BU_TILE_WORDS]; // candidate masks, then new visited bits
        unsigned long long f1acc[XB_BU_TILE_WORDS]; // level L+1 discoveries
        unsigned long long f2acc[XB_BU_TILE_WORDS]; // +2 discoveries
        int woff[XB_BU_TILE_WORDS + 1];            // exclusive prefix of candidates per word
        uint32_t tile;
    } bu;
    struct {
        long long pref[XB_GCAP + 1]; // exclusive degree prefix of grid-bin vertices
        int32_t u[XB_GCAP];
    } td;
    struct {
        int32_t list[XB_BU_TILE_WORDS * 64];       // frontier vertices of the tile
        unsigned long long words[XB_BU_TILE_WORDS];
        int woff[XB_BU_TILE_WORDS + 1];
        int wcnt[16 * XB_BINS];                    // per-wave bin counts, then offsets
        int base[XB_BINS];                         // reserved queue positions
    } cp;
    struct {
        float x[8], h1[32], h2[32], y[4];
    } mlp;
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

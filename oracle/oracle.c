/*
 * CPU oracle for XBFS.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * The product path (libxbfs.so, the ./bfs driver) never links or calls it.
 *
 * Functions and what pins each one (tests/test_oracle.py):
 *   oracle_bfs_scan      literal restatement of the reference's serial
 *                        level-synchronous BFS, bfs_gpu_update.cu:104-127.
 *                        Pinned by closed forms (path, grid, star, complete,
 *                        binary tree, disconnected, degree-0 source), brute
 *                        force (Floyd-Warshall) on random graphs <= 64 vertices,
 *                        and agreement with oracle_bfs_queue.
 *   oracle_bfs_queue     serial FIFO-queue BFS (levels + parents).  Same pins,
 *                        plus oracle_validate on every output.
 *   oracle_validate      Graph500-style invariants; pinned by planted
 *                        mistakes (wrong level, wrong parent, unreached
 *                        neighbour, src level != 0) that each must fail.
 *   oracle_level_counts  count[L] = sum of degrees over level==L, the
 *                        reference's per-level frontier-edge count
 *                        (bfs_adaptiveFQ_wb_async.cuh:1042-1046); pinned by
 *                        closed forms (star, path).
 *   oracle_heuristic_tags  restatement of the reference's direction/queue
 *                        decision block bfs_adaptiveFQ_wb_async.cuh:1047-1078
 *                        (with mode branches 1078-1167).  Pinned by
 *                        hand-derived tag sequences on constructed count
 *                        sequences that hit every branch (TD_adaptive
 *                        triggers, hub_before / hub_after, last phase).
 *   oracle_omp_td_bfs    OpenMP top-down frontier-queue BFS with CAS on
 *                        int32 levels: the CPU baseline.  Pinned by
 *                        agreement with oracle_bfs_queue.
 *   oracle_exclusive_scan  std::exclusive_scan restated (int32 in, int64 out).
 *   oracle_bfs_queue_rows  oracle_bfs_queue from each of k sources, rows in
 *                        parallel threads.  Pinned by row-by-row equality
 *                        with oracle_bfs_queue.
 *   oracle_bfs_multi     serial FIFO BFS from a set of seeds (the batched
 *                        modes' oracle; no reference counterpart).  Pinned by
 *                        the identity level = min over seeds of the
 *                        single-source oracle level, and a closed form (path
 *                        with two seeds: min(|i-a|, |i-b|)).
 *   oracle_validate_multi  oracle_validate with a seed set (level 0 exactly
 *                        at the seeds); pinned by planted mistakes.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <time.h>

/* bfs_gpu_update.cu:104-127: status[]=-1, status[src]=0, then for k=0.. scan
 * every vertex; a vertex at level k marks each unvisited neighbour k+1; stop
 * after a pass that marked nothing.  Returns k at loop exit: passes 0..maxl
 * run, the last marks nothing, so k = max level + 1 (the number of levels),
 * the value the reference prints as "k =" (main:128). */
int32_t oracle_bfs_scan(int64_t n, const int64_t *beg, const int32_t *adj,
                        int32_t src, int32_t *status)
{
    for (int64_t i = 0; i < n; i++) status[i] = -1;
    status[src] = 0;
    int done = 0;
    int32_t k;
    for (k = 0; !done; k++) {
        done = 1;
        for (int64_t i = 0; i < n; i++) {
            if (status[i] == k) {
                for (int64_t j = beg[i]; j < beg[i + 1]; j++) {
                    if (status[adj[j]] == -1) {
                        done = 0;
                        status[adj[j]] = k + 1;
                    }
                }
            }
        }
    }
    return k;
}

/* Serial FIFO BFS.  parent[src] = src, unreached parent = -1.
 * Returns the number of levels (max level + 1). */
int32_t oracle_bfs_queue(int64_t n, const int64_t *beg, const int32_t *adj,
                         int32_t src, int32_t *level, int32_t *parent)
{
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) level[i] = -1;
    if (parent)
        for (int64_t i = 0; i < n; i++) parent[i] = -1;
    int64_t head = 0, tail = 0;
    level[src] = 0;
    if (parent) parent[src] = src;
    q[tail++] = src;
    int32_t maxl = 0;
    while (head < tail) {
        int32_t u = q[head++];
        int32_t lu = level[u];
        if (lu > maxl) maxl = lu;
        for (int64_t j = beg[u]; j < beg[u + 1]; j++) {
            int32_t v = adj[j];
            if (level[v] == -1) {
                level[v] = lu + 1;
                if (parent) parent[v] = u;
                q[tail++] = v;
            }
        }
    }
    free(q);
    return maxl + 1;
}

/* k independent single-source BFSs: row i of levels (n entries) is
 * oracle_bfs_queue from srcs[i].  Rows run in parallel (one OpenMP thread per
 * row at a time), each exactly the serial FIFO BFS, so full-size batches can
 * be checked bit for bit in reasonable time. */
void oracle_bfs_queue_rows(int64_t n, const int64_t *beg, const int32_t *adj,
                           const int32_t *srcs, int32_t k, int32_t *levels)
{
    #pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < k; i++)
        oracle_bfs_queue(n, beg, adj, srcs[i], levels + (int64_t)i * n, NULL);
}

/* Serial FIFO BFS from every distinct seed at once (level 0, parent = the
 * seed itself); level[v] = hops to the nearest seed.  Returns max level + 1
 * (0 when k == 0). */
int32_t oracle_bfs_multi(int64_t n, const int64_t *beg, const int32_t *adj,
                         const int32_t *srcs, int32_t k, int32_t *level, int32_t *parent)
{
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) level[i] = -1;
    if (parent)
        for (int64_t i = 0; i < n; i++) parent[i] = -1;
    int64_t head = 0, tail = 0;
    for (int32_t i = 0; i < k; i++) {
        int32_t s = srcs[i];
        if (level[s] == -1) {
            level[s] = 0;
            if (parent) parent[s] = s;
            q[tail++] = s;
        }
    }
    int32_t maxl = k > 0 ? 0 : -1;
    while (head < tail) {
        int32_t u = q[head++];
        int32_t lu = level[u];
        if (lu > maxl) maxl = lu;
        for (int64_t j = beg[u]; j < beg[u + 1]; j++) {
            int32_t v = adj[j];
            if (level[v] == -1) {
                level[v] = lu + 1;
                if (parent) parent[v] = u;
                q[tail++] = v;
            }
        }
    }
    free(q);
    return maxl + 1;
}

/* Validation of a multi-source result; returns the number of violated checks.
 *  (1) level[v] == 0 exactly for the seeds (parent[seed] == seed);
 *  (2) every edge (u,v) with u reached has v reached, |level[u]-level[v]| <= 1;
 *  (3) every reached non-seed v has a neighbour at level[v]-1 (and, with
 *      parents, parent[v] is such a neighbour); unreached v has parent -1.
 * Together these make level the distance to the nearest seed. */
int64_t oracle_validate_multi(int64_t n, const int64_t *beg, const int32_t *adj,
                              const int32_t *srcs, int32_t k, const int32_t *level,
                              const int32_t *parent)
{
    int64_t bad = 0;
    unsigned char *seed = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
    for (int32_t i = 0; i < k; i++) seed[srcs[i]] = 1;
    #pragma omp parallel for reduction(+:bad) schedule(dynamic, 4096)
    for (int64_t u = 0; u < n; u++) {
        int32_t lu = level[u];
        if (seed[u]) {
            if (lu != 0) bad++;
            if (parent && parent[u] != u) bad++;
        } else if (lu == 0) {
            bad++;
        }
        if (lu < -1) { bad++; continue; }
        if (lu == -1) {
            if (parent && parent[u] != -1) bad++;
            continue;
        }
        int has_pred = seed[u];
        for (int64_t j = beg[u]; j < beg[u + 1]; j++) {
            int32_t v = adj[j];
            int32_t lv = level[v];
            if (lv == -1) { bad++; continue; }
            int32_t d = lu - lv;
            if (d > 1 || d < -1) bad++;
            if (lv == lu - 1) has_pred = 1;
        }
        if (!has_pred) bad++;
        if (parent && !seed[u]) {
            int32_t p = parent[u];
            if (p < 0 || p >= n || level[p] != lu - 1) { bad++; continue; }
            int found = 0;
            for (int64_t j = beg[u]; j < beg[u + 1]; j++)
                if (adj[j] == p) { found = 1; break; }
            if (!found) bad++;
        }
    }
    free(seed);
    return bad;
}

/* Graph500-style validation of a BFS result.  Returns the number of violated
 * checks (0 = valid).  Checks:
 *  (1) level[src]==0;
 *  (2) every edge (u,v) with u reached has v reached and |level[u]-level[v]|<=1;
 *  (3) every reached v != src has a neighbour at level[v]-1;
 *  (4) if parent != NULL: parent[src]==src, every reached v != src has
 *      parent[v] adjacent to v with level[parent[v]]==level[v]-1, and
 *      unreached v has parent -1.
 * (1)+(2)+(3) together imply level == shortest-path distance, so a result
 * that passes is bit-exact with any correct BFS. */
int64_t oracle_validate(int64_t n, const int64_t *beg, const int32_t *adj,
                        int32_t src, const int32_t *level, const int32_t *parent)
{
    int64_t bad = 0;
    if (level[src] != 0) bad++;
    if (parent && parent[src] != src) bad++;
    #pragma omp parallel for reduction(+:bad) schedule(dynamic, 4096)
    for (int64_t u = 0; u < n; u++) {
        int32_t lu = level[u];
        if (lu < -1) { bad++; continue; }
        if (lu == -1) {
            if (parent && parent[u] != -1) bad++;
            continue;
        }
        int has_pred = (u == src);
        for (int64_t j = beg[u]; j < beg[u + 1]; j++) {
            int32_t v = adj[j];
            int32_t lv = level[v];
            if (lv == -1) { bad++; continue; }
            int32_t d = lu - lv;
            if (d > 1 || d < -1) bad++;
            if (lv == lu - 1) has_pred = 1;
        }
        if (!has_pred) bad++;
        if (parent && u != src) {
            int32_t p = parent[u];
            if (p < 0 || p >= n || level[p] != lu - 1) { bad++; continue; }
            int found = 0;
            for (int64_t j = beg[u]; j < beg[u + 1]; j++)
                if (adj[j] == p) { found = 1; break; }
            if (!found) bad++;
        }
    }
    return bad;
}

/* count[L] = sum over level[v]==L of deg(v), for L in [0, cap).  This is the
 * host loop at bfs_adaptiveFQ_wb_async.cuh:1042-1046 evaluated on final
 * levels (levels <= L are final at the start of level L).  Also fills
 * nverts[L] (vertices per level) when non-NULL.  Returns max level + 1. */
int32_t oracle_level_counts(int64_t n, const int64_t *beg, const int32_t *level,
                            int64_t *count, int64_t *nverts, int32_t cap)
{
    for (int32_t i = 0; i < cap; i++) { count[i] = 0; if (nverts) nverts[i] = 0; }
    int32_t maxl = -1;
    for (int64_t v = 0; v < n; v++) {
        int32_t l = level[v];
        if (l < 0) continue;
        if (l > maxl) maxl = l;
        if (l < cap) {
            count[l] += beg[v + 1] - beg[v];
            if (nverts) nverts[l]++;
        }
    }
    return maxl + 1;
}

#define ORACLE_INFINITE 10000 /* bfs_adaptiveFQ_wb_async.cuh:29 */

/* Restatement of the reference's per-level decision
 * (bfs_adaptiveFQ_wb_async.cuh:1047-1078 and the mode branches 1078-1167).
 * Input: count[0..nlev-1] frontier-edge counts, E = total directed edges,
 * alpha.  Output tag[L] in {0: top-down scan-free (incl. "last phase"),
 * 1: top-down single-scan, -1: bottom-up} and sub[L] in {0: scan-free,
 * 2: last phase, 1: single-scan, -1: bottom-up} when sub != NULL.
 * The reference reads tag[-1] and count[-1] at level 0; both reads are inert
 * (they only assign hub_* = 0 or are short-circuited), modelled here as
 * tag[-1] = -2 and count[-1] = 0.
 * Returns TD_adaptive's final value. */
int32_t oracle_heuristic_tags(const int64_t *count, int32_t nlev, int64_t E,
                              float alpha, int32_t *tag, int32_t *sub)
{
    int32_t TD_adaptive = ORACLE_INFINITE;
    int is_change = 0;
    int32_t hub_before = 0, hub_after = 0;
    for (int32_t L = 0; L < nlev; L++) {
        int64_t c = count[L];
        int64_t cp = (L > 0) ? count[L - 1] : 0;
        int32_t tp = (L > 0) ? tag[L - 1] : -2;
        if (!is_change) {
            if (L == 0 && c >= ORACLE_INFINITE) {
                is_change = 1;
                TD_adaptive = L + 1;
            } else if (L > 0 && (cp >= ORACLE_INFINITE && c >= (int64_t)ORACLE_INFINITE * 10 &&
                                 c / cp >= 8)) {
                is_change = 1;
                TD_adaptive = L;
            } else if ((L > 0 && c >= (int64_t)ORACLE_INFINITE * 100) ||
                       (cp < ORACLE_INFINITE && c >= (int64_t)ORACLE_INFINITE * 10 &&
                        cp != 0 && c / cp >= 1000)) {
                /* cp != 0 guards a division the reference never reaches with
                 * cp == 0: a non-empty level L needs an edge out of level L-1,
                 * and at L == 0 the first branch already took c >= 10000. */
                is_change = 1;
                TD_adaptive = L;
            }
        }
        float ratio = (float)c / (float)E;
        if (ratio >= alpha && tp >= 0) hub_before = L;
        if (ratio < alpha && tp == -1) hub_after = L;
        int32_t t, s;
        if (ratio < alpha) {
            t = 0;
            if (hub_before == 0 && L < TD_adaptive) {
                s = 0;
            } else if (hub_after > 0 && L > hub_after) {
                s = 2;
            } else {
                t = 1;
                s = 1;
            }
        } else {
            t = -1;
            s = -1;
        }
        tag[L] = t;
        if (sub) sub[L] = s;
    }
    return TD_adaptive;
}

/* OpenMP top-down BFS: frontier queue, parallel loop over the frontier,
 * CAS on int32 levels, per-thread local next-queue buffers merged by an
 * atomic reservation.  Returns wall seconds (steady clock) of the traversal
 * (level init to empty frontier). */
static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double oracle_omp_td_bfs(int64_t n, const int64_t *beg, const int32_t *adj,
                         int32_t src, int32_t *level, int32_t nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t *nxt = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    double t0 = now_s();
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) level[i] = -1;
    level[src] = 0;
    int64_t ncur = 1;
    cur[0] = src;
    int32_t L = 0;
    while (ncur > 0) {
        int64_t nnext = 0;
        #pragma omp parallel
        {
            int32_t buf[1024];
            int nb = 0;
            #pragma omp for schedule(dynamic, 64)
            for (int64_t f = 0; f < ncur; f++) {
                int32_t u = cur[f];
                for (int64_t j = beg[u]; j < beg[u + 1]; j++) {
                    int32_t v = adj[j];
                    if (__atomic_load_n(&level[v], __ATOMIC_RELAXED) == -1) {
                        int32_t expect = -1;
                        if (__atomic_compare_exchange_n(&level[v], &expect, L + 1, 0,
                                                        __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                            buf[nb++] = v;
                            if (nb == 1024) {
                                int64_t at = __atomic_fetch_add(&nnext, nb, __ATOMIC_RELAXED);
                                memcpy(nxt + at, buf, sizeof(int32_t) * nb);
                                nb = 0;
                            }
                        }
                    }
                }
            }
            if (nb) {
                int64_t at = __atomic_fetch_add(&nnext, nb, __ATOMIC_RELAXED);
                memcpy(nxt + at, buf, sizeof(int32_t) * nb);
            }
        }
        int32_t *t = cur; cur = nxt; nxt = t;
        ncur = nnext;
        L++;
    }
    double t1 = now_s();
    free(cur);
    free(nxt);
    return t1 - t0;
}

/* Exclusive prefix sum of int32 into int64 (std::exclusive_scan restated). */
void oracle_exclusive_scan(int64_t n, const int32_t *in, int64_t *out)
{
    int64_t s = 0;
    for (int64_t i = 0; i < n; i++) { out[i] = s; s += in[i]; }
}

int32_t oracle_max_threads(void) { return omp_get_max_threads(); }

/*
 * xbfs.h - C ABI of the MI355X-native XBFS library (libxbfs.so).
 *
 * Plain C, plain pointers and sizes.  Every entry point returns an int error
 * code (XBFS_OK == 0) instead of exiting the process the way the reference's
 * H_ERR macro does (bfs_single_scan.cuh:11-19).
 *
 * Entry points and the reference interface each one replaces:
 *   xbfs_graph_load / xbfs_graph_free
 *       graph<int,long,int,int,long,char>::graph(beg, csr, weight)
 *       graph.hpp:4-104, graph.h:18-36 (same size rules and checks).
 *   xbfs_graph_write
 *       the output side of CSRGenerator/tuple_text_to_bin.cpp:359-360, but in
 *       the int32 adjacency layout the loader reads (bfs_gpu_update.cu:52).
 *   xbfs_graph_from_tuples (and the text_to_bin command line)
 *       CSRGenerator/tuple_text_to_bin.cpp:59-426 (text edge list -> CSR).
 *   xbfs_graph_from_edge_text
 *       no reference counterpart: a deduplicated, loop-free edge-list reader.
 *   xbfs_gen_rmat / xbfs_gen_er / xbfs_gen_grid
 *       no reference counterpart (the reference ships no generator); seeded
 *       inputs for the benchmark configurations.
 *   xbfs_create / xbfs_destroy
 *       the device allocations and uploads of bfs_gpu_update.cu:138-240
 *       (beg_cu, csr_cu, Status_cu, frontier queues, counters).
 *   xbfs_bfs
 *       BFS_ControlFlow<int,long>(45 args), bfs_adaptiveFQ_wb_async.cuh:934-1258,
 *       called at bfs_gpu_update.cu:251: one adaptive direction-optimizing
 *       BFS from src; per-vertex int32 levels (-1 = unreached) as in the
 *       Status array copied back at bfs_gpu_update.cu:255; per-level
 *       ratio/tag/iter records (bfs_adaptiveFQ_wb_async.cuh:1067,1077,1097).
 *   xbfs_degree_stats
 *       the degree histogram of bfs_gpu_update.cu:76-89, workload_gap1
 *       (workload_gap.cuh:27-38) and hub_num (bfs_adaptiveFQ_wb_async.cuh:953-958).
 *   xbfs_exclusive_scan_i32
 *       my_prefixsum / IntraBlockSum / InterBlockSum, prefix_sum1.cuh:13-108,
 *       as a device-only scan (no host round trip).
 *   xbfs_unvisited_list
 *       ballot_counting + FQ_gen, bfs_adaptiveFQ_wb_async.cuh:189-284: the
 *       ascending list of unvisited vertices, built by 64-bit wave ballots.
 *   xbfs_bfs_multi, xbfs_bfs_multigraph, xbfs_graph_concat,
 *   xbfs_bfs_batch64, xbfs_autotune, xbfs_policy_*, xbfs_mlp_*
 *       no reference counterpart (batched traversal, the bin/grid
 *       autotuner, and the learned direction/queue policy that replaces the
 *       decision block bfs_adaptiveFQ_wb_async.cuh:1047-1078).
 */
#ifndef XBFS_H
#define XBFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define XBFS_OK 0
#define XBFS_ERR_ARG 1          /* invalid argument (null, out of range)      */
#define XBFS_ERR_HIP 2          /* HIP runtime call failed                    */
#define XBFS_ERR_NOMEM 3        /* host or device allocation failed           */
#define XBFS_ERR_IO 4           /* file missing, unreadable or truncated      */
#define XBFS_ERR_FORMAT 5       /* file contents violate the CSR format       */
#define XBFS_ERR_UNSUPPORTED 6  /* configuration not supported                */
#define XBFS_ERR_LIMIT 7        /* a size limit was exceeded                  */
#define XBFS_ERR_DEVICE 8       /* device-side check failed (barrier timeout) */

const char *xbfs_strerror(int code);

/* ---- host graphs (CSR: beg_pos int64[n+1], adj int32[m]) ---------------- */
typedef struct {
    int64_t n;        /* vertices (vert_count)                                */
    int64_t m;        /* directed CSR entries (edge_count = beg_pos[n])       */
    int64_t *beg_pos; /* int64[n+1]                                           */
    int32_t *adj;     /* int32[m]                                             */
    int8_t *weight;   /* int8[m] when a weight file was read, else NULL       */
} xbfs_host_graph;

/* Loads <beg_file> (int64 beg_pos, n = filesize/8 - 1) and <csr_file>
 * (int32 adjacency, beg_pos[n] entries).  weight_file may be NULL or missing
 * (the reference then prints "Weight file cannot open" and continues). */
int xbfs_graph_load(const char *beg_file, const char *csr_file,
                    const char *weight_file, xbfs_host_graph *g);
/* Writes <prefix>_beg_pos.bin and <prefix>_csr.bin, and <prefix>_weight.bin
 * (int32 per entry) when g->weight is set. */
int xbfs_graph_write(const char *prefix, const xbfs_host_graph *g);
/* Reads a whitespace-separated "src dst [weight]" text edge list (lines
 * starting with '#' or '%' skipped).  reverse != 0 also inserts dst->src.
 * Self-loops dropped, rows sorted and deduplicated: a clean undirected graph
 * for traversal (no reference counterpart; compare xbfs_graph_from_tuples). */
int xbfs_graph_from_edge_text(const char *path, int reverse, xbfs_host_graph *g);
/* The reference converter's semantics, CSRGenerator/tuple_text_to_bin.cpp:59-426:
 * skip_lines header lines, then whitespace-separated ids read as (src, dst)
 * pairs; n = largest id + 1; rows in file order, duplicates and self-loops
 * kept; reverse != 0 also appends src to dst's row (m = 2 x pairs, else m =
 * pairs); a weight 1..63 per pair in g->weight.  Written by xbfs_graph_write
 * as int32 adjacency and weights, the layout the loader reads (the reference
 * writes int64, which its own int32 loader cannot read). */
int xbfs_graph_from_tuples(const char *path, int reverse, int64_t skip_lines, xbfs_host_graph *g);
void xbfs_graph_free(xbfs_host_graph *g);

/* RMAT/Kronecker generator: 2^scale vertices, ef*2^scale sampled edges with
 * quadrant probabilities a, b, c (d = 1-a-b-c), counter-based RNG seeded by
 * seed; vertex ids permuted by a seeded shuffle when relabel != 0;
 * symmetrized, self-loops and duplicates removed, rows sorted ascending. */
int xbfs_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c,
                  uint64_t seed, int32_t relabel, xbfs_host_graph *g);
/* Erdos-Renyi G(n, m): m uniform endpoint pairs, symmetrized, deduplicated. */
int xbfs_gen_er(int64_t n, int64_t m, uint64_t seed, xbfs_host_graph *g);
/* rows x cols 4-neighbour grid, vertex id r*cols+c. */
int xbfs_gen_grid(int32_t rows, int32_t cols, xbfs_host_graph *g);

/* Degree statistics printed by the reference driver. */
typedef struct {
    int64_t small;      /* deg <= 32                 (bfs_gpu_update.cu:79)  */
    int64_t medium;     /* 32 < deg <= 256           (bfs_gpu_update.cu:81)  */
    int64_t large;      /* deg > 256                 (bfs_gpu_update.cu:84)  */
    int64_t max_degree; /* workload_gap1             (workload_gap.cuh:27)   */
    int64_t max_vertex; /* vertex id of max_degree                            */
    int64_t hub_num;    /* deg > 1000 (theta, bfs_adaptiveFQ_wb_async.cuh:954) */
    int64_t isolated;   /* deg == 0                                           */
} xbfs_degree_info;
int xbfs_degree_stats(const xbfs_host_graph *g, xbfs_degree_info *out);

/* ---- device context ----------------------------------------------------- */
typedef struct xbfs_ctx xbfs_ctx;

/* Uploads the CSR to device `device` and allocates all traversal state.
 * The library owns every device buffer until xbfs_destroy. */
int xbfs_create(int device, const int64_t *beg_pos, const int32_t *adj,
                int64_t n, int64_t m, xbfs_ctx **out);
void xbfs_destroy(xbfs_ctx *ctx);
int64_t xbfs_num_vertices(const xbfs_ctx *ctx);
/* Persistent traversals are launched as plain kernels with a residency check
 * (a plain launch costs about 6 us, a cooperative one about 19 us); when not
 * every block was resident at once, the traversal is relaunched
 * cooperatively and the context launches cooperatively from then on.  Returns
 * how many traversals were relaunched (XBFS_LAUNCH=coop: every launch is
 * cooperative; XBFS_RESIDENCY_US bounds the check, default 20000).  No
 * reference counterpart. */
int64_t xbfs_launch_fallbacks(const xbfs_ctx *ctx);
int64_t xbfs_num_edges(const xbfs_ctx *ctx);

/* direction / queue modes (the reference's per-level tag in parentheses) */
#define XBFS_MODE_TD_SF 0 /* top-down scan-free: CAS claim + binned enqueue (tag 0) */
#define XBFS_MODE_TD_SS 1 /* top-down single-scan: plain stores, no enqueue (tag 1) */
#define XBFS_MODE_BU 2    /* bottom-up over a frontier bitmap (tag -1)              */

/* policies */
#define XBFS_POLICY_HEURISTIC 0 /* the reference's alpha / TD_adaptive rule */
#define XBFS_POLICY_LEARNED 1   /* MLP cost model; falls back to HEURISTIC  */
#define XBFS_POLICY_FORCE_TD_SF 2
#define XBFS_POLICY_FORCE_TD_SS 3
#define XBFS_POLICY_FORCE_BU 4
#define XBFS_POLICY_SCHEDULE 5  /* mode per level from opts.schedule       */

typedef struct {
    int32_t block_dim;      /* accepted for CLI compatibility (64..1024, power of
                               two); the persistent kernel runs 1024-thread blocks,
                               the host-driven phases 256-thread blocks           */
    int32_t grid_dim;       /* blocks; 0 = sized to the GPU (persistent: all
                               co-resident blocks, a smaller value caps it)        */
    float alpha;            /* frontier-edge ratio that switches to bottom-up       */
    float beta;             /* reserved (declared, unused by the reference)         */
    int32_t padding;        /* accepted for CLI compatibility; work is split per wave */
    int32_t policy;         /* XBFS_POLICY_*                                         */
    int32_t persistent;     /* 1 = level loop inside one cooperative kernel (default) */
    int32_t want_parent;    /* 1 = also produce the BFS tree (parent_out)            */
    int32_t bin_thread_max; /* thread bin: 0 < deg <= this; 0 = default              */
    int32_t bin_wave_max;   /* wave bin: <= this, block bin above; 0 = default       */
    const int8_t *schedule; /* XBFS_POLICY_SCHEDULE: mode for level L = schedule[min(L, len-1)] */
    int32_t schedule_len;
    int32_t time_levels;    /* 1 = per-level device timing in rec[].t_s              */
    int32_t bin_block_max;  /* block bin: <= this, grid bin above; 0 = default       */
    int32_t max_levels;     /* > 0: stop with XBFS_ERR_LIMIT at this many levels     */
    int32_t wait_limit_us;  /* > 0: bound of one grid-barrier wait in the persistent
                               kernel (default 4 s); past it XBFS_ERR_DEVICE          */
} xbfs_opts;

void xbfs_opts_default(xbfs_opts *o);

typedef struct {
    int32_t level;   /* L                                                     */
    int32_t mode;    /* XBFS_MODE_* executed                                   */
    int32_t tag;     /* reference tag: 0 scan-free, 1 single-scan, -1 bottom-up */
    int32_t policy;  /* policy that decided (LEARNED may fall back)           */
    int64_t n_f;     /* frontier vertices (level == L)                        */
    int64_t m_f;     /* frontier edges, the reference's count[L]              */
    int64_t n_u;     /* unvisited non-isolated vertices at level start        */
    int64_t m_u;     /* edges of unvisited vertices at level start            */
    float ratio;     /* (float)m_f / (float)E                                 */
    float pad;
    double t_s;      /* device seconds for this level (when time_levels)      */
    /* work a bottom-up level executed (0 for top-down levels): candidates
     * probed at their largest-degree neighbour, candidates whose rows were
     * then scanned, adjacency words those scans read.  A top-down level
     * expands exactly m_f edges. */
    int64_t bu_cand, bu_scan_v, bu_scan_e;
} xbfs_level_rec;

typedef struct {
    double t_bfs_s;   /* device time: level init .. final frontier check      */
    int32_t n_levels; /* levels executed                                       */
    int32_t max_level;/* largest level assigned                                */
    int64_t v_cc;     /* vertices reached                                      */
    int64_t m_cc;     /* sum of degrees of reached vertices (TEPS numerator)   */
    double td_s, bu_s, gen_s; /* per-direction device time (when time_levels)  */
} xbfs_stats;

/* One BFS from src.  level_out: host int32[n] or NULL (levels stay on the
 * device); parent_out: host int32[n] or NULL (needs opts->want_parent).
 * rec: host array of rec_cap records or NULL.  st: may be NULL. */
int xbfs_bfs(xbfs_ctx *ctx, int32_t src, const xbfs_opts *opts,
             int32_t *level_out, int32_t *parent_out,
             xbfs_level_rec *rec, int32_t rec_cap, xbfs_stats *st);

/* ---- allocation-free traversal -------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the traversal of xbfs_bfs on a CSR the caller already
 * holds in device memory, with every buffer beyond the inputs and outputs in
 * a library-owned device array of XBFS_NOALLOC_SCRATCH_BYTES per GPU (held
 * from the library's module load until the process ends).  Nothing is
 * allocated, host or device.
 *   d_beg_pos int64[n+1], d_adj int32[m]: device pointers, n <= XBFS_NOALLOC_MAX_N
 *     (larger n returns XBFS_ERR_LIMIT before touching the GPU);
 *   d_level int32[n] (device, required), d_parent int32[n] (device, or NULL);
 *   policy_weights: host pointer to XBFS_MLP_NPARAM floats for policy LEARNED,
 *     or NULL (LEARNED then decides as HEURISTIC);
 *   st (may be NULL): t_bfs_s is host wall time around the launch (no events:
 *     creating one allocates).
 * Rows are scanned in their stored order (no degree-ordered copy) and every
 * traversal runs until its frontier is empty (no create-time component).
 * Synchronous, on the null stream of `device`.  Calls that overlap on one GPU
 * share its array and race: serialise them. */
#define XBFS_NOALLOC_MAX_N (1 << 20)
#define XBFS_NOALLOC_SCRATCH_BYTES (60u << 20)
int xbfs_bfs_noalloc(int device, const int64_t *d_beg_pos, const int32_t *d_adj, int64_t n,
                     int64_t m, int32_t src, const xbfs_opts *opts, const float *policy_weights,
                     int32_t *d_level, int32_t *d_parent, xbfs_stats *st);

/* ---- batched traversal (no reference counterpart) ------------------------ */

/* Multi-source BFS: level[v] = hops from v to the nearest of the nsrc seeds
 * (duplicates allowed), -1 when no seed reaches v; parents form a BFS forest
 * rooted at the seeds (parent[seed] = seed).  Same level loop, policies,
 * records and stats as xbfs_bfs (v_cc counts every distinct seed). */
int xbfs_bfs_multi(xbfs_ctx *ctx, const int32_t *srcs, int32_t nsrc, const xbfs_opts *opts,
                   int32_t *level_out, int32_t *parent_out,
                   xbfs_level_rec *rec, int32_t rec_cap, xbfs_stats *st);

/* Packs k host graphs side by side into one CSR: vertex v of gs[i] becomes
 * vert_off[i] + v (vert_off: k+1 entries, vert_off[k] = total vertices). */
int xbfs_graph_concat(const xbfs_host_graph *gs, int32_t k, xbfs_host_graph *out,
                      int64_t *vert_off);

/* Multigraph batch: ctx holds k graphs packed by xbfs_graph_concat; one BFS
 * per graph from its local source local_src[i], all k in one traversal (one
 * persistent launch).  level_out: host int32[total n]; the slice
 * [vert_off[i], vert_off[i+1]) is graph i's level array in its own vertex
 * numbering.  Components are disjoint, so every slice equals graph i's
 * single-source BFS. */
int xbfs_bfs_multigraph(xbfs_ctx *ctx, const int64_t *vert_off, int32_t k,
                        const int32_t *local_src, const xbfs_opts *opts,
                        int32_t *level_out, xbfs_stats *st);

/* 64-source batch: up to 64 single-source BFSs at once, source s as bit s of
 * a 64-bit frontier / visited mask per vertex, expanded top-down (push) or
 * bottom-up (pull) per level by the union frontier's size, in one cooperative
 * launch over a copy of the graph relabelled by degree (built on the first
 * call, about 12 bytes per edge and 40 per vertex).  The graph is taken as
 * symmetric (bottom-up reads a row as its vertex's in-neighbours), as the
 * single-source bottom-up does.
 * levels_out: host int32[nsrc * n], row s = levels from srcs[s], or NULL
 * (rows stay on the device).  st (may be NULL): t_bfs_s = device time of the
 * whole batch, v_cc / m_cc = sums over the sources, n_levels = levels run,
 * td_s / bu_s = device time of its push / pull levels, gen_s = of the level
 * rows. */
int xbfs_bfs_batch64(xbfs_ctx *ctx, const int32_t *srcs, int32_t nsrc, const xbfs_opts *opts,
                     int32_t *levels_out, xbfs_stats *st);

/* Autotuner for the degree-bin thresholds and the grid size: coordinate
 * descent (two passes) over candidate values of bin_thread_max,
 * bin_wave_max, bin_block_max and grid_dim.  Objective: multi = 0, the sum
 * over the nsrc sample sources of single-source t_bfs; multi = 1, the t_bfs
 * of one multi-source traversal from all of them (the multigraph batch).
 * Each measurement is the median of reps runs.  opts is the starting point
 * on entry and the best setting found on return.  trials (may be NULL)
 * receives up to cap records; *ntrials (may be NULL) = records written. */
typedef struct {
    int32_t bin_thread_max, bin_wave_max, bin_block_max, grid_dim;
    double t_s;          /* summed traversal time of the sample sources */
} xbfs_tune_rec;
int xbfs_autotune(xbfs_ctx *ctx, const int32_t *srcs, int32_t nsrc, int32_t multi, int32_t reps,
                  xbfs_opts *opts, xbfs_tune_rec *trials, int32_t cap, int32_t *ntrials);

/* Device scan used by the frontier generators, exposed for testing:
 * out[i] = sum(in[0..i-1]) (int64), n up to 2^31-1.  Host pointers. */
int xbfs_exclusive_scan_i32(xbfs_ctx *ctx, const int32_t *in, int64_t *out, int64_t n);

/* Ascending list of vertices whose level is -1 and degree > 0, for a level
 * array given on the host (ballot + popcount + device scan).  Returns the
 * count in *count; list must hold n entries. */
int xbfs_unvisited_list(xbfs_ctx *ctx, const int32_t *level, int32_t *list,
                        int64_t *count);

/* ---- learned direction / queue policy ------------------------------------
 * No reference counterpart: a cost model that replaces the decision block
 * bfs_adaptiveFQ_wb_async.cuh:1047-1078 when loaded (policy LEARNED).
 * FP32 MLP 8 -> 32 -> 32 -> 3 (ReLU), XBFS_MLP_NPARAM floats packed as
 * W1[32][8], b1[32], W2[32][32], b2[32], W3[3][32], b3[3]; output j is the
 * predicted log2 microseconds of mode j at the level (XBFS_MODE_* order). */
#define XBFS_MLP_NPARAM 1443

/* The 8 features of a level (the same function the kernels evaluate). */
int xbfs_policy_features(int32_t level, int64_t n_f, int64_t m_f, int64_t n_u, int64_t m_u,
                         int32_t prev_mode, float *x8);
/* Loads weights into a context (copied; used by policy LEARNED). */
int xbfs_policy_load(xbfs_ctx *ctx, const float *weights, int32_t nweights);

/* HIP kernels of the model, host pointers in and out, on GPU `device`. */
int xbfs_mlp_forward(int device, const float *params, const float *x, int32_t n, float *out);
int xbfs_mlp_loss_grad(int device, const float *params, const float *x, const float *y,
                       const float *mask, int32_t n, float *grad, float *loss);
/* Full-batch Adam for `steps` steps; params is the initial value on entry
 * and the trained value on return; loss_curve (steps floats) may be NULL. */
int xbfs_mlp_train(int device, float *params, const float *x, const float *y, const float *mask,
                   int32_t n, int32_t steps, float lr, float *loss_curve);

/* Per-level probe: at every level the three modes are each run `reps` times
 * from a device snapshot of the traversal state and timed; then the level is
 * executed with the mode `opts->policy` picks (follow_best = 0) or with the
 * fastest measured mode (follow_best = 1, the per-level oracle-best
 * schedule).  Host-driven loop; parents are not produced. */
typedef struct {
    int32_t level;
    int32_t prev_mode;   /* -1 at level 0                                     */
    int32_t mode;        /* mode executed                                     */
    int32_t htag;        /* reference heuristic tag at this level              */
    int64_t n_f, m_f, n_u, m_u;
    double t_mode[3];    /* seconds per mode (median of reps), XBFS_MODE_* order.
                            TD single-scan includes the frontier sweep the next
                            level then needs. */
} xbfs_probe_rec;
int xbfs_bfs_probe(xbfs_ctx *ctx, int32_t src, const xbfs_opts *opts, int32_t follow_best,
                   int32_t reps, xbfs_probe_rec *rec, int32_t rec_cap, xbfs_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* XBFS_H */

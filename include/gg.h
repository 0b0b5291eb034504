/*
 * gg.h -- C ABI of the MI355X-native Gemini edge-processing engine.
 *
 * One GPU, one partition.  Every entry point returns an int status
 * (GG_OK = 0, negative on error) and leaves a message for gg_last_error()
 * (thread-local).  The reference aborts through assert()/exit(-1) instead
 * (core/graph.hpp and the toolkits' main()); the C++ facade in
 * geminigraph_amd/core turns a non-zero status back into an abort.
 *
 * Pointers named d_* are device pointers (hipMalloc'd, e.g. by gg_malloc);
 * pointers named h_* are host pointers.  No torch or HIP types appear here:
 * streams are opaque (void*), NULL meaning the library's own stream.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the reference tree).
 */
#ifndef GG_H
#define GG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GG_OK 0
#define GG_EINVAL (-1)   /* bad argument / shape */
#define GG_ENOMEM (-2)   /* device or host allocation failed */
#define GG_EHIP (-3)     /* HIP runtime error */
#define GG_EIO (-4)      /* edge-list file error */
#define GG_ENODEV (-5)   /* no usable GPU */
#define GG_EBUDGET (-6)  /* request exceeds the device-memory budget */

typedef struct gg_graph gg_graph;
typedef struct gg_subset gg_subset;

/* ------------------------------------------------------------ runtime */

/* Message of the last failing call on this thread. */
const char *gg_last_error(void);
/* Select and initialise a device (MPI_Instance ctor, core/mpi.hpp:53-57). */
int gg_init(int device);
/* Block until all work queued by the library has finished. */
int gg_synchronize(void);
/* Device memory helpers (alloc_vertex_array / dealloc_vertex_array,
 * core/graph.hpp:229-243; gather_vertex_array, core/graph.hpp:310-324).
 * gg_free waits for the work queued on the device (as hipFree does) and
 * keeps the buffer, per device and size, for the next gg_malloc of the same
 * size (up to the "alloc_cache_bytes" option in all; hipMalloc failures
 * release the kept buffers first).  Contents are not cleared. */
int gg_malloc(void **d_ptr, size_t bytes);
int gg_free(void *d_ptr);
int gg_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
int gg_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
int gg_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes);
/* Free / total device memory in bytes. */
int gg_mem_info(size_t *free_bytes, size_t *total_bytes);
/* Streaming device copy by a plain HIP kernel ("gg_copy_kernel").  No
 * counterpart in the reference: it calibrates HBM byte counters
 * (FETCH_SIZE / WRITE_SIZE) against a kernel of known traffic. */
int gg_device_copy(void *d_dst, const void *d_src, size_t bytes);

/* Tunables (names below).  The switch divisor reproduces the reference's
 * "sparse iff active_edges < edges / 20" (core/graph.hpp:1509) at 20; the
 * default, 200, is the crossover measured on MI355X (results do not depend
 * on it).
 *   "switch_divisor"   push/pull crossover divisor (default 200)
 *   "plan_switch_divisor" the divisor gg_cc and gg_sssp use when their
 *                      dense steps run on the bin plan (default 50: those
 *                      cost a few ps per arc, so they push longer); set
 *                      both divisors to 20 for the reference's switch
 *   "force_mode"       0 = auto, 1 = always push, 2 = always pull
 *   "push_small_max"   largest out-degree handled by in-wave load balancing
 *   "push_big_max"     largest out-degree handled by one wave per vertex
 *   "pull_light_max"   largest in-degree handled inside a 64-vertex group
 *   "pull_serial_max"  batched BFS pull without a bin plan: largest in-degree
 *                      scanned by a single lane
 *   "huge_grid_waves"  waves sharing the adjacency of the largest-degree bin
 *   "count_examined"   1 = count adjacency entries read (gg_run_stats)
 *   "bfs_alpha"        BFS: also push while the frontier's out-arcs times
 *                      alpha are at most the in-arcs of the unvisited
 *                      vertices (what a pull would scan); 0 = the divisor
 *                      rule alone (default 24)
 *   "pr_hot_sources"   relabelled graphs: H > 0 builds the destination-bin
 *                      plan (arcs from the H hottest sources gathered in
 *                      source order per bin, H capped at 2^(32 -
 *                      pr_bin_bits); the other arcs streamed by propagation
 *                      blocking; weighted graphs keep each arc's weight),
 *                      built on first use and kept with the graph (default
 *                      2^19).  PageRank, CC, SSSP and the batched BFS run
 *                      their dense steps on it; 0 = the plain pulls
 *   "pr_bin_bits"      that plan: 2^bits destinations per bin (8..14,
 *                      default 13)
 *   "pr_range_bits"    that plan: 2^bits sources per cold range (8..14,
 *                      default 14)
 *   "heavy_waves"      waves of the dense heavy-chunk grids (0 = one per chunk)
 *   "bfs_pull_waves"   waves of the BFS bottom-up grid (default 8192, one per
 *                      wave slot of MI355X; 0 = one per 64-vertex group);
 *                      the heavy-chunk step uses half as many
 *   "level_masks"      batched BFS/BC: 1 = keep per-level root masks
 *                      (default; depth rows are built once at the end and
 *                      BC selects backward arcs exactly), 0 = write depth
 *                      rows level by level
 *   "level_mask_bytes" size limit of that mask store (default 4 GiB)
 *   "batch_device_loop" batched BFS and SSSP: 1 = each iteration's mode is
 *                      decided on the device and the host issues iterations
 *                      ahead without waiting (default); 0 = the host reads
 *                      the frontier counts back after every iteration
 *   "bfs_ahead"        BFS: iterations the host issues beyond the last
 *                      decided one (1..8, default 2); more hides the host's
 *                      reaction time on short levels, each one past the end
 *                      costs launches that return at once
 *   "alloc_cache_bytes" bytes of freed gg_malloc buffers kept for reuse
 *                      (default 4 GiB; 0 = every gg_free calls hipFree);
 *                      lowering it below what is kept releases the kept
 *                      buffers at once (gg_mem_info counts them as used)
 *   "bfs_record_starts" BFS bottom-up step: 1 = take each in-list's start
 *                      from the vertex's 16-byte record, read anyway, when
 *                      the graph has fewer than 2^32 arcs (default); 0 =
 *                      read the offsets array, as graphs with more arcs do
 *   "relabel_hubs"     GG_BUILD_REORDER builds (read at build time): K > 0
 *                      keeps the K highest-out-degree vertices first, in
 *                      degree order, and the others in file order, those
 *                      without in-arcs last, so that BFS parents and depths
 *                      (indexed by file id) found by one wave share lines;
 *                      0 = every vertex in degree order (default; PageRank,
 *                      CC and the batched BC run faster on it)
 *   "serve_bfs_workers" gg_server_create: workers that take BFS queries only
 *                      (default 2, at most workers - 1; 0 = every worker takes
 *                      every kind): BFS batches take milliseconds, SSSP and
 *                      BC batches of the same graph tens to hundreds
 *   "batch_graphs"     batched BFS: 1 = replay each level (after the first)
 *                      from a hipGraph captured once per buffer parity;
 *                      0 = launch the level's kernels (default: measured
 *                      faster on MI355X, the host already runs a level
 *                      ahead of the device) */
int gg_set_option(const char *name, int64_t value);
int gg_get_option(const char *name, int64_t *value);

/* ------------------------------------------------------ synthetic input */

/* RMAT edge list written to device memory: num_edges packed records of
 * 8 bytes <u32 src,u32 dst> or 12 bytes <u32 src,u32 dst,f32 w>
 * (EdgeUnit<E>, core/type.hpp:27-41).  The reference ships no generator;
 * this one is counter-based so the CPU oracle reproduces it bit for bit. */
int gg_rmat_generate(void *d_edges, uint64_t num_edges, int scale, double a, double b,
                     double c, uint64_t seed, int permute, int weighted);

/* -------------------------------------------------------------- graphs */

typedef struct gg_graph_info {
  uint32_t vertices;
  uint64_t edges;       /* file edge count, used by the switch */
  uint64_t arcs;        /* adjacency entries per direction */
  int weighted;
  int symmetric;
  int transposed;
  size_t device_bytes;  /* device memory held by the graph */
  const uint32_t *d_out_degree;
  const uint32_t *d_in_degree;
  const uint64_t *d_out_index;   /* V+1 */
  const uint32_t *d_out_nbr;
  const float *d_out_w;          /* NULL when unweighted */
  const uint64_t *d_in_index;
  const uint32_t *d_in_nbr;
  const float *d_in_w;
  /* GG_BUILD_REORDER: the arrays above are in internal order;
   * d_perm[original] = internal id, d_iperm[internal] = original id
   * (both NULL when the graph keeps original ids). */
  int reordered;
  const uint32_t *d_perm;
  const uint32_t *d_iperm;
} gg_graph_info;

/* Build CSR (outgoing, keyed by src) and CSC (incoming, keyed by dst) from a
 * packed edge list, on the device with a radix sort.  Duplicates and
 * self-loops are kept.  undirected != 0 symmetrises first and aliases CSC to
 * CSR, with `edges` kept at the input count.
 * Replaces Graph<E>::load_directed (core/graph.hpp:772-1339) and
 * Graph<E>::load_undirected_from_directed (core/graph.hpp:350-757) minus the
 * file read.  edges_on_device != 0: `edges` is a device pointer.
 * Weighted records must carry weights >= 0 and not NaN (GG_EINVAL
 * otherwise, from build, load and gg_graph_open alike): the SSSP kernels
 * order f32 distances by their bit patterns, which holds for non-negative
 * floats only.  The reference's CAS-loop write_min (core/atomic.hpp:35-41)
 * takes any sign; its SSSP never relies on that. */
int gg_graph_build(const void *edges, int edges_on_device, uint64_t num_edges,
                   uint32_t vertices, int weighted, int undirected, gg_graph **out);
/* Same, reading a binary edge-list file: |E| = file size / record size
 * (core/graph.hpp:359-360, :781-782). */
int gg_graph_load(const char *path, uint32_t vertices, int weighted, int undirected,
                  gg_graph **out);
/* Build flags.  GG_BUILD_REORDER relabels vertices internally by descending
 * out-degree (ties by id) before the CSR/CSC build, so the hot sources of the
 * pull gathers share cache lines; every entry point still takes and returns
 * original vertex ids (roots in, per-vertex outputs out), translating at the
 * boundary.  The reference's loaders keep file ids (core/graph.hpp:772-1339);
 * its locality step is the vertex chunking of core/graph.hpp:824-871. */
#define GG_BUILD_REORDER 1u
/* GG_BUILD_HUB_FIRST keeps file ids but orders every adjacency list by
 * descending out-degree of the neighbour (ties as the relabeling orders
 * them), so a bottom-up step meets the likeliest frontier member first; the
 * reference's list order is arbitrary (core/graph.hpp:1064-1075).  Implied
 * by GG_BUILD_REORDER. */
#define GG_BUILD_HUB_FIRST 2u
int gg_graph_build_ex(const void *edges, int edges_on_device, uint64_t num_edges,
                      uint32_t vertices, int weighted, int undirected, unsigned flags,
                      gg_graph **out);
int gg_graph_load_ex(const char *path, uint32_t vertices, int weighted, int undirected,
                     unsigned flags, gg_graph **out);
int gg_graph_destroy(gg_graph *g);
/* Swap directions without moving data (Graph<E>::transpose, core/graph.hpp:760-769). */
int gg_graph_transpose(gg_graph *g);
int gg_graph_get_info(const gg_graph *g, gg_graph_info *info);

/* ------------------------------------------------------ vertex subsets */
/* Device bitmaps of u64 words, WORD_OFFSET(i) = i >> 6
 * (Bitmap / VertexSubset, core/bitmap.hpp:20-61;
 *  Graph<E>::alloc_vertex_subset, core/graph.hpp:327-329). */
int gg_subset_create(const gg_graph *g, gg_subset **out);
int gg_subset_destroy(gg_subset *s);
int gg_subset_clear(gg_subset *s);
int gg_subset_fill(gg_subset *s);
int gg_subset_set_bit(gg_subset *s, uint32_t v);
int gg_subset_get_bit(const gg_subset *s, uint32_t v, int *bit);
int gg_subset_count(const gg_subset *s, uint64_t *count);
uint64_t *gg_subset_device_words(gg_subset *s);

/* ------------------------------------------------------------ toolkits */

/* Per-call records kept in gg_run_stats (process_edges calls beyond this
 * are counted in `iterations` but not recorded). */
#define GG_STATS_RECORDS 256
typedef struct gg_run_stats {
  int iterations;          /* process_edges calls */
  int push_iterations;     /* of which sparse (push) */
  uint64_t edges_examined; /* adjacency entries read by edge kernels */
  uint64_t m_reach;        /* sum of out-degrees of reached vertices */
  uint64_t reached;        /* reached vertices (BFS/SSSP/BC) or components (CC) */
  double device_ms;        /* device time of the compute, by HIP events
                              (gg_bfs: host wall time of the call, until the
                              device published the traversal's end) */
  /* per process_edges call (the first GG_STATS_RECORDS calls): */
  uint32_t modes[GG_STATS_RECORDS];          /* 0 = push (sparse), 1 = pull (dense) */
  uint32_t it_active[GG_STATS_RECORDS];      /* active vertices passed in */
  uint32_t it_new[GG_STATS_RECORDS];         /* vertices activated by the call */
  uint64_t it_examined[GG_STATS_RECORDS];    /* adjacency entries read (option count_examined) */
  uint64_t it_unvisited[GG_STATS_RECORDS];   /* BFS: vertices not yet visited before the call */
} gg_run_stats;

/* PageRank, toolkits/pagerank.cpp:24-129: FP64, d = 0.85, non-normalised,
 * `iterations` pull sweeps.  d_curr_out[V] receives the final ranks;
 * h_deltas (iterations+1 entries, may be NULL) the printed deltas. */
int gg_pagerank(gg_graph *g, int iterations, double *d_curr_out, double *h_deltas,
                gg_run_stats *stats);
/* BFS, toolkits/bfs.cpp:22-109.  d_parent[V] (sentinel V; the root is its
 * own parent), d_depth[V] (0xFFFFFFFF when unreached, else the level) or
 * NULL for parents only, the reference's output, which writes 4 bytes per
 * vertex found instead of 8; h_frontier[cap] receives the
 * active-vertex count of every level (the reference's "active(i)" lines).
 * Returns the number of levels in stats->iterations. */
int gg_bfs(gg_graph *g, uint32_t root, uint32_t *d_parent, uint32_t *d_depth,
           uint32_t *h_frontier, int cap, gg_run_stats *stats);
/* gg_bfs without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device or pinned memory during a call (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and statistics as gg_bfs (depths and
 * level sizes identical; parents valid, as races may pick other ones), but
 * every traversal buffer lives in one fixed scratch array per GPU that the
 * library holds until the process ends (about 13 MiB), so the call makes no
 * hipMalloc, hipFree, pinned-memory or event calls; device_ms is host wall
 * time.  Limits: at most 2^20 vertices and cap <= 65536 (GG_EINVAL beyond).
 * gg_init (or any earlier call) must have initialised the library.  Calls
 * on one GPU share the array: calls that overlap on one GPU race, so
 * serialise them (one stream, or a lock). */
int gg_bfs_noalloc(gg_graph *g, uint32_t root, uint32_t *d_parent, uint32_t *d_depth,
                   uint32_t *h_frontier, int cap, gg_run_stats *stats);
/* Connected components by min-label propagation, toolkits/cc.cpp:22-109,
 * on a graph built with undirected != 0. */
int gg_cc(gg_graph *g, uint32_t *d_label, uint32_t *h_frontier, int cap,
          gg_run_stats *stats);
/* Single-source shortest paths, toolkits/sssp.cpp:22-105 (f32 weights,
 * sentinel 1e9). */
int gg_sssp(gg_graph *g, uint32_t root, float *d_dist, uint32_t *h_frontier, int cap,
            gg_run_stats *stats);
/* Betweenness dependencies from one root, toolkits/bc.cpp:24-196 (FP64).
 * d_dep[V] = dependencies, d_inv_paths[V] = 1 / num_paths. */
int gg_bc(gg_graph *g, uint32_t root, double *d_dep, double *d_inv_paths,
          uint32_t *h_frontier, int cap, gg_run_stats *stats);

/* ------------------------------------------------------- batched queries */
/* Up to 64 BFS roots in one traversal: every vertex carries a 64-bit visited
 * mask and a 64-bit frontier mask (bit i = root i), so one sweep of the
 * CSR/CSC serves all roots.  No counterpart in the reference, which runs one
 * root per compute() (toolkits/bfs.cpp:22-109); each root's depths equal what
 * gg_bfs returns for it.  d_depth is vertex-major: d_depth[v * 64 + i] = level
 * of v from roots[i], 0xFFFFFFFF when unreached (64 entries per vertex
 * whatever n_roots is).  stats->m_reach sums over the roots; the per-level
 * records (modes, it_active, it_new) describe the union of the roots'
 * frontiers.  The levels run on a device-driven loop (option
 * "batch_device_loop"): the host never waits for a level before issuing the
 * next. */
int gg_bfs_batch(gg_graph *g, const uint32_t *h_roots, int n_roots, uint32_t *d_depth,
                 gg_run_stats *stats);
/* Up to 64 SSSP roots in one traversal (toolkits/sssp.cpp semantics per
 * root): one wave per vertex, lane i relaxes root i, so each edge moves one
 * 256-byte row of distances.  d_dist[v * 64 + i] (f32, 1e9 when unreached)
 * equals what gg_sssp returns for roots[i], bit for bit.  Relaxations within
 * an iteration see each other, so the per-iteration records depend on
 * timing; only the distances are fixed.  Iterations run on the device-driven
 * loop like gg_bfs_batch's. */
int gg_sssp_batch(gg_graph *g, const uint32_t *h_roots, int n_roots, float *d_dist,
                  gg_run_stats *stats);
/* Up to 64 BC roots in one traversal (toolkits/bc.cpp semantics per root):
 * d_dep[v * 64 + i] and d_inv_paths[v * 64 + i] (FP64) as gg_bc returns them
 * for roots[i] (within 1e-9 relative; NaN where unreached). */
int gg_bc_batch(gg_graph *g, const uint32_t *h_roots, int n_roots, double *d_dep,
                double *d_inv_paths, gg_run_stats *stats);

/* ---------------------------------------------------------------- serving */
/* A query server over graphs resident in HBM (no counterpart in the
 * reference, whose Graph is not re-entrant: core/graph.hpp:132-141).
 * Queries (BFS depths, SSSP distances, BC dependencies from one root) are
 * admitted as they arrive; each worker, on its own HIP stream and its own
 * traversal buffers, takes the oldest pending query whose graph and kind
 * have no batch running (BFS queries only, for the serve_bfs_workers workers
 * reserved for them) and every pending query of the same graph and kind (up
 * to 64) and runs them as one batched traversal, so batches from different
 * workers overlap on the device while new queries queue.  Device memory is admitted against `budget_bytes`: the
 * resident graphs plus each worker's buffers for the graphs it serves. */
typedef struct gg_server gg_server;
#define GG_QUERY_BFS 0
#define GG_QUERY_SSSP 1
#define GG_QUERY_BC 2
typedef struct gg_server_info {
  size_t budget_bytes;
  size_t used_bytes;     /* graphs + worker buffers admitted so far */
  int graphs;
  int workers;
  uint64_t submitted;
  uint64_t completed;
  uint64_t batches;      /* batched traversals run */
  uint64_t max_batch;    /* largest batch formed */
} gg_server_info;
int gg_server_create(size_t budget_bytes, int workers, gg_server **out);
int gg_server_destroy(gg_server *s);
/* Makes a graph resident; GG_EBUDGET if it does not fit the budget.  The
 * server does not own the graph (destroy it after the server). */
int gg_server_add_graph(gg_server *s, gg_graph *g, int *graph_id);
/* Queues one query; want_result != 0 keeps its per-vertex result for
 * gg_server_wait (BFS: u32 depth, SSSP: f32 distance, BC: f64 dependency). */
int gg_server_submit(gg_server *s, int graph_id, int kind, uint32_t root, int want_result,
                     uint64_t *query_id);
/* Waits for a query; copies its result (if kept) into h_result
 * (result_bytes >= V * element size) and reports arrival-to-completion
 * latency.  The query is released afterwards. */
int gg_server_wait(gg_server *s, uint64_t query_id, void *h_result, size_t result_bytes,
                   double *latency_ms);
/* Open-loop replay: query i arrives at arrival_s[i] seconds after the call
 * starts; latency_ms[i] = completion - arrival; *wall_s = time to the last
 * completion.  Results are not kept. */
int gg_server_replay(gg_server *s, int n, const double *arrival_s, const int *graph_id,
                     const int *kind, const uint32_t *root, double *latency_ms, double *wall_s);
int gg_server_get_info(gg_server *s, gg_server_info *info);

/* ------------------------------------------------- vertex arrays, files */
/* dump_vertex_array / restore_vertex_array (core/graph.hpp:253-307) with one
 * partition: a file of elem_bytes x vertices bytes holding the whole array
 * (indexed by original vertex ids, as every entry point returns them).
 * Restore refuses a file of another size (GG_EIO; the reference asserts).
 * d_array holds `vertices` elements of elem_bytes bytes each. */
int gg_dump_vertex_array(const gg_graph *g, const void *d_array, size_t elem_bytes,
                         const char *path);
int gg_restore_vertex_array(const gg_graph *g, void *d_array, size_t elem_bytes,
                            const char *path);
/* gather_vertex_array (core/graph.hpp:310-324): with one partition the whole
 * array is local; gathering means the device-to-host copy into h_out. */
int gg_gather_vertex_array(const gg_graph *g, const void *d_array, size_t elem_bytes,
                           void *h_out);
/* A built graph to a binary file (both directions, weights, the relabeling)
 * and back without the edge sort: the reference rebuilds its CSR/CSC from
 * the edge list on every launch (core/graph.hpp:772-1339).  Heavy chunks and
 * per-vertex records are rebuilt on open with the options in force then. */
int gg_graph_save(const gg_graph *g, const char *path);
int gg_graph_open(const char *path, gg_graph **out);

/* ------------------------------------------------------------ autotuner */
/* Times candidate values of the engine's options on a sample of this graph's
 * own work and keeps the fastest, one option at a time (no counterpart in
 * the reference, whose switch is fixed at |E| / 20, core/graph.hpp:1509).
 * GG_TUNE_BFS: `samples` BFS roots (spread over the ids, with out-arcs);
 * tunes switch_divisor, bfs_alpha, push_big_max, push_small_max,
 * bfs_pull_waves.  GG_TUNE_PAGERANK: `samples` iterations (the plan build
 * is not timed); tunes pr_hot_sources, pr_bin_bits, pr_range_bits.
 * Each candidate is timed twice (best of two) with device events.  apply != 0
 * keeps the chosen values (options are process-wide), else the previous
 * values are restored.  No result depends on these options. */
#define GG_TUNE_BFS 0
#define GG_TUNE_PAGERANK 1
#define GG_TUNE_MAX_OPTIONS 8
typedef struct gg_tune_result {
  int kind;
  int n_options;                        /* entries used below */
  char names[GG_TUNE_MAX_OPTIONS][24];  /* option names, in tuning order */
  int64_t values[GG_TUNE_MAX_OPTIONS];  /* chosen values */
  double baseline_ms;                   /* the sample with the options in force before */
  double tuned_ms;                      /* the sample with the chosen values */
  int trials;                           /* configurations timed */
} gg_tune_result;
int gg_tune(gg_graph *g, int kind, int samples, int apply, gg_tune_result *out);

/* ----------------------------------------------------------- profiling */

/* Per-kernel device time accumulated with HIP events on the launching
 * stream while enabled.  kernel: "pull_light", "pull_heavy", "push",
 * "pr_pull", ...  Returns total milliseconds and launch count. */
int gg_prof_enable(int on);
int gg_prof_reset(void);
int gg_prof_read(const char *kernel, double *total_ms, uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif

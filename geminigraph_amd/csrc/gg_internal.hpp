// Internal definitions shared by the HIP sources of libgg.
//
// Layout in HBM (one graph, one partition):
//   per direction (out = CSR keyed by src, in = CSC keyed by dst):
//     index[V+1]  u64 offsets      nbr[arcs] u32 neighbour ids
//     w[arcs]     f32 weights (weighted graphs only, structure of arrays)
//     degree[V]   u32
//     heavy chunk list: vertices whose degree exceeds the pull group limit,
//     cut into chunks of at most kChunkEdges adjacency entries
//   frontiers: u64 bitmap words (V/64 + 1) plus three degree-binned queues.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/gg.h"

namespace gg {
struct Work;

constexpr int kWave = 64;
constexpr int kBlock = 256;                 // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kChunkEdges = 1024;      // heavy-vertex chunk (one wave, 16 x 64)
constexpr uint32_t kUnset = 0xFFFFFFFFu;
constexpr uint32_t kDegEscape = 0xFFFFu;   // a vertex record's capped 16-bit degree
constexpr uint32_t kScalarSlots = 256;      // spread slots of one scalar accumulator
constexpr size_t kScalarStride = kScalarSlots * 16;  // doubles per accumulator

// ------------------------------------------------------------------ errors
void set_error(int code, const std::string &msg);
int fail(int code, const std::string &msg);
#define GG_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e__ = (call);                                                      \
    if (e__ != hipSuccess)                                                        \
      return ::gg::fail(GG_EHIP, std::string(#call) + ": " + hipGetErrorString(e__)); \
  } while (0)
#define GG_TRY(expr)          \
  do {                        \
    int r__ = (expr);         \
    if (r__ != GG_OK) return r__; \
  } while (0)

// ----------------------------------------------------------------- weights
// An edge weight (f32 bits) the SSSP kernels can order: >= 0 (or -0.0) and
// not NaN.  Those kernels compare f32 distances as unsigned bit patterns,
// which orders non-negative floats only.
__host__ __device__ inline bool weight_ok(uint32_t bits) {
  return bits <= 0x7F800000u || bits == 0x80000000u;
}
// Weights of a host edge list of packed <u32, u32, f32> records.
inline bool host_weights_ok(const uint32_t *rec, uint64_t n) {
  for (uint64_t i = 0; i < n; i++)
    if (!weight_ok(rec[i * 3 + 2])) return false;
  return true;
}
constexpr const char *kBadWeight =
    "negative or NaN edge weight: weights must be >= 0 (SSSP orders f32 distances by their "
    "bit patterns)";

// ----------------------------------------------------------------- options
struct Options {
  // push/pull crossover: sparse iff active out-edges < |E| / switch_divisor.
  // The reference uses 20 (core/graph.hpp:1509); on gfx950 the pull sweep is
  // cheap relative to push atomics and the measured optimum for RMAT-24 BFS
  // lies between 200 and 2000 (results do not depend on the mode).
  int64_t switch_divisor = 200;
  int64_t plan_switch_divisor = 50;  // CC / SSSP crossover when their dense steps use the plan
  int64_t force_mode = 0;        // 0 auto, 1 push, 2 pull
  int64_t push_small_max = 64;   // in-wave load-balanced bin
  int64_t push_big_max = 1024;   // one wave per vertex; above: shared by the grid
  int64_t pull_light_max = 1024; // in-group; above: heavy chunks (fixed at build)
  int64_t pull_serial_max = 128; // BFS pull: lane-serial scan (4 probes in flight)
  int64_t huge_grid_waves = 4096;
  int64_t count_examined = 0;    // count adjacency entries read (roofline model)
  int64_t level_masks = 1;       // batched BFS/BC: per-level root masks (else depth rows)
  int64_t level_mask_bytes = int64_t(4) << 30;  // size limit of that store
  int64_t bfs_alpha = 24;        // BFS: push while frontier out-arcs * alpha <= unvisited in-arcs
  int64_t pr_hot_sources = 1 << 19;  // PageRank bin plan: hot sources (0 = plain pull)
  int64_t pr_bin_bits = 13;      // PageRank bin plan: 2^bits destinations per bin
  int64_t pr_range_bits = 14;    // PageRank bin plan: 2^bits sources per cold range
  int64_t heavy_waves = 8192;    // dense heavy-chunk grids: waves (0 = one per chunk)
  // BFS pull grid: waves (0 = one per group); the heavy-chunk step takes half.
  // 8192 = one wave per wave slot of the chip (256 CUs x 4 SIMDs x 8 waves)
  int64_t bfs_pull_waves = 8192;
  int64_t batch_graphs = 0;      // batched BFS levels replayed from captured hipGraphs
  int64_t batch_device_loop = 1; // batched BFS/SSSP iterations decided on the device
  int64_t bfs_ahead = 2;         // BFS: iterations the host issues ahead of the decisions
  int64_t bfs_record_starts = 1; // BFS pull: in-list starts from the vertex records (arcs < 2^32)
  int64_t alloc_cache_bytes = int64_t(4) << 30;  // gg_free keeps up to this much for gg_malloc
  // relabeling (fixed at build): the relabel_hubs highest-out-degree vertices
  // first, in degree order, then the others in file order, those without
  // in-arcs last (0: every vertex in degree order)
  int64_t relabel_hubs = 0;
  int64_t serve_bfs_workers = 2;  // servers: workers that take BFS queries only (read at create)
};
Options &options();
bool prof_on();  // per-kernel event timing (gg_prof_enable) is on

// ----------------------------------------------------------------- runtime
int ensure_init();
// The stream and traversal-context slot of the calling thread: slot 0 and
// the library stream by default; serving workers bind their own so that
// their traversals run concurrently on separate streams and buffers.
hipStream_t stream();
int context_slot();
void bind_thread(hipStream_t s, int slot);
constexpr int kMaxSlots = 33;
int device_alloc(void **p, size_t bytes);
void device_free(void *p);

// Profiling: event pairs around launches of named kernels.
struct ProfScope {
  const char *name;
  hipEvent_t start = nullptr, stop = nullptr;
  explicit ProfScope(const char *n);
  ~ProfScope();
};

// ------------------------------------------------------------------- graph
struct Dir {
  uint64_t *index = nullptr;  // V+1
  uint32_t *nbr = nullptr;    // arcs
  float *w = nullptr;         // arcs or null
  uint32_t *degree = nullptr; // V
  uint64_t *empty = nullptr;  // V/64 + 1 words: bit v set iff degree[v] == 0
  // V: per vertex {original id, degrees, first (smallest) neighbour here or
  // kUnset, the list's start index[v] (low 32 bits)}: what a BFS step needs
  // of a vertex it tests or finds, in one 16-byte read.  degrees = the degree
  // in the other direction << 16 | the degree here, each capped at
  // kDegEscape (a capped field means: read `degree` of that direction); the
  // start is whole while arcs < 2^32
  uint4 *vinfo = nullptr;
  uint32_t *first_orig = nullptr;  // V: original id of the first neighbour (V if none)
  // heavy chunks (degree > light limit at build time)
  uint32_t n_chunks = 0;
  uint32_t *chunk_v = nullptr;
  uint64_t *chunk_beg = nullptr;
  uint64_t *chunk_end = nullptr;
  uint32_t n_heavy = 0;
  uint32_t *heavy_v = nullptr;     // heavy vertex list
  uint32_t *heavy_first = nullptr; // n_heavy + 1: first chunk of each heavy vertex
  uint32_t *chunk_h = nullptr;     // heavy index of each chunk
  uint32_t light_max = 0;
  // 64-vertex groups up to the last vertex with arcs here: the groups after
  // it hold no vertex a bottom-up step can find
  uint32_t live_groups = 0;
};

// PageRank bin plan of a relabelled graph (built on first use, see
// prplan.hip): per destination bin the in-arcs from the H hottest sources,
// sorted by source; the other (cold) arcs stored for propagation blocking.
struct PrSeg;
struct PrPlan {
  uint64_t serial = 0;         // distinct for every plan built in the process
  int transposed = 0;          // orientation it was built for
  uint32_t H = 0;              // hot sources: internal id < H (at most 2^(32 - bin_bits))
  int bin_bits = 14;           // destination bins of 2^bin_bits vertices
  int range_bits = 14;         // cold source ranges of 2^range_bits sources
  uint32_t nbins = 0;
  uint64_t cold_base = 0;      // first cold out-arc (out.index[H])
  uint64_t n_cold = 0;         // cold arcs
  uint64_t n_hot = 0;          // hot arcs
  uint32_t *hot = nullptr;     // per bin, sorted: source << bin_bits | destination - bin start
  float *hot_w = nullptr;      // weighted graphs: each hot arc's weight
  float *cold_w = nullptr;     // weighted graphs: each cold arc's weight (cold order)
  uint32_t nranges = 0;
  uint16_t *src16 = nullptr;   // cold arcs by (range, bin, source): source - range start
  uint16_t *dst16 = nullptr;   //   and destination - bin start
  uint32_t *run_beg = nullptr; // nbins x nranges: run of range r in bin b (bin-major)
  uint32_t *run_end = nullptr;
  uint32_t nparts = 0;         // emit workgroups: (range, first arc, end arc)
  uint32_t *parts = nullptr;
  uint32_t nseg = 0;           // bin kernel workgroups
  PrSeg *segs = nullptr;
  uint32_t nsplit = 0;         // bins split over several segments
  uint32_t *split_bin = nullptr;
  std::vector<void *> allocations;
};

}  // namespace gg

struct gg_graph {
  int device = 0;
  uint32_t vertices = 0;
  uint64_t edges = 0;
  uint64_t arcs = 0;
  int weighted = 0;
  int symmetric = 0;
  int transposed = 0;
  size_t device_bytes = 0;
  gg::Dir out, in;   // in aliases out's arrays when symmetric
  // degree-sorted relabeling (GG_BUILD_REORDER): internal = perm[original]
  int reordered = 0;
  uint32_t *perm = nullptr;   // device, original -> internal
  uint32_t *iperm = nullptr;  // device, internal -> original
  std::vector<uint32_t> h_perm;
  uint32_t to_internal(uint32_t v) const { return reordered ? h_perm[v] : v; }
  std::vector<void *> allocations;
  // traversal contexts owned by the graph, one per context slot (lazily sized)
  gg::Work *works[gg::kMaxSlots] = {};
  // the current bin plan; every call that uses it holds a reference for its
  // duration, so a plan replaced under other threads' traversals (new plan
  // options) is freed only when the last of them is done with it
  std::shared_ptr<gg::PrPlan> pr_plan;
};

struct gg_subset {
  uint32_t size = 0;
  size_t words = 0;
  uint64_t *d = nullptr;
  const gg_graph *g = nullptr;  // bits are indexed by the graph's internal ids
};

namespace gg {

// Device counters written by the frontier compaction (one set per frontier).
struct Counters {
  uint32_t n[3];      // binned queue sizes
  uint32_t total;     // active vertices (popcount)
  uint64_t edges;     // sum of bin degrees of active vertices
};

struct Frontier {
  uint64_t *bits = nullptr;
  uint32_t *q[3] = {nullptr, nullptr, nullptr};
  float *snap_f[3] = {nullptr, nullptr, nullptr};   // message snapshots (u32/f32 bits)
  Counters *cnt = nullptr;                          // device
};

struct Work {
  uint32_t V = 0;
  size_t words = 0;
  Frontier f[2];
  uint64_t *visited = nullptr;
  uint64_t *aux_bits = nullptr;
  void *msg = nullptr;         // V x 8 bytes two-phase messages
  void *res_a = nullptr;       // V x 8 bytes results in internal order (reordered graphs)
  void *res_b = nullptr;       // V x 8 bytes
  double *hacc = nullptr;      // heavy accumulators (max of both directions)
  void *hpart = nullptr;       // per-chunk partials of a dense heavy pass (8 bytes each)
  uint64_t hpart_cap = 0;
  double *dacc = nullptr;      // per-iteration spread accumulators (kScalarStride each)
  double *dsum = nullptr;      // reduced per-iteration scalars
  int dacc_cap = 0;            // iterations the buffers hold
  unsigned long long *exam = nullptr;  // 64 slots x 16 (one 128-byte line each)
  // batched (64-root) traversals, allocated on first use
  uint64_t *ms_vis = nullptr;          // V visited masks
  uint64_t *ms_f[2] = {nullptr, nullptr};  // V frontier masks (current, next)
  uint32_t *ms_roots = nullptr;        // 64 roots
  void *ms_tmp = nullptr;              // V x 64 x 4 bytes (depth / distance rows)
  void *mb_a = nullptr, *mb_b = nullptr;  // V x 64 x 8 bytes each (BC rows, reordered outputs)
  uint64_t *ms_levels = nullptr;       // per-level frontier bitmaps (BC backward)
  size_t ms_levels_cap = 0;            // levels the store holds
  uint64_t *ms_lmask = nullptr;        // per-level root masks, V each (BC backward)
  size_t ms_lmask_cap = 0;
  size_t bytes = 0;                    // device bytes held (fixed part)
  void *bfs = nullptr;         // device-driven BFS loop state (bfs.hip)
  void *msd = nullptr;         // device-driven batched BFS loop state (algos.hip)
  double *pr_vals = nullptr;   // PageRank cold contributions, in cold-arc storage order
  double *pr_gsum = nullptr;   // PageRank sums of split bins (kept at 0 between iterations)
  uint64_t pr_vals_cap = 0;
  uint64_t pr_gsum_cap = 0;
  Counters *h_cnt = nullptr;   // pinned host copy
  Counters *cnt_slots = nullptr;    // spread partial counts of the count pass
  Counters *h_cnt_slots = nullptr;  // pinned host copy of those
  unsigned long long *h_exam = nullptr;
  std::vector<void *> allocations;
};

// Graph construction (build.hip): device memory owned by the graph, and the
// derived structures every graph needs once its CSR/CSC are in place.
int graph_alloc(gg_graph *g, void **p, size_t bytes);
int graph_finish(gg_graph *g);
int work_get(gg_graph *g, Work **w);
void work_free(Work *w);
void bfs_state_free(void *state);
// Looks up the device address of gg_bfs_noalloc's scratch array on `device`
// (current); called once per device at initialisation.
int bfs_noalloc_setup(int device);
void msd_state_free(void *state);
// PageRank plan (prplan.hip): *out = null when the graph gets the plain pull
// (not relabelled, option pr_hot_sources = 0, or 2^32 arcs or more).
using PlanRef = std::shared_ptr<const PrPlan>;
int pr_plan_get(gg_graph *g, PlanRef *out);
void pr_plan_free(PrPlan *p);
// One PageRank iteration on the plan: per destination bin the sums of its
// hot and cold arcs, then finish(v, sum) of the PageRank op for every vertex.
struct PrPull;
int pr_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const PrPull &op,
                      double *dacc);
// One dense step of the batched BFS on the plan: for every vertex the OR of
// its in-neighbours' 64-root frontier masks (fcur); the bits not in vis (and
// within full) go to fnext[v] and v's bit to the bitmap `out` (every word is
// written).  With a gate, the kernels return at once unless *gate is
// kGateDense when they run (a device-driven loop's mode word).  Arcs into
// vertices whose vis covers full read no source value; `settled`, if given,
// is a bitmap of those vertices (else vis is tested per bin).
constexpr uint32_t kGateDense = 1;
// One dense CC step on the plan: label[v] = min(label[v], min over v's
// in-neighbours u of ~nlab[u]) with nlab = ~label copied at the start of the
// step; lowered vertices go to `out` (every word is written).
int cc_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const uint32_t *nlab,
                      uint32_t *label, uint64_t *out);
// One dense SSSP step on the plan of a weighted graph: dist[v] is lowered to
// the minimum over v's in-arcs of fl(snap[u] + w) when that is below 1e9,
// where snap holds the distances' f32 bits at the step's start; lowered
// vertices go to `out` (every word is written).
int sssp_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const uint32_t *snap,
                        float *dist, uint64_t *out);
int ms_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const uint64_t *fcur,
                      const uint64_t *vis, uint64_t full, uint64_t *fnext, uint64_t *out,
                      const uint32_t *gate = nullptr, const uint64_t *settled = nullptr);
// Device bytes of one traversal context that runs batched queries of `kind`
// (GG_QUERY_*) on g, counting the first 64 stored BFS levels.
size_t work_footprint(const gg_graph *g, int kind);
// Batched queries (gg_bfs_batch, gg_sssp_batch, gg_bc_batch).  With a null
// output the results stay in the calling context's buffers, in internal
// vertex order (row of internal vertex n at rows[n * 64]), and *rows points
// at them (BC: the dependencies); they stay valid until the context's next
// batched query on g.
int bfs_batch_run(gg_graph *g, const uint32_t *h_roots, int n_roots, uint32_t *d_depth,
                  gg_run_stats *stats, const void **rows);
int sssp_batch_run(gg_graph *g, const uint32_t *h_roots, int n_roots, float *d_dist,
                   gg_run_stats *stats, const void **rows);
int bc_batch_run(gg_graph *g, const uint32_t *h_roots, int n_roots, double *d_dep,
                 double *d_inv_paths, gg_run_stats *stats, const void **rows);

// Host-side view of one direction, swapped by transpose.
inline const Dir &out_dir(const gg_graph *g) { return g->transposed ? g->in : g->out; }
inline const Dir &in_dir(const gg_graph *g) { return g->transposed ? g->out : g->in; }

}  // namespace gg

/*
 * CPU oracle for the Gemini edge-processing engine (test infrastructure).
 *
 * This library is a plain-C restatement of the reference's single-partition
 * engine and of its five toolkits.  It exists only to check the HIP product
 * path: tests, the smoke check and the bench's CPU-baseline leg load it, and
 * nothing in the product links it.
 *
 * Reference anchors (paths relative to the reference tree):
 *   core/type.hpp:27-41      EdgeUnit<E>: packed <u32 src, u32 dst[, f32 w]>
 *   core/bitmap.hpp:20-61    Bitmap / VertexSubset
 *   core/atomic.hpp:24-51    cas / write_min / write_add
 *   core/graph.hpp:772-1339  load_directed (degrees, CSR keyed by src, CSC keyed by dst)
 *   core/graph.hpp:350-757   load_undirected_from_directed (symmetric CSR, CSC alias)
 *   core/graph.hpp:760-769   transpose
 *   core/graph.hpp:1406-1469 process_vertices
 *   core/graph.hpp:1492-1867 process_edges (switch, sparse/dense, BSP phases)
 *   toolkits/{pagerank,bfs,cc,sssp,bc}.cpp  compute()
 *
 * Pinning status (each function is checked in tests/test_oracle_pins.py):
 *   og_rmat_*          determinism, bijective permutation, quadrant frequencies
 *   og_graph_build     brute-force lexsort/bincount CSR/CSC
 *   og_pagerank        closed forms (cycle, complete, star) + dense matrix recurrence
 *   og_bfs             scipy BFS distances, parent validity, level-set sizes
 *   og_cc              scipy weak components (min id per component)
 *   og_sssp            float32 Dijkstra (heap), bit-exact
 *   og_bc              networkx Brandes dependencies, root = reached-1, hand multigraph
 */
#ifndef GG_ORACLE_H
#define GG_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic RMAT input (the build's own generator; the reference has none) ---- */
typedef struct og_rmat_params {
  int scale;              /* V = 2^scale */
  uint64_t num_edges;     /* E */
  double a, b, c;         /* quadrant probabilities; d = 1-a-b-c */
  uint64_t seed;
  int permute;            /* apply the seeded vertex permutation to both endpoints */
  int weighted;           /* 12-byte records with f32 weight in (0,1] */
} og_rmat_params;

/* Writes E records (8 or 12 bytes each, packed) into out. */
void og_rmat_generate(const og_rmat_params *p, void *out);
uint32_t og_rmat_permute(uint32_t x, int scale, uint64_t seed);

/* ---- graph ---- */
typedef struct og_graph {
  uint32_t vertices;
  uint64_t edges;          /* file edge count (graph.hpp:359-360, :781-782) */
  uint64_t arcs;           /* stored adjacency entries per direction */
  int weighted;
  int symmetric;
  uint32_t *out_degree;
  uint32_t *in_degree;     /* aliases out_degree when symmetric (graph.hpp:494) */
  uint64_t *out_index;     /* V+1 */
  uint32_t *out_nbr;
  float *out_w;
  uint64_t *in_index;      /* V+1; aliases out_* when symmetric (graph.hpp:733-738) */
  uint32_t *in_nbr;
  float *in_w;
} og_graph;

/* edges: packed records; undirected != 0 follows load_undirected_from_directed. */
og_graph *og_graph_build(const void *edges, uint64_t num_edges, uint32_t vertices,
                         int weighted, int undirected);
void og_graph_free(og_graph *g);
void og_graph_transpose(og_graph *g);

/* Per-process_edges trace: mode (0 = sparse/push, 1 = dense/pull) and the
 * popcount of the active subset passed in. */
typedef struct og_trace {
  int32_t *modes;
  uint32_t *active;
  int cap;
  int n;
} og_trace;

/* Thread count used by the OpenMP regions (0 = OpenMP default). */
void og_set_threads(int n);
int og_get_threads(void);

/* toolkits/pagerank.cpp:26-129.  curr_out[V]; deltas_out[iterations] (may be NULL). */
void og_pagerank(og_graph *g, int iterations, double *curr_out, double *deltas_out,
                 og_trace *tr);
/* toolkits/bfs.cpp:22-109.  parent_out[V] (sentinel V), depth_out[V] (UINT32_MAX if
 * unreached), frontier_out[...] sizes per level (returns number of levels). */
int og_bfs(og_graph *g, uint32_t root, uint32_t *parent_out, uint32_t *depth_out,
           uint32_t *frontier_out, int frontier_cap, og_trace *tr);
/* toolkits/cc.cpp:22-109 (graph built with undirected=1).  Returns iterations. */
int og_cc(og_graph *g, uint32_t *label_out, uint32_t *frontier_out, int frontier_cap,
          og_trace *tr);
/* toolkits/sssp.cpp:22-105 (weighted graph).  Returns iterations. */
int og_sssp(og_graph *g, uint32_t root, float *dist_out, uint32_t *frontier_out,
            int frontier_cap, og_trace *tr);
/* Checker for SSSP results too large for og_sssp: counts the vertices at
 * which `dist` is not a fixed point of the relaxation the reference iterates
 * (sssp.cpp:44-71, f32 adds fl(dist[u] + w)).  With m(v) = the minimum of
 * fl(dist[u] + w) over v's in-arcs (CSC: index[V+1], src, w; +inf if none),
 * v violates when m(v) < dist[v] (an arc still relaxes), or when v != root,
 * dist[v] < 1e9 and m(v) != dist[v] (no arc gives the distance), or when
 * v == root and dist[v] != 0. */
uint64_t og_sssp_fixed_point_violations(const uint64_t *index, const uint32_t *src,
                                        const float *w, uint32_t vertices,
                                        const float *dist, uint32_t root);
/* toolkits/bc.cpp:24-196.  dep_out[V], inv_paths_out[V] (=1/sigma), levels sizes. */
int og_bc(og_graph *g, uint32_t root, double *dep_out, double *inv_paths_out,
          uint32_t *frontier_out, int frontier_cap, og_trace *tr);

/* Wall-clock seconds (core/time.hpp:22-26). */
double og_time(void);

#ifdef __cplusplus
}
#endif
#endif

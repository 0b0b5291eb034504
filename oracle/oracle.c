/*
 * CPU oracle: restatement of the reference's single-partition engine and its
 * five toolkits (test infrastructure only; see oracle.h for the anchors and
 * for how each function is pinned).
 *
 * The structure follows the reference: a Bitmap of 64-bit words, a
 * process_vertices reduction over set bits, and a process_edges that counts
 * active out-edges, switches on |E|/20, and runs the toolkit's signal and
 * slot callbacks in two BSP phases with an emitted-message buffer in
 * between.  OpenMP supplies the threads, as in the reference.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

/* ------------------------------------------------------------------ misc */

double og_time(void) { /* core/time.hpp:22-26 */
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + tv.tv_usec / 1e6;
}

static int og_threads = 0;
void og_set_threads(int n) {
  og_threads = n;
  if (n > 0) omp_set_num_threads(n);
}
int og_get_threads(void) { return og_threads > 0 ? og_threads : omp_get_max_threads(); }

static void *og_xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) abort();
  return p;
}
static void *og_xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) abort();
  return p;
}

/* --------------------------------------------------------------- atomics */
/* core/atomic.hpp:24-51 */

static inline int og_cas_u32(uint32_t *p, uint32_t o, uint32_t n) {
  return __atomic_compare_exchange_n(p, &o, n, 0, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
}
static inline int og_write_min_u32(uint32_t *p, uint32_t v) {
  int done = 0;
  uint32_t cur;
  do {
    cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  } while (cur > v && !(done = og_cas_u32(p, cur, v)));
  return done;
}
static inline int og_write_min_f32(float *p, float v) {
  int done = 0;
  uint32_t *pb = (uint32_t *)p;
  uint32_t cb, vb;
  float cur;
  memcpy(&vb, &v, 4);
  do {
    cb = __atomic_load_n(pb, __ATOMIC_SEQ_CST);
    memcpy(&cur, &cb, 4);
  } while (cur > v && !(done = og_cas_u32(pb, cb, vb)));
  return done;
}
static inline void og_write_add_f64(double *p, double v) {
  uint64_t *pb = (uint64_t *)p;
  uint64_t ob, nb;
  double o, n;
  do {
    ob = __atomic_load_n(pb, __ATOMIC_SEQ_CST);
    memcpy(&o, &ob, 8);
    n = o + v;
    memcpy(&nb, &n, 8);
  } while (!__atomic_compare_exchange_n(pb, &ob, nb, 0, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
}

/* ---------------------------------------------------------------- bitmap */
/* core/bitmap.hpp:20-61: WORD_OFFSET(i) = i >> 6, words = WORD_OFFSET(size)+1 */

typedef struct og_bitmap {
  uint32_t size;
  size_t words;
  uint64_t *data;
} og_bitmap;

static og_bitmap *bm_new(uint32_t size) {
  og_bitmap *b = (og_bitmap *)og_xmalloc(sizeof(og_bitmap));
  b->size = size;
  b->words = ((size_t)size >> 6) + 1;
  b->data = (uint64_t *)og_xcalloc(b->words, 8);
  return b;
}
static void bm_free(og_bitmap *b) {
  if (!b) return;
  free(b->data);
  free(b);
}
static void bm_clear(og_bitmap *b) {
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)b->words; i++) b->data[i] = 0;
}
static void bm_fill(og_bitmap *b) {
  size_t last = (size_t)b->size >> 6;
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)last; i++) b->data[i] = ~0ULL;
  b->data[last] = 0;
  for (size_t i = last << 6; i < b->size; i++) b->data[last] |= 1ULL << (i & 63);
}
static inline int bm_get(const og_bitmap *b, uint32_t i) {
  return (int)((b->data[i >> 6] >> (i & 63)) & 1ULL);
}
static inline void bm_set(og_bitmap *b, uint32_t i) {
  __atomic_fetch_or(&b->data[i >> 6], 1ULL << (i & 63), __ATOMIC_RELAXED);
}
static uint64_t bm_count(const og_bitmap *b) {
  uint64_t n = 0;
#pragma omp parallel for reduction(+ : n)
  for (int64_t i = 0; i < (int64_t)b->words; i++) n += (uint64_t)__builtin_popcountll(b->data[i]);
  return n;
}

/* ----------------------------------------------------------- graph build */

static void og_counting_fill(const unsigned char *edges, uint64_t ne, size_t rec, uint32_t V,
                             int key_is_dst, int both, uint64_t *index, uint32_t *nbr,
                             float *w, int weighted) {
  /* index[] must hold exclusive prefix sums; uses a cursor copy, stable in
   * input order (the reference's own order is nondeterministic,
   * graph.hpp:1064-1075, and no result depends on it). */
  uint64_t *cur = (uint64_t *)og_xmalloc(sizeof(uint64_t) * ((size_t)V + 1));
  memcpy(cur, index, sizeof(uint64_t) * ((size_t)V + 1));
  for (uint64_t e = 0; e < ne; e++) {
    uint32_t s, d;
    float wt = 0.f;
    memcpy(&s, edges + e * rec, 4);
    memcpy(&d, edges + e * rec + 4, 4);
    if (weighted) memcpy(&wt, edges + e * rec + 8, 4);
    for (int pass = 0; pass < (both ? 2 : 1); pass++) {
      uint32_t k = key_is_dst ? d : s, n = key_is_dst ? s : d;
      if (pass == 1) {
        uint32_t t = k;
        k = n;
        n = t;
      }
      uint64_t pos = cur[k]++;
      nbr[pos] = n;
      if (weighted) w[pos] = wt;
    }
  }
  free(cur);
}

og_graph *og_graph_build(const void *edges_v, uint64_t ne, uint32_t V, int weighted,
                         int undirected) {
  const unsigned char *edges = (const unsigned char *)edges_v;
  size_t rec = weighted ? 12 : 8;
  og_graph *g = (og_graph *)og_xcalloc(1, sizeof(og_graph));
  g->vertices = V;
  g->edges = ne;
  g->weighted = weighted;
  g->symmetric = undirected;
  g->out_degree = (uint32_t *)og_xcalloc((size_t)V + 1, 4);
  if (undirected) {
    /* graph.hpp:393-399: both endpoints count; CSR = CSC (graph.hpp:733-738) */
    for (uint64_t e = 0; e < ne; e++) {
      uint32_t s, d;
      memcpy(&s, edges + e * rec, 4);
      memcpy(&d, edges + e * rec + 4, 4);
      g->out_degree[s]++;
      g->out_degree[d]++;
    }
    g->in_degree = g->out_degree;
    g->arcs = 2 * ne;
    g->out_index = (uint64_t *)og_xmalloc(sizeof(uint64_t) * ((size_t)V + 1));
    g->out_index[0] = 0;
    for (uint32_t v = 0; v < V; v++) g->out_index[v + 1] = g->out_index[v] + g->out_degree[v];
    g->out_nbr = (uint32_t *)og_xmalloc(sizeof(uint32_t) * g->arcs);
    g->out_w = weighted ? (float *)og_xmalloc(sizeof(float) * g->arcs) : NULL;
    og_counting_fill(edges, ne, rec, V, 0, 1, g->out_index, g->out_nbr, g->out_w, weighted);
    g->in_index = g->out_index;
    g->in_nbr = g->out_nbr;
    g->in_w = g->out_w;
  } else {
    /* graph.hpp:815-820 out-degree; :966 in-degree; CSR keyed by src listing
     * dst, CSC keyed by dst listing src. */
    g->in_degree = (uint32_t *)og_xcalloc((size_t)V + 1, 4);
    for (uint64_t e = 0; e < ne; e++) {
      uint32_t s, d;
      memcpy(&s, edges + e * rec, 4);
      memcpy(&d, edges + e * rec + 4, 4);
      g->out_degree[s]++;
      g->in_degree[d]++;
    }
    g->arcs = ne;
    g->out_index = (uint64_t *)og_xmalloc(sizeof(uint64_t) * ((size_t)V + 1));
    g->in_index = (uint64_t *)og_xmalloc(sizeof(uint64_t) * ((size_t)V + 1));
    g->out_index[0] = g->in_index[0] = 0;
    for (uint32_t v = 0; v < V; v++) {
      g->out_index[v + 1] = g->out_index[v] + g->out_degree[v];
      g->in_index[v + 1] = g->in_index[v] + g->in_degree[v];
    }
    g->out_nbr = (uint32_t *)og_xmalloc(sizeof(uint32_t) * ne);
    g->in_nbr = (uint32_t *)og_xmalloc(sizeof(uint32_t) * ne);
    g->out_w = weighted ? (float *)og_xmalloc(sizeof(float) * ne) : NULL;
    g->in_w = weighted ? (float *)og_xmalloc(sizeof(float) * ne) : NULL;
    og_counting_fill(edges, ne, rec, V, 0, 0, g->out_index, g->out_nbr, g->out_w, weighted);
    og_counting_fill(edges, ne, rec, V, 1, 0, g->in_index, g->in_nbr, g->in_w, weighted);
  }
  return g;
}

void og_graph_free(og_graph *g) {
  if (!g) return;
  free(g->out_degree);
  free(g->out_index);
  free(g->out_nbr);
  free(g->out_w);
  if (!g->symmetric) {
    free(g->in_degree);
    free(g->in_index);
    free(g->in_nbr);
    free(g->in_w);
  }
  free(g);
}

void og_graph_transpose(og_graph *g) { /* graph.hpp:760-769 */
  uint32_t *td = g->out_degree;
  g->out_degree = g->in_degree;
  g->in_degree = td;
  uint64_t *ti = g->out_index;
  g->out_index = g->in_index;
  g->in_index = ti;
  uint32_t *tn = g->out_nbr;
  g->out_nbr = g->in_nbr;
  g->in_nbr = tn;
  float *tw = g->out_w;
  g->out_w = g->in_w;
  g->in_w = tw;
}

/* ---------------------------------------------------------------- engine */

typedef union og_val {
  double f64;
  float f32;
  uint32_t u32;
} og_val;

typedef struct og_msg { /* MsgUnit<M>, graph.hpp:61-65 */
  uint32_t vertex;
  og_val data;
} og_msg;

typedef struct og_adj { /* VertexAdjList<E>, type.hpp:62-68 */
  const uint32_t *nbr;
  const float *w;
  uint64_t begin, end;
} og_adj;

typedef struct og_emitter {
  og_msg *buf;
  size_t n, cap;
} og_emitter;

static void og_emit(og_emitter *em, uint32_t v, og_val m) { /* graph.hpp:1480-1490 */
  if (em->n == em->cap) {
    em->cap = em->cap ? em->cap * 2 : 256;
    em->buf = (og_msg *)realloc(em->buf, em->cap * sizeof(og_msg));
    if (!em->buf) abort();
  }
  em->buf[em->n].vertex = v;
  em->buf[em->n].data = m;
  em->n++;
}

typedef struct og_ops {
  void (*sparse_signal)(void *ctx, og_emitter *em, uint32_t v);
  uint64_t (*sparse_slot)(void *ctx, uint32_t v, og_val msg, og_adj adj);
  void (*dense_signal)(void *ctx, og_emitter *em, uint32_t v, og_adj adj);
  uint64_t (*dense_slot)(void *ctx, uint32_t v, og_val msg);
} og_ops;

typedef uint64_t (*og_vfn_u64)(void *ctx, uint32_t v);
typedef double (*og_vfn_f64)(void *ctx, uint32_t v);

/* graph.hpp:1406-1469: sum of fn(v) over set bits of active. */
static uint64_t og_process_vertices_u64(og_graph *g, og_vfn_u64 fn, void *ctx,
                                        const og_bitmap *active) {
  uint64_t red = 0;
  int64_t words = (int64_t)(((size_t)g->vertices + 63) >> 6);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : red)
  for (int64_t wi = 0; wi < words; wi++) {
    uint64_t word = active->data[wi];
    uint32_t v = (uint32_t)(wi << 6);
    while (word) {
      if (word & 1ULL) red += fn(ctx, v);
      v++;
      word >>= 1;
    }
  }
  return red;
}
static double og_process_vertices_f64(og_graph *g, og_vfn_f64 fn, void *ctx,
                                      const og_bitmap *active) {
  double red = 0;
  int64_t words = (int64_t)(((size_t)g->vertices + 63) >> 6);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : red)
  for (int64_t wi = 0; wi < words; wi++) {
    uint64_t word = active->data[wi];
    uint32_t v = (uint32_t)(wi << 6);
    while (word) {
      if (word & 1ULL) red += fn(ctx, v);
      v++;
      word >>= 1;
    }
  }
  return red;
}

static uint64_t og_outdeg_fn(void *ctx, uint32_t v) { return ((og_graph *)ctx)->out_degree[v]; }

static void og_trace_push(og_trace *tr, int mode, uint32_t active) {
  if (!tr || tr->n >= tr->cap) {
    if (tr) tr->n++;
    return;
  }
  tr->modes[tr->n] = mode;
  tr->active[tr->n] = active;
  tr->n++;
}

/* graph.hpp:1492-1867 with partitions = sockets = 1. */
static uint64_t og_process_edges(og_graph *g, const og_ops *ops, void *ctx,
                                 const og_bitmap *active, og_trace *tr) {
  uint64_t active_edges = og_process_vertices_u64(g, og_outdeg_fn, g, active);
  int sparse = active_edges < g->edges / 20; /* graph.hpp:1509 */
  if (tr) og_trace_push(tr, sparse ? 0 : 1, (uint32_t)bm_count(active));
  int nt = og_get_threads();
  og_emitter *ems = (og_emitter *)og_xcalloc((size_t)nt, sizeof(og_emitter));
  uint64_t red = 0;
  int64_t words = (int64_t)(((size_t)g->vertices + 63) >> 6);
  if (sparse) {
    /* signal phase (graph.hpp:1541-1556) */
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t wi = 0; wi < words; wi++) {
      og_emitter *em = &ems[omp_get_thread_num()];
      uint64_t word = active->data[wi];
      uint32_t v = (uint32_t)(wi << 6);
      while (word) {
        if (word & 1ULL) ops->sparse_signal(ctx, em, v);
        v++;
        word >>= 1;
      }
    }
    /* slot phase (graph.hpp:1600-1659): only vertices with out-edges */
    for (int t = 0; t < nt; t++) {
      og_emitter *em = &ems[t];
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : red) num_threads(nt)
      for (int64_t b = 0; b < (int64_t)em->n; b++) {
        uint32_t v = em->buf[b].vertex;
        if (g->out_index[v + 1] > g->out_index[v]) {
          og_adj adj = {g->out_nbr, g->out_w, g->out_index[v], g->out_index[v + 1]};
          red += ops->sparse_slot(ctx, v, em->buf[b].data, adj);
        }
      }
    }
  } else {
    /* signal phase over vertices with in-edges (graph.hpp:1755-1793) */
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t vv = 0; vv < (int64_t)g->vertices; vv++) {
      uint32_t v = (uint32_t)vv;
      if (g->in_index[v + 1] > g->in_index[v]) {
        og_adj adj = {g->in_nbr, g->in_w, g->in_index[v], g->in_index[v + 1]};
        ops->dense_signal(ctx, &ems[omp_get_thread_num()], v, adj);
      }
    }
    /* slot phase (graph.hpp:1801-1850) */
    for (int t = 0; t < nt; t++) {
      og_emitter *em = &ems[t];
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : red) num_threads(nt)
      for (int64_t b = 0; b < (int64_t)em->n; b++)
        red += ops->dense_slot(ctx, em->buf[b].vertex, em->buf[b].data);
    }
  }
  for (int t = 0; t < nt; t++) free(ems[t].buf);
  free(ems);
  return red;
}

/* ------------------------------------------------------------- pagerank */
/* toolkits/pagerank.cpp:24-129 */

typedef struct pr_ctx {
  og_graph *g;
  double *curr, *next;
} pr_ctx;

static const double PR_D = 0.85;

static double pr_init(void *c, uint32_t v) {
  pr_ctx *x = (pr_ctx *)c;
  x->curr[v] = 1.0;
  if (x->g->out_degree[v] > 0) x->curr[v] /= x->g->out_degree[v];
  return 1.0;
}
static void pr_sparse_signal(void *c, og_emitter *em, uint32_t v) {
  og_val m;
  m.f64 = ((pr_ctx *)c)->curr[v];
  og_emit(em, v, m);
}
static uint64_t pr_sparse_slot(void *c, uint32_t v, og_val msg, og_adj adj) {
  pr_ctx *x = (pr_ctx *)c;
  (void)v;
  for (uint64_t e = adj.begin; e < adj.end; e++) og_write_add_f64(&x->next[adj.nbr[e]], msg.f64);
  return 0;
}
static void pr_dense_signal(void *c, og_emitter *em, uint32_t v, og_adj adj) {
  pr_ctx *x = (pr_ctx *)c;
  double sum = 0;
  for (uint64_t e = adj.begin; e < adj.end; e++) sum += x->curr[adj.nbr[e]];
  og_val m;
  m.f64 = sum;
  og_emit(em, v, m);
}
static uint64_t pr_dense_slot(void *c, uint32_t v, og_val msg) {
  og_write_add_f64(&((pr_ctx *)c)->next[v], msg.f64);
  return 0;
}
static double pr_update_last(void *c, uint32_t v) {
  pr_ctx *x = (pr_ctx *)c;
  x->next[v] = 1 - PR_D + PR_D * x->next[v];
  return 0;
}
static double pr_update(void *c, uint32_t v) {
  pr_ctx *x = (pr_ctx *)c;
  x->next[v] = 1 - PR_D + PR_D * x->next[v];
  uint32_t od = x->g->out_degree[v];
  if (od > 0) {
    x->next[v] /= od;
    return fabs(x->next[v] - x->curr[v]) * od;
  }
  return fabs(x->next[v] - x->curr[v]);
}

void og_pagerank(og_graph *g, int iterations, double *curr_out, double *deltas_out,
                 og_trace *tr) {
  size_t V = g->vertices;
  pr_ctx x = {g, (double *)og_xmalloc(V * 8), (double *)og_xmalloc(V * 8)};
  og_bitmap *active = bm_new(g->vertices);
  bm_fill(active);
  double delta = og_process_vertices_f64(g, pr_init, &x, active);
  delta /= g->vertices;
  og_ops ops = {pr_sparse_signal, pr_sparse_slot, pr_dense_signal, pr_dense_slot};
  for (int it = 0; it < iterations; it++) {
    if (deltas_out) deltas_out[it] = delta;
#pragma omp parallel for
    for (int64_t v = 0; v < (int64_t)V; v++) x.next[v] = 0;
    og_process_edges(g, &ops, &x, active, tr);
    if (it == iterations - 1)
      delta = og_process_vertices_f64(g, pr_update_last, &x, active);
    else
      delta = og_process_vertices_f64(g, pr_update, &x, active);
    delta /= g->vertices;
    double *t = x.curr;
    x.curr = x.next;
    x.next = t;
  }
  if (deltas_out) deltas_out[iterations] = delta;
  memcpy(curr_out, x.curr, V * 8);
  free(x.curr);
  free(x.next);
  bm_free(active);
}

/* ------------------------------------------------------------------ bfs */
/* toolkits/bfs.cpp:22-109; depth is the level at which a vertex entered
 * active_out (not stored by the reference, derived here). */

typedef struct bfs_ctx {
  og_graph *g;
  uint32_t *parent, *depth;
  og_bitmap *visited, *active_in, *active_out;
  uint32_t level;
} bfs_ctx;

static void bfs_sparse_signal(void *c, og_emitter *em, uint32_t v) {
  (void)c;
  og_val m;
  m.u32 = v;
  og_emit(em, v, m);
}
static uint64_t bfs_sparse_slot(void *c, uint32_t src, og_val msg, og_adj adj) {
  bfs_ctx *x = (bfs_ctx *)c;
  (void)msg;
  uint32_t V = x->g->vertices;
  uint64_t act = 0;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t dst = adj.nbr[e];
    if (__atomic_load_n(&x->parent[dst], __ATOMIC_RELAXED) == V &&
        og_cas_u32(&x->parent[dst], V, src)) {
      bm_set(x->active_out, dst);
      act++;
    }
  }
  return act;
}
static void bfs_dense_signal(void *c, og_emitter *em, uint32_t dst, og_adj adj) {
  bfs_ctx *x = (bfs_ctx *)c;
  if (bm_get(x->visited, dst)) return;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t src = adj.nbr[e];
    if (bm_get(x->active_in, src)) {
      og_val m;
      m.u32 = src;
      og_emit(em, dst, m);
      break;
    }
  }
}
static uint64_t bfs_dense_slot(void *c, uint32_t dst, og_val msg) {
  bfs_ctx *x = (bfs_ctx *)c;
  if (og_cas_u32(&x->parent[dst], x->g->vertices, msg.u32)) {
    bm_set(x->active_out, dst);
    return 1;
  }
  return 0;
}
static uint64_t bfs_visit(void *c, uint32_t v) {
  bfs_ctx *x = (bfs_ctx *)c;
  bm_set(x->visited, v);
  x->depth[v] = x->level;
  return 1;
}

int og_bfs(og_graph *g, uint32_t root, uint32_t *parent_out, uint32_t *depth_out,
           uint32_t *frontier_out, int frontier_cap, og_trace *tr) {
  uint32_t V = g->vertices;
  bfs_ctx x;
  x.g = g;
  x.parent = parent_out;
  x.depth = depth_out;
  x.visited = bm_new(V);
  x.active_in = bm_new(V);
  x.active_out = bm_new(V);
  for (uint32_t v = 0; v < V; v++) {
    x.parent[v] = V;
    x.depth[v] = UINT32_MAX;
  }
  bm_set(x.visited, root);
  bm_set(x.active_in, root);
  x.parent[root] = root;
  x.depth[root] = 0;
  og_ops ops = {bfs_sparse_signal, bfs_sparse_slot, bfs_dense_signal, bfs_dense_slot};
  uint64_t active = 1;
  int it;
  for (it = 0; active > 0; it++) {
    if (it < frontier_cap) frontier_out[it] = (uint32_t)active;
    bm_clear(x.active_out);
    og_process_edges(g, &ops, &x, x.active_in, tr);
    x.level = (uint32_t)it + 1;
    active = og_process_vertices_u64(g, bfs_visit, &x, x.active_out);
    og_bitmap *t = x.active_in;
    x.active_in = x.active_out;
    x.active_out = t;
  }
  bm_free(x.visited);
  bm_free(x.active_in);
  bm_free(x.active_out);
  return it;
}

/* ------------------------------------------------------------------- cc */
/* toolkits/cc.cpp:22-109 */

typedef struct cc_ctx {
  og_graph *g;
  uint32_t *label;
  og_bitmap *active_out;
} cc_ctx;

static uint64_t cc_init(void *c, uint32_t v) {
  ((cc_ctx *)c)->label[v] = v;
  return 1;
}
static void cc_sparse_signal(void *c, og_emitter *em, uint32_t v) {
  og_val m;
  m.u32 = ((cc_ctx *)c)->label[v];
  og_emit(em, v, m);
}
static uint64_t cc_sparse_slot(void *c, uint32_t src, og_val msg, og_adj adj) {
  cc_ctx *x = (cc_ctx *)c;
  (void)src;
  uint64_t act = 0;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t dst = adj.nbr[e];
    if (msg.u32 < __atomic_load_n(&x->label[dst], __ATOMIC_RELAXED)) {
      og_write_min_u32(&x->label[dst], msg.u32);
      bm_set(x->active_out, dst);
      act++;
    }
  }
  return act;
}
static void cc_dense_signal(void *c, og_emitter *em, uint32_t dst, og_adj adj) {
  cc_ctx *x = (cc_ctx *)c;
  uint32_t msg = dst;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t l = x->label[adj.nbr[e]];
    if (l < msg) msg = l;
  }
  if (msg < dst) {
    og_val m;
    m.u32 = msg;
    og_emit(em, dst, m);
  }
}
static uint64_t cc_dense_slot(void *c, uint32_t dst, og_val msg) {
  cc_ctx *x = (cc_ctx *)c;
  if (msg.u32 < x->label[dst]) {
    og_write_min_u32(&x->label[dst], msg.u32);
    bm_set(x->active_out, dst);
    return 1;
  }
  return 0;
}

int og_cc(og_graph *g, uint32_t *label_out, uint32_t *frontier_out, int frontier_cap,
          og_trace *tr) {
  uint32_t V = g->vertices;
  cc_ctx x = {g, label_out, NULL};
  og_bitmap *active_in = bm_new(V), *active_out = bm_new(V);
  bm_fill(active_in);
  uint64_t active = og_process_vertices_u64(g, cc_init, &x, active_in);
  og_ops ops = {cc_sparse_signal, cc_sparse_slot, cc_dense_signal, cc_dense_slot};
  int it;
  for (it = 0; active > 0; it++) {
    if (it < frontier_cap) frontier_out[it] = (uint32_t)bm_count(active_in);
    bm_clear(active_out);
    x.active_out = active_out;
    og_process_edges(g, &ops, &x, active_in, tr);
    /* The reference loops on the reducer (cc.cpp:44, :39), which is nonzero
     * exactly when active_out is non-empty. */
    active = bm_count(active_out);
    og_bitmap *t = active_in;
    active_in = active_out;
    active_out = t;
  }
  bm_free(active_in);
  bm_free(active_out);
  return it;
}

/* ----------------------------------------------------------------- sssp */
/* toolkits/sssp.cpp:22-105 */

uint64_t og_sssp_fixed_point_violations(const uint64_t *index, const uint32_t *src,
                                        const float *w, uint32_t vertices,
                                        const float *dist, uint32_t root) {
  uint64_t bad = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad) num_threads(og_get_threads())
  for (int64_t v = 0; v < (int64_t)vertices; v++) {
    float m = INFINITY;
    for (uint64_t e = index[v]; e < index[v + 1]; e++) {
      const float c = dist[src[e]] + w[e]; /* sssp.cpp:50, :65 */
      if (c < m) m = c;
    }
    const float d = dist[v];
    if (v == (int64_t)root)
      bad += d != 0.0f || m < d;
    else
      bad += m < d || (d < 1e9f && m != d);
  }
  return bad;
}

typedef struct sssp_ctx {
  og_graph *g;
  float *dist;
  og_bitmap *active_out;
} sssp_ctx;

static void sssp_sparse_signal(void *c, og_emitter *em, uint32_t v) {
  og_val m;
  m.f32 = ((sssp_ctx *)c)->dist[v];
  og_emit(em, v, m);
}
static uint64_t sssp_sparse_slot(void *c, uint32_t src, og_val msg, og_adj adj) {
  sssp_ctx *x = (sssp_ctx *)c;
  (void)src;
  uint64_t act = 0;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t dst = adj.nbr[e];
    float relax = msg.f32 + adj.w[e];
    uint32_t cb = __atomic_load_n((uint32_t *)&x->dist[dst], __ATOMIC_RELAXED);
    float cur;
    memcpy(&cur, &cb, 4);
    if (relax < cur) {
      if (og_write_min_f32(&x->dist[dst], relax)) {
        bm_set(x->active_out, dst);
        act++;
      }
    }
  }
  return act;
}
static void sssp_dense_signal(void *c, og_emitter *em, uint32_t dst, og_adj adj) {
  sssp_ctx *x = (sssp_ctx *)c;
  float msg = 1e9f;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    float relax = x->dist[adj.nbr[e]] + adj.w[e];
    if (relax < msg) msg = relax;
  }
  if (msg < 1e9f) {
    og_val m;
    m.f32 = msg;
    og_emit(em, dst, m);
  }
}
static uint64_t sssp_dense_slot(void *c, uint32_t dst, og_val msg) {
  sssp_ctx *x = (sssp_ctx *)c;
  if (msg.f32 < x->dist[dst]) {
    og_write_min_f32(&x->dist[dst], msg.f32);
    bm_set(x->active_out, dst);
    return 1;
  }
  return 0;
}

int og_sssp(og_graph *g, uint32_t root, float *dist_out, uint32_t *frontier_out,
            int frontier_cap, og_trace *tr) {
  uint32_t V = g->vertices;
  sssp_ctx x = {g, dist_out, NULL};
  og_bitmap *active_in = bm_new(V), *active_out = bm_new(V);
  bm_set(active_in, root);
  for (uint32_t v = 0; v < V; v++) x.dist[v] = 1e9f;
  x.dist[root] = 0.f;
  og_ops ops = {sssp_sparse_signal, sssp_sparse_slot, sssp_dense_signal, sssp_dense_slot};
  uint64_t active = 1;
  int it;
  for (it = 0; active > 0; it++) {
    if (it < frontier_cap) frontier_out[it] = (uint32_t)bm_count(active_in);
    bm_clear(active_out);
    x.active_out = active_out;
    og_process_edges(g, &ops, &x, active_in, tr);
    active = bm_count(active_out);
    og_bitmap *t = active_in;
    active_in = active_out;
    active_out = t;
  }
  bm_free(active_in);
  bm_free(active_out);
  return it;
}

/* ------------------------------------------------------------------- bc */
/* toolkits/bc.cpp:24-196 (compute(), COMPACT 0) */

typedef struct bc_ctx {
  og_graph *g;
  double *num_paths, *dep;
  og_bitmap *visited, *active_in, *active_out;
} bc_ctx;

static void bcf_sparse_signal(void *c, og_emitter *em, uint32_t v) {
  og_val m;
  m.f64 = ((bc_ctx *)c)->num_paths[v];
  og_emit(em, v, m);
}
static uint64_t bcf_sparse_slot(void *c, uint32_t src, og_val msg, og_adj adj) {
  bc_ctx *x = (bc_ctx *)c;
  (void)src;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t dst = adj.nbr[e];
    if (!bm_get(x->visited, dst)) {
      uint64_t bits = __atomic_load_n((uint64_t *)&x->num_paths[dst], __ATOMIC_RELAXED);
      double cur;
      memcpy(&cur, &bits, 8);
      if (cur == 0) bm_set(x->active_out, dst);
      og_write_add_f64(&x->num_paths[dst], msg.f64);
    }
  }
  return 0;
}
static void bcf_dense_signal(void *c, og_emitter *em, uint32_t dst, og_adj adj) {
  bc_ctx *x = (bc_ctx *)c;
  if (bm_get(x->visited, dst)) return;
  double sum = 0;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t src = adj.nbr[e];
    if (bm_get(x->active_in, src)) sum += x->num_paths[src];
  }
  if (sum > 0) {
    og_val m;
    m.f64 = sum;
    og_emit(em, dst, m);
  }
}
static uint64_t bcf_dense_slot(void *c, uint32_t dst, og_val msg) {
  bc_ctx *x = (bc_ctx *)c;
  if (!bm_get(x->visited, dst)) {
    bm_set(x->active_out, dst);
    og_write_add_f64(&x->num_paths[dst], msg.f64);
  }
  return 0;
}
static uint64_t bc_visit(void *c, uint32_t v) {
  bm_set(((bc_ctx *)c)->visited, v);
  return 1;
}
static void bcb_sparse_signal(void *c, og_emitter *em, uint32_t v) {
  og_val m;
  m.f64 = ((bc_ctx *)c)->dep[v];
  og_emit(em, v, m);
}
static uint64_t bcb_sparse_slot(void *c, uint32_t src, og_val msg, og_adj adj) {
  bc_ctx *x = (bc_ctx *)c;
  (void)src;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t dst = adj.nbr[e];
    if (!bm_get(x->visited, dst)) og_write_add_f64(&x->dep[dst], msg.f64);
  }
  return 0;
}
static void bcb_dense_signal(void *c, og_emitter *em, uint32_t dst, og_adj adj) {
  bc_ctx *x = (bc_ctx *)c;
  if (bm_get(x->visited, dst)) return;
  double sum = 0;
  for (uint64_t e = adj.begin; e < adj.end; e++) {
    uint32_t src = adj.nbr[e];
    if (bm_get(x->active_in, src)) sum += x->dep[src]; /* active_in = levels.back() */
  }
  og_val m;
  m.f64 = sum;
  og_emit(em, dst, m);
}
static uint64_t bcb_dense_slot(void *c, uint32_t dst, og_val msg) {
  bc_ctx *x = (bc_ctx *)c;
  if (!bm_get(x->visited, dst)) og_write_add_f64(&x->dep[dst], msg.f64);
  return 0;
}
static uint64_t bc_inv(void *c, uint32_t v) {
  bc_ctx *x = (bc_ctx *)c;
  x->num_paths[v] = 1 / x->num_paths[v];
  x->dep[v] = 0;
  return 1;
}
static uint64_t bc_back_visit(void *c, uint32_t v) {
  bc_ctx *x = (bc_ctx *)c;
  bm_set(x->visited, v);
  x->dep[v] += x->num_paths[v];
  return 1;
}
static uint64_t bc_final(void *c, uint32_t v) {
  bc_ctx *x = (bc_ctx *)c;
  x->dep[v] = (x->dep[v] - x->num_paths[v]) / x->num_paths[v];
  return 1;
}

int og_bc(og_graph *g, uint32_t root, double *dep_out, double *inv_paths_out,
          uint32_t *frontier_out, int frontier_cap, og_trace *tr) {
  uint32_t V = g->vertices;
  bc_ctx x;
  x.g = g;
  x.num_paths = inv_paths_out;
  x.dep = dep_out;
  og_bitmap *active_all = bm_new(V);
  bm_fill(active_all);
  x.visited = bm_new(V);
  int nlev = 0, caplev = 64;
  og_bitmap **levels = (og_bitmap **)og_xmalloc(sizeof(og_bitmap *) * (size_t)caplev);
  og_bitmap *active_in = bm_new(V);
  bm_set(x.visited, root);
  bm_set(active_in, root);
  levels[nlev++] = active_in;
  for (uint32_t v = 0; v < V; v++) x.num_paths[v] = 0.0;
  x.num_paths[root] = 1.0;
  og_ops fops = {bcf_sparse_signal, bcf_sparse_slot, bcf_dense_signal, bcf_dense_slot};
  uint64_t active = 1;
  int it;
  for (it = 0; active > 0; it++) {
    if (it < frontier_cap) frontier_out[it] = (uint32_t)active;
    og_bitmap *active_out = bm_new(V);
    x.active_in = active_in;
    x.active_out = active_out;
    og_process_edges(g, &fops, &x, active_in, tr);
    active = og_process_vertices_u64(g, bc_visit, &x, active_out);
    if (nlev == caplev) {
      caplev *= 2;
      levels = (og_bitmap **)realloc(levels, sizeof(og_bitmap *) * (size_t)caplev);
    }
    levels[nlev++] = active_out;
    active_in = active_out;
  }
  og_process_vertices_u64(g, bc_inv, &x, active_all);
  bm_clear(x.visited);
  og_process_vertices_u64(g, bc_back_visit, &x, levels[nlev - 1]);
  og_graph_transpose(g);
  og_ops bops = {bcb_sparse_signal, bcb_sparse_slot, bcb_dense_signal, bcb_dense_slot};
  while (nlev > 1) {
    x.active_in = levels[nlev - 1];
    og_process_edges(g, &bops, &x, levels[nlev - 1], tr);
    bm_free(levels[nlev - 1]);
    nlev--;
    og_process_vertices_u64(g, bc_back_visit, &x, levels[nlev - 1]);
  }
  og_process_vertices_u64(g, bc_final, &x, active_all);
  og_graph_transpose(g);
  bm_free(levels[0]);
  free(levels);
  bm_free(x.visited);
  bm_free(active_all);
  return it;
}

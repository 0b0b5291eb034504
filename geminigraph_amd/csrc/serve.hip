// Query server: resident graphs, HBM-budget admission and continuous
// batching over the batched traversals (gg_bfs_batch / gg_sssp_batch /
// gg_bc_batch).
//
// Workers are threads, each bound to its own HIP stream and traversal-context
// slot, so the batches they run overlap on the device.  A free worker takes
// the oldest pending query whose (graph, kind) has no batch running (of kind
// BFS only, for the serve_bfs_workers workers reserved for it), and
// every other pending query of the same graph and kind (up to 64) --
// queries that arrived while the previous batch of their kind ran join the
// next one.  One batch per (graph, kind) at a time keeps batches full (a
// batched traversal costs nearly as much for a few roots as for 64) and
// stops one graph and kind from taking every worker.  Device memory is charged against the budget when a
// graph is added and the first time a worker serves a graph (its traversal
// context and 64-root result rows); a worker that would exceed the budget
// first drops its own buffers for other graphs, and fails the batch with
// GG_EBUDGET if that is not enough.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "gg_internal.hpp"

namespace gg {

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Column `col` of 64-root result rows kept in internal vertex order, in
// original vertex order: out[v] = rows[perm[v] * 64 + col].
template <class T>
__global__ void extract_column_kernel(const T *rows, const uint32_t *perm, uint32_t V, int col,
                                      T *out) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s) {
    uint32_t n = perm ? perm[v] : v;
    out[v] = rows[(size_t)n * 64 + col];
  }
}

struct SQuery {
  uint64_t id = 0;
  int graph = 0, kind = 0;
  uint32_t root = 0;
  bool want = false;
  double t_arrive = 0, t_done = 0;
  int status = GG_OK;
  std::string err;
  bool done = false;
  std::vector<unsigned char> result;
};

struct WorkerBuf {
  void *col = nullptr;  // one extracted result column (8 x V bytes)
  size_t charged = 0;   // budget charged for this worker on this graph
};

}  // namespace gg

struct gg_server {
  size_t budget = 0, used = 0;
  int device = 0;
  std::vector<gg_graph *> graphs;
  std::vector<std::thread> threads;
  std::vector<hipStream_t> streams;
  // [worker][graph]; a deque so add_graph's growth keeps references valid
  std::vector<std::deque<gg::WorkerBuf>> bufs;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<gg::SQuery *> pending;
  std::vector<unsigned char> running;  // [graph * 3 + kind]: a batch is running
  int bfs_workers = 0;  // workers 0 .. bfs_workers-1 take BFS queries only
  std::map<uint64_t, gg::SQuery *> queries;
  uint64_t next_id = 1, submitted = 0, completed = 0, batches = 0, max_batch = 0;
  bool stop = false;
};

namespace gg {

static size_t elem_bytes(int kind) { return kind == GG_QUERY_BC ? 8 : 4; }

// Device bytes a worker needs to serve `kind` on g: its traversal context
// (which keeps the 64-root result rows) plus one result column.
static size_t serve_bytes(const gg_graph *g, int kind) {
  return work_footprint(g, kind) + 8 * (size_t)g->vertices;
}

static void release_buf(WorkerBuf &b) {
  device_free(b.col);
  b = WorkerBuf{};
}

static void release_worker_graph(gg_server *s, int k, size_t gi) {
  WorkerBuf &o = s->bufs[k][gi];
  s->used -= o.charged;
  release_buf(o);
  gg_graph *g = s->graphs[gi];
  if (g->works[k + 1]) {
    work_free(g->works[k + 1]);
    g->works[k + 1] = nullptr;
  }
}

// Charges worker k's buffers for `kind` on graph gi against the budget,
// first dropping the worker's buffers for other graphs if it has to (caller
// holds mu; the allocation itself happens on the worker, outside the lock).
static int admit_locked(gg_server *s, int k, int gi, int kind) {
  WorkerBuf &b = s->bufs[k][gi];
  size_t need = std::max(b.charged, serve_bytes(s->graphs[gi], kind));
  size_t extra = need - b.charged;
  if (!extra) return GG_OK;
  for (size_t j = 0; j < s->bufs[k].size() && s->used + extra > s->budget; j++)
    if ((int)j != gi && s->bufs[k][j].charged) release_worker_graph(s, k, j);
  if (s->used + extra > s->budget)
    return fail(GG_EBUDGET, "serving: query working set exceeds the device-memory budget");
  s->used += extra;
  b.charged = need;
  return GG_OK;
}

static int run_batch(gg_graph *g, WorkerBuf &b, int kind, std::vector<SQuery *> &batch) {
  if (!b.col) GG_TRY(device_alloc(&b.col, 8 * (size_t)g->vertices));
  std::vector<uint32_t> roots(batch.size());
  for (size_t i = 0; i < batch.size(); i++) roots[i] = batch[i]->root;
  int n = (int)roots.size();
  const void *rows = nullptr;
  if (kind == GG_QUERY_BFS)
    GG_TRY(bfs_batch_run(g, roots.data(), n, nullptr, nullptr, &rows));
  else if (kind == GG_QUERY_SSSP)
    GG_TRY(sssp_batch_run(g, roots.data(), n, nullptr, nullptr, &rows));
  else
    GG_TRY(bc_batch_run(g, roots.data(), n, nullptr, nullptr, nullptr, &rows));
  const uint32_t V = g->vertices;
  const uint32_t *perm = g->reordered ? g->perm : nullptr;
  const size_t eb = elem_bytes(kind);
  const dim3 grid(std::min<uint32_t>((V + 255) / 256, 65536));
  for (size_t i = 0; i < batch.size(); i++) {
    if (!batch[i]->want) continue;
    if (eb == 8)
      hipLaunchKernelGGL(extract_column_kernel<double>, grid, dim3(256), 0, stream(),
                         (const double *)rows, perm, V, (int)i, (double *)b.col);
    else
      hipLaunchKernelGGL(extract_column_kernel<uint32_t>, grid, dim3(256), 0, stream(),
                         (const uint32_t *)rows, perm, V, (int)i, (uint32_t *)b.col);
    batch[i]->result.resize(eb * V);
    GG_HIP(hipMemcpyAsync(batch[i]->result.data(), b.col, eb * V, hipMemcpyDeviceToHost,
                          stream()));
    GG_HIP(hipStreamSynchronize(stream()));
  }
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

// The oldest pending query whose graph and kind have no batch running, of
// kind BFS only for a BFS worker (caller holds mu); pending.end() if there is
// none.
static std::deque<SQuery *>::iterator runnable_locked(gg_server *s, bool bfs_only) {
  for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
    if (!s->running[(size_t)(*it)->graph * 3 + (*it)->kind] &&
        (!bfs_only || (*it)->kind == GG_QUERY_BFS))
      return it;
  return s->pending.end();
}

static void worker_main(gg_server *s, int k) {
  hipSetDevice(s->device);
  bind_thread(s->streams[k], k + 1);
  std::unique_lock<std::mutex> lk(s->mu);
  const bool bfs_only = k < s->bfs_workers;
  while (true) {
    std::deque<SQuery *>::iterator first;
    s->cv_work.wait(lk, [&] {
      if (s->stop && s->pending.empty()) return true;
      first = runnable_locked(s, bfs_only);
      return first != s->pending.end();
    });
    if (s->stop && s->pending.empty()) break;  // stop requested and drained
    SQuery *head = *first;
    int gi = head->graph, kind = head->kind;
    s->running[(size_t)gi * 3 + kind] = 1;
    std::vector<SQuery *> batch;
    for (auto it = first; it != s->pending.end() && batch.size() < 64;) {
      if ((*it)->graph == gi && (*it)->kind == kind) {
        batch.push_back(*it);
        it = s->pending.erase(it);
      } else {
        ++it;
      }
    }
    int r = admit_locked(s, k, gi, kind);
    std::string err = r == GG_OK ? "" : gg_last_error();
    gg_graph *g = s->graphs[gi];
    WorkerBuf &buf = s->bufs[k][gi];
    lk.unlock();
    if (r == GG_OK) {
      r = run_batch(g, buf, kind, batch);
      if (r != GG_OK) err = gg_last_error();
    }
    double t = now_s();
    lk.lock();
    s->running[(size_t)gi * 3 + kind] = 0;
    s->cv_work.notify_all();  // queries of this graph and kind may be waiting
    for (SQuery *q : batch) {
      q->t_done = t;
      q->status = r;
      q->err = err;
      q->done = true;
    }
    s->batches++;
    s->completed += batch.size();
    if (batch.size() > s->max_batch) s->max_batch = batch.size();
    s->cv_done.notify_all();
  }
}

}  // namespace gg

using namespace gg;

extern "C" {

int gg_server_create(size_t budget_bytes, int workers, gg_server **out) {
  if (!out || workers < 1 || workers > kMaxSlots - 1)
    return fail(GG_EINVAL, "gg_server_create: 1 to 32 workers");
  GG_TRY(ensure_init());
  gg_server *s = new gg_server;
  s->budget = budget_bytes;
  hipGetDevice(&s->device);
  s->streams.resize(workers);
  s->bufs.resize(workers);
  // BFS batches take milliseconds where SSSP and BC batches of the same
  // graph take tens to hundreds: reserved workers keep BFS queries from
  // waiting behind them (at least one worker takes every kind)
  s->bfs_workers = (int)std::min<int64_t>(options().serve_bfs_workers, workers - 1);
  for (int k = 0; k < workers; k++) {
    if (hipStreamCreateWithFlags(&s->streams[k], hipStreamNonBlocking) != hipSuccess) {
      for (int j = 0; j < k; j++) hipStreamDestroy(s->streams[j]);
      delete s;
      return fail(GG_EHIP, "gg_server_create: stream");
    }
  }
  for (int k = 0; k < workers; k++) s->threads.emplace_back(worker_main, s, k);
  *out = s;
  return GG_OK;
}

int gg_server_destroy(gg_server *s) {
  if (!s) return GG_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->threads) t.join();
  for (size_t k = 0; k < s->bufs.size(); k++)
    for (size_t gi = 0; gi < s->bufs[k].size(); gi++) release_worker_graph(s, (int)k, gi);
  for (hipStream_t st : s->streams) hipStreamDestroy(st);
  for (auto &kv : s->queries) delete kv.second;
  for (SQuery *q : s->pending) delete q;
  delete s;
  return GG_OK;
}

int gg_server_add_graph(gg_server *s, gg_graph *g, int *graph_id) {
  if (!s || !g || !graph_id) return fail(GG_EINVAL, "gg_server_add_graph: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->used + g->device_bytes + serve_bytes(g, GG_QUERY_BFS) > s->budget)
    return fail(GG_EBUDGET, "gg_server_add_graph: graph and one worker's buffers exceed the budget");
  for (gg_graph *h : s->graphs)
    if (h == g) return fail(GG_EINVAL, "gg_server_add_graph: graph already added");
  s->used += g->device_bytes;
  s->graphs.push_back(g);
  s->running.resize(s->graphs.size() * 3, 0);
  for (auto &v : s->bufs) v.emplace_back();
  *graph_id = (int)s->graphs.size() - 1;
  return GG_OK;
}

int gg_server_submit(gg_server *s, int graph_id, int kind, uint32_t root, int want_result,
                     uint64_t *query_id) {
  if (!s || !query_id) return fail(GG_EINVAL, "gg_server_submit: null argument");
  if (kind < GG_QUERY_BFS || kind > GG_QUERY_BC) return fail(GG_EINVAL, "unknown query kind");
  std::lock_guard<std::mutex> lk(s->mu);
  if (graph_id < 0 || graph_id >= (int)s->graphs.size())
    return fail(GG_EINVAL, "gg_server_submit: unknown graph");
  gg_graph *g = s->graphs[graph_id];
  if (root >= g->vertices) return fail(GG_EINVAL, "gg_server_submit: root out of range");
  if (kind == GG_QUERY_SSSP && !g->weighted)
    return fail(GG_EINVAL, "gg_server_submit: SSSP needs a weighted graph");
  SQuery *q = new SQuery;
  q->id = s->next_id++;
  q->graph = graph_id;
  q->kind = kind;
  q->root = root;
  q->want = want_result != 0;
  q->t_arrive = now_s();
  s->queries[q->id] = q;
  s->pending.push_back(q);
  s->submitted++;
  *query_id = q->id;
  s->cv_work.notify_one();
  return GG_OK;
}

int gg_server_wait(gg_server *s, uint64_t query_id, void *h_result, size_t result_bytes,
                   double *latency_ms) {
  if (!s) return fail(GG_EINVAL, "gg_server_wait: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->queries.find(query_id);
  if (it == s->queries.end()) return fail(GG_EINVAL, "gg_server_wait: unknown query");
  SQuery *q = it->second;
  s->cv_done.wait(lk, [&] { return q->done; });
  s->queries.erase(it);
  lk.unlock();
  int r = q->status;
  std::string err = q->err;
  if (r == GG_OK && h_result) {
    if (result_bytes < q->result.size()) {
      delete q;
      return fail(GG_EINVAL, "gg_server_wait: result buffer too small");
    }
    if (q->result.empty()) {
      delete q;
      return fail(GG_EINVAL, "gg_server_wait: result was not kept (want_result = 0)");
    }
    memcpy(h_result, q->result.data(), q->result.size());
  }
  if (latency_ms) *latency_ms = (q->t_done - q->t_arrive) * 1e3;
  delete q;
  if (r != GG_OK) return fail(r, "query failed in its batch: " + err);
  return GG_OK;
}

int gg_server_replay(gg_server *s, int n, const double *arrival_s, const int *graph_id,
                     const int *kind, const uint32_t *root, double *latency_ms, double *wall_s) {
  if (!s || n < 0 || (n && (!arrival_s || !graph_id || !kind || !root || !latency_ms)))
    return fail(GG_EINVAL, "gg_server_replay: bad argument");
  std::vector<uint64_t> ids(n);
  double t0 = now_s();
  for (int i = 0; i < n; i++) {
    double due = t0 + arrival_s[i];
    double dt = due - now_s();
    if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(dt));
    GG_TRY(gg_server_submit(s, graph_id[i], kind[i], root[i], 0, &ids[i]));
    std::lock_guard<std::mutex> lk(s->mu);
    s->queries[ids[i]]->t_arrive = due;  // the scheduled arrival, not the submit call
  }
  int status = GG_OK;
  std::string err;
  double last = t0;
  for (int i = 0; i < n; i++) {
    std::unique_lock<std::mutex> lk(s->mu);
    SQuery *q = s->queries[ids[i]];
    s->cv_done.wait(lk, [&] { return q->done; });
    latency_ms[i] = (q->t_done - q->t_arrive) * 1e3;
    if (q->t_done > last) last = q->t_done;
    if (q->status != GG_OK) {
      status = q->status;
      err = q->err;
    }
    s->queries.erase(ids[i]);
    delete q;
  }
  if (wall_s) *wall_s = last - t0;
  if (status != GG_OK) return fail(status, "a replayed query failed: " + err);
  return GG_OK;
}

int gg_server_get_info(gg_server *s, gg_server_info *info) {
  if (!s || !info) return fail(GG_EINVAL, "gg_server_get_info: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  info->budget_bytes = s->budget;
  info->used_bytes = s->used;
  info->graphs = (int)s->graphs.size();
  info->workers = (int)s->threads.size();
  info->submitted = s->submitted;
  info->completed = s->completed;
  info->batches = s->batches;
  info->max_batch = s->max_batch;
  return GG_OK;
}

}  // extern "C"

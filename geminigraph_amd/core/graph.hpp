// Graph<EdgeData> facade: the reference's host API (core/graph.hpp:91-1869)
// for one partition on one GPU, on top of the C ABI in gg.h.
//
// The graph is built by libgg (device radix-sort CSR/CSC, original vertex
// ids).  process_vertices / process_edges keep the reference's signatures
// except that the callbacks are template parameters, not std::function: they
// run on the device, so toolkits pass `[=] __device__` lambdas and reach
// device state through views (GraphView, BitmapView) instead of host
// pointers (a host pointer cannot be followed on the device).
//
// Engine semantics follow the reference:
//   * switch: sparse iff sum of out-degrees over `active` < edges / 20
//     (core/graph.hpp:1503-1509);
//   * BSP: every signal (and its emits) completes before any slot runs
//     (core/graph.hpp:1541-1556 then :1585; :1755-1793 then :1801);
//   * sparse slots get (vertex, message, out-adjacency), dense signals get
//     (vertex, in-adjacency); long lists are cut into disjoint sub-lists, one
//     callback each, as the reference does per socket (core/graph.hpp:1600-1633,
//     :1767-1770), so callbacks must combine partials the way Gemini's do;
//   * the reducer is the sum of the slot / process return values.
// Vertex arrays are device memory: read them on the host with to_host()
// (the gather_vertex_array step) and poke single entries with set_value().
#ifndef GG_CORE_GRAPH_HPP
#define GG_CORE_GRAPH_HPP

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "atomic.hpp"
#include "bitmap.hpp"
#include "filesystem.hpp"
#include "gg.h"
#include "mpi.hpp"
#include "time.hpp"
#include "type.hpp"

namespace gg_facade {

inline void check(int code, const char *what) {
  if (code != GG_OK) {
    fprintf(stderr, "gg: %s: %s\n", what, gg_last_error());
    abort();
  }
}

constexpr int kLightMax = 256;  // lists above this go to one wave with lane sub-lists
constexpr int kSlots = 256;     // spread reducer slots (16 x 8 bytes apart)

template <typename M>
struct MsgUnit {  // core/graph.hpp:61-65
  VertexId vertex;
  M msg_data;
} __attribute__((packed));

struct EmitBlock {
  void *buf;
  unsigned long long count;
  unsigned long long cap;
};

// spread reduction of R into kSlots slots
template <typename R>
__device__ inline void slot_add(R *slots, R v) {
  if (v == R(0)) return;
  R *p = slots + (size_t)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kSlots) * 16;
  if constexpr (sizeof(R) == 8 && __is_same(R, double))
    unsafeAtomicAdd(p, v);
  else if constexpr (sizeof(R) == 8)
    atomicAdd((unsigned long long *)p, (unsigned long long)v);
  else if constexpr (__is_same(R, float))
    atomicAdd(p, v);
  else
    atomicAdd((unsigned *)p, (unsigned)v);
}

template <typename R, typename F>
__global__ void process_vertices_kernel(F f, const unsigned long *active, VertexId V, R *slots) {
  for (VertexId base = blockIdx.x * blockDim.x; base < V; base += gridDim.x * blockDim.x) {
    VertexId v = base + threadIdx.x;
    if (v < V && (active[WORD_OFFSET(v)] >> BIT_OFFSET(v)) & 1ul) slot_add<R>(slots, (R)f(v));
  }
}

template <typename F>
__global__ void sparse_signal_kernel(F f, const unsigned long *active, VertexId V) {
  for (VertexId base = blockIdx.x * blockDim.x; base < V; base += gridDim.x * blockDim.x) {
    VertexId v = base + threadIdx.x;
    if (v < V && (active[WORD_OFFSET(v)] >> BIT_OFFSET(v)) & 1ul) f(v);
  }
}

// One thread per message for short lists; long lists are queued for waves.
template <typename R, typename M, typename E, typename F>
__global__ void sparse_slot_kernel(F f, const MsgUnit<M> *msgs, unsigned long long n,
                                   const EdgeId *index, AdjUnit<E> *adj, R *slots,
                                   unsigned *heavy, unsigned *n_heavy) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x) {
    VertexId v = msgs[i].vertex;
    EdgeId b = index[v], e = index[v + 1];
    if (e == b) continue;  // outgoing_adj_bitmap test (core/graph.hpp:1630)
    if (e - b > (EdgeId)kLightMax) {
      heavy[atomicAdd(n_heavy, 1u)] = (unsigned)i;
      continue;
    }
    slot_add<R>(slots, (R)f(v, msgs[i].msg_data, VertexAdjList<E>(adj + b, adj + e)));
  }
}

template <typename R, typename M, typename E, typename F>
__global__ void sparse_slot_heavy_kernel(F f, const MsgUnit<M> *msgs, const unsigned *heavy,
                                         unsigned n_heavy, const EdgeId *index, AdjUnit<E> *adj,
                                         R *slots) {
  unsigned wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  unsigned lane = threadIdx.x & 63, waves = gridDim.x * (blockDim.x >> 6);
  for (unsigned h = wave; h < n_heavy; h += waves) {
    const MsgUnit<M> &m = msgs[heavy[h]];
    EdgeId b = index[m.vertex], e = index[m.vertex + 1];
    EdgeId per = (e - b + 63) / 64;
    EdgeId lb = b + lane * per, le = lb + per < e ? lb + per : e;
    if (lb < le) slot_add<R>(slots, (R)f(m.vertex, m.msg_data, VertexAdjList<E>(adj + lb, adj + le)));
  }
}

template <typename E, typename F>
__global__ void dense_signal_kernel(F f, VertexId V, const EdgeId *index, AdjUnit<E> *adj) {
  for (VertexId base = blockIdx.x * blockDim.x; base < V; base += gridDim.x * blockDim.x) {
    VertexId v = base + threadIdx.x;
    if (v >= V) continue;
    EdgeId b = index[v], e = index[v + 1];
    if (e > b && e - b <= (EdgeId)kLightMax) f(v, VertexAdjList<E>(adj + b, adj + e));
  }
}

template <typename E, typename F>
__global__ void dense_signal_heavy_kernel(F f, const VertexId *heavy, unsigned n_heavy,
                                          const EdgeId *index, AdjUnit<E> *adj) {
  unsigned wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  unsigned lane = threadIdx.x & 63, waves = gridDim.x * (blockDim.x >> 6);
  for (unsigned h = wave; h < n_heavy; h += waves) {
    VertexId v = heavy[h];
    EdgeId b = index[v], e = index[v + 1];
    EdgeId per = (e - b + 63) / 64;
    EdgeId lb = b + lane * per, le = lb + per < e ? lb + per : e;
    if (lb < le) f(v, VertexAdjList<E>(adj + lb, adj + le));
  }
}

template <typename R, typename M, typename F>
__global__ void dense_slot_kernel(F f, const MsgUnit<M> *msgs, unsigned long long n, R *slots) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    slot_add<R>(slots, (R)f(msgs[i].vertex, msgs[i].msg_data));
}

template <typename T>
__global__ void fill_kernel(T *p, T v, VertexId n) {
  for (VertexId i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = v;
}

template <int = 0>
__global__ void pack_adj_kernel(const VertexId *nbr, const float *w, EdgeId n,
                                       AdjUnit<float> *out) {
  for (EdgeId i = blockIdx.x * (EdgeId)blockDim.x + threadIdx.x; i < n;
       i += (EdgeId)gridDim.x * blockDim.x) {
    out[i].neighbour = nbr[i];
    out[i].edge_data = w[i];
  }
}

inline unsigned grid_for(unsigned long long n) {
  unsigned long long g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace gg_facade

// Device-copyable view used inside callbacks: graph fields plus emit().
template <typename EdgeData>
struct GraphView {
  VertexId vertices;
  EdgeId edges;
  const VertexId *out_degree;
  const VertexId *in_degree;
  int partition_id;
  int partitions;
  gg_facade::EmitBlock *eb;
  // emit a message to a vertex (core/graph.hpp:1479-1490); legal inside signals
  template <typename M>
  __device__ void emit(VertexId vtx, M msg) const {
    unsigned long long i = atomicAdd(&eb->count, 1ull);
    if (i < eb->cap) {
      gg_facade::MsgUnit<M> *b = (gg_facade::MsgUnit<M> *)eb->buf;
      b[i].vertex = vtx;
      b[i].msg_data = msg;
    }
  }
};

template <typename EdgeData = Empty>
class Graph {
 public:
  int partition_id = 0;
  int partitions = 1;
  VertexId vertices = 0;
  EdgeId edges = 0;
  VertexId *out_degree = nullptr;  // device
  VertexId *in_degree = nullptr;   // device
  bool symmetric = false;

  Graph() {
    gg_facade::hip_check(hipMalloc((void **)&eb_, sizeof(gg_facade::EmitBlock)), "emit block");
    gg_facade::hip_check(hipMalloc((void **)&slots_, 16 * 8 * gg_facade::kSlots), "slots");
    gg_facade::hip_check(hipMalloc((void **)&n_heavy_, 4), "n_heavy");
  }
  ~Graph() {
    for (void *p : owned_) (void)hipFree(p);
    (void)hipFree(eb_);
    (void)hipFree(slots_);
    (void)hipFree(n_heavy_);
    (void)hipFree(msg_buf_);
    (void)hipFree(heavy_msgs_);
    if (g_) gg_graph_destroy(g_);
  }

  // core/graph.hpp:772-1339
  void load_directed(std::string path, VertexId vertices) {
    load(path, vertices, false);
  }
  // core/graph.hpp:350-757
  void load_undirected_from_directed(std::string path, VertexId vertices) {
    load(path, vertices, true);
  }
  // core/graph.hpp:760-769
  void transpose() {
    std::swap(out_degree, in_degree);
    std::swap(out_, in_);
  }

  // ---- vertex arrays (core/graph.hpp:219-324) ----
  template <typename T>
  T *alloc_vertex_array() {
    T *p = nullptr;
    gg_facade::hip_check(hipMalloc((void **)&p, sizeof(T) * (size_t)vertices), "alloc_vertex_array");
    return p;
  }
  template <typename T>
  T *alloc_interleaved_vertex_array() {
    return alloc_vertex_array<T>();
  }
  template <typename T>
  void fill_vertex_array(T *array, T value) {
    hipLaunchKernelGGL(gg_facade::fill_kernel<T>, dim3(gg_facade::grid_for(vertices)), dim3(256), 0,
                       0, array, value, vertices);
    gg_facade::hip_check(hipDeviceSynchronize(), "fill_vertex_array");
  }
  template <typename T>
  void dealloc_vertex_array(T *array) {
    (void)hipFree(array);
  }
  // one partition: nothing to move between ranks
  template <typename T>
  void gather_vertex_array(T *array, int root) {
    (void)array;
    (void)root;
    gg_facade::hip_check(hipDeviceSynchronize(), "gather_vertex_array");
  }
  template <typename T>
  std::vector<T> to_host(const T *array) const {
    std::vector<T> h(vertices);
    gg_facade::hip_check(hipMemcpy(h.data(), array, sizeof(T) * (size_t)vertices,
                                   hipMemcpyDeviceToHost),
                         "to_host");
    return h;
  }
  template <typename T>
  void set_value(T *array, VertexId v, T value) {
    gg_facade::hip_check(hipMemcpy(array + v, &value, sizeof(T), hipMemcpyHostToDevice), "set_value");
  }
  template <typename T>
  T get_value(const T *array, VertexId v) const {
    T x;
    gg_facade::hip_check(hipMemcpy(&x, array + v, sizeof(T), hipMemcpyDeviceToHost), "get_value");
    return x;
  }
  VertexSubset *alloc_vertex_subset() { return new VertexSubset(vertices); }

  GraphView<EdgeData> view() const {
    return GraphView<EdgeData>{vertices, edges, out_degree, in_degree, partition_id, partitions, eb_};
  }

  // ---- engine (core/graph.hpp:1406-1867) ----
  template <typename R, typename F>
  R process_vertices(F process, Bitmap *active) {
    clear_slots<R>();
    hipLaunchKernelGGL((gg_facade::process_vertices_kernel<R, F>), dim3(gg_facade::grid_for(vertices)),
                       dim3(256), 0, 0, process, (const unsigned long *)active->data, vertices,
                       (R *)slots_);
    return read_slots<R>();
  }

  template <typename R, typename M, typename SS, typename SL, typename DS, typename DL>
  R process_edges(SS sparse_signal, SL sparse_slot, DS dense_signal, DL dense_slot, Bitmap *active,
                  Bitmap *dense_selective = nullptr) {
    (void)dense_selective;  // only synchronises replicas across partitions
    const VertexId *od = out_degree;
    EdgeId active_edges = process_vertices<EdgeId>(
        [=] __device__(VertexId v) { return (EdgeId)od[v]; }, active);
    bool sparse = active_edges < edges / 20;
    last_mode_sparse = sparse;
    clear_slots<R>();
    if (sparse) {
      reset_emits<M>((size_t)vertices);
      hipLaunchKernelGGL((gg_facade::sparse_signal_kernel<SS>), dim3(gg_facade::grid_for(vertices)),
                         dim3(256), 0, 0, sparse_signal, (const unsigned long *)active->data, vertices);
      unsigned long long n = emitted();
      gg_facade::hip_check(hipMemset(n_heavy_, 0, 4), "n_heavy");
      if (n) {
        hipLaunchKernelGGL((gg_facade::sparse_slot_kernel<R, M, EdgeData, SL>),
                           dim3(gg_facade::grid_for(n)), dim3(256), 0, 0, sparse_slot,
                           (const gg_facade::MsgUnit<M> *)msg_buf_, n, out_.index, out_.adj,
                           (R *)slots_, heavy_msgs_, n_heavy_);
        unsigned nh = 0;
        gg_facade::hip_check(hipMemcpy(&nh, n_heavy_, 4, hipMemcpyDeviceToHost), "n_heavy");
        if (nh)
          hipLaunchKernelGGL((gg_facade::sparse_slot_heavy_kernel<R, M, EdgeData, SL>),
                             dim3(gg_facade::grid_for(64ull * nh)), dim3(256), 0, 0, sparse_slot,
                             (const gg_facade::MsgUnit<M> *)msg_buf_, heavy_msgs_, nh, out_.index,
                             out_.adj, (R *)slots_);
      }
    } else {
      reset_emits<M>((size_t)vertices + 64ull * in_.n_heavy);
      hipLaunchKernelGGL((gg_facade::dense_signal_kernel<EdgeData, DS>),
                         dim3(gg_facade::grid_for(vertices)), dim3(256), 0, 0, dense_signal, vertices,
                         in_.index, in_.adj);
      if (in_.n_heavy)
        hipLaunchKernelGGL((gg_facade::dense_signal_heavy_kernel<EdgeData, DS>),
                           dim3(gg_facade::grid_for(64ull * in_.n_heavy)), dim3(256), 0, 0,
                           dense_signal, in_.heavy, in_.n_heavy, in_.index, in_.adj);
      unsigned long long n = emitted();
      if (n)
        hipLaunchKernelGGL((gg_facade::dense_slot_kernel<R, M, DL>), dim3(gg_facade::grid_for(n)),
                           dim3(256), 0, 0, dense_slot, (const gg_facade::MsgUnit<M> *)msg_buf_, n,
                           (R *)slots_);
    }
    return read_slots<R>();
  }

  bool last_mode_sparse = false;

 private:
  struct DirView {
    EdgeId *index = nullptr;
    AdjUnit<EdgeData> *adj = nullptr;
    VertexId *heavy = nullptr;  // vertices with more than kLightMax entries
    unsigned n_heavy = 0;
  };
  gg_graph *g_ = nullptr;
  DirView out_, in_;
  gg_facade::EmitBlock *eb_ = nullptr;
  void *msg_buf_ = nullptr;
  size_t msg_cap_bytes_ = 0;
  void *slots_ = nullptr;
  unsigned *heavy_msgs_ = nullptr;
  unsigned *n_heavy_ = nullptr;
  std::vector<void *> owned_;

  void load(const std::string &path, VertexId v, bool undirected) {
    int weighted = sizeof(EdgeData) == 0 || __is_same(EdgeData, Empty) ? 0 : 1;
    static_assert(__is_same(EdgeData, Empty) || __is_same(EdgeData, float),
                  "edge data: Empty or float");
    gg_facade::check(gg_graph_load_ex(path.c_str(), v, weighted, undirected ? 1 : 0, 0u, &g_),
                     "load");
    gg_graph_info info;
    gg_facade::check(gg_graph_get_info(g_, &info), "info");
    gg_facade::check(gg_synchronize(), "sync");
    vertices = info.vertices;
    edges = info.edges;
    symmetric = info.symmetric;
    out_degree = (VertexId *)info.d_out_degree;
    in_degree = (VertexId *)info.d_in_degree;
    out_ = make_dir(info.d_out_index, info.d_out_nbr, info.d_out_w, info.arcs, out_degree);
    in_ = symmetric ? out_ : make_dir(info.d_in_index, info.d_in_nbr, info.d_in_w, info.arcs, in_degree);
    gg_facade::hip_check(hipMalloc((void **)&heavy_msgs_, sizeof(unsigned) * ((size_t)vertices + 1)),
                         "heavy msgs");
  }

  DirView make_dir(const uint64_t *index, const uint32_t *nbr, const float *w, uint64_t arcs,
                   const VertexId *deg) {
    DirView d;
    d.index = (EdgeId *)index;
    if constexpr (__is_same(EdgeData, Empty)) {
      d.adj = (AdjUnit<EdgeData> *)nbr;  // AdjUnit<Empty> is the bare neighbour id
    } else {
      AdjUnit<float> *a = nullptr;
      gg_facade::hip_check(hipMalloc((void **)&a, sizeof(AdjUnit<float>) * (arcs ? arcs : 1)), "adj");
      owned_.push_back(a);
      if (arcs)
        hipLaunchKernelGGL(gg_facade::pack_adj_kernel<0>, dim3(gg_facade::grid_for(arcs)), dim3(256), 0, 0,
                           nbr, w, (EdgeId)arcs, a);
      d.adj = (AdjUnit<EdgeData> *)a;
    }
    std::vector<VertexId> hd(vertices), hv;
    gg_facade::hip_check(hipMemcpy(hd.data(), deg, sizeof(VertexId) * vertices, hipMemcpyDeviceToHost),
                         "degrees");
    for (VertexId v = 0; v < vertices; v++)
      if (hd[v] > (VertexId)gg_facade::kLightMax) hv.push_back(v);
    d.n_heavy = (unsigned)hv.size();
    gg_facade::hip_check(hipMalloc((void **)&d.heavy, sizeof(VertexId) * (hv.size() + 1)), "heavy");
    owned_.push_back(d.heavy);
    if (!hv.empty())
      gg_facade::hip_check(hipMemcpy(d.heavy, hv.data(), sizeof(VertexId) * hv.size(),
                                     hipMemcpyHostToDevice),
                           "heavy");
    gg_facade::hip_check(hipDeviceSynchronize(), "make_dir");
    return d;
  }

  template <typename M>
  void reset_emits(size_t cap) {
    size_t bytes = sizeof(gg_facade::MsgUnit<M>) * (cap ? cap : 1);
    if (bytes > msg_cap_bytes_) {
      (void)hipFree(msg_buf_);
      gg_facade::hip_check(hipMalloc(&msg_buf_, bytes), "message buffer");
      msg_cap_bytes_ = bytes;
    }
    gg_facade::EmitBlock h{msg_buf_, 0ull, (unsigned long long)cap};
    gg_facade::hip_check(hipMemcpy(eb_, &h, sizeof(h), hipMemcpyHostToDevice), "emit block");
  }
  unsigned long long emitted() {
    gg_facade::EmitBlock h;
    gg_facade::hip_check(hipMemcpy(&h, eb_, sizeof(h), hipMemcpyDeviceToHost), "emitted");
    if (h.count > h.cap) {
      fprintf(stderr, "gg: message buffer overflow (%llu > %llu)\n", h.count, h.cap);
      abort();
    }
    return h.count;
  }
  template <typename R>
  void clear_slots() {
    gg_facade::hip_check(hipMemset(slots_, 0, 16 * 8 * gg_facade::kSlots), "slots");
  }
  template <typename R>
  R read_slots() {
    std::vector<unsigned char> h(16 * 8 * gg_facade::kSlots);
    gg_facade::hip_check(hipMemcpy(h.data(), slots_, h.size(), hipMemcpyDeviceToHost), "slots");
    R s = 0;
    for (int i = 0; i < gg_facade::kSlots; i++) s += *(R *)(h.data() + (size_t)i * 16 * sizeof(R));
    return s;
  }
};

#endif

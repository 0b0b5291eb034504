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

constexpr int kLightMax = 256;  // dense: longer in-lists go to one wave with lane sub-lists
constexpr int kSubList = 32;    // sparse: out-lists reach slots in sub-lists of at most 32
constexpr int kWaveList = 64 * kSubList;  // sparse: longer out-lists go to the heavy kernel
constexpr int kSlots = 256;     // spread reducer slots (16 x 8 bytes apart)
constexpr int kPhaseGrid = 2048;  // blocks of the persistent phase kernels (one round of residency)

template <typename M>
struct MsgUnit {  // core/graph.hpp:61-65
  VertexId vertex;
  M msg_data;
} __attribute__((packed));

constexpr int kEmitSegs = 1024;  // message segments, each with its own counter
constexpr int kPart = 2048;      // dense: heavy in-lists reach signals in parts of this many

// Messages of one call.  A signal's emit goes to one of kEmitSegs segments
// (chosen by the emitting wave, so no counter takes more than a small share
// of the atomics; one counter for all would serialise them at one L2
// address), or, once its segment is full, to an overflow region; a pack
// pass then lays the messages out densely in `buf` for the slots.
struct EmitBlock {
  void *buf;                    // packed messages the slots read
  unsigned long long count;     // messages packed
  void *seg_buf;                // kEmitSegs x seg_cap messages, then ovf_cap more
  unsigned *seg_cnt;            // kEmitSegs counters, 32 apart (one 128-byte line each)
  unsigned long long seg_cap, ovf_cap;
  unsigned long long ovf_count;
  unsigned long long dropped;   // messages beyond every region (the host aborts)
};

// Device state of one process_edges / process_vertices call.  The switch is
// decided on the device (`mode`), every phase kernel of both modes is
// launched and returns at once in the mode not chosen, and the message and
// heavy-list counts stay on the device, so a call makes one host round
// trip: the read-back of this block, for the reducer the API returns.
struct Ctl {
  EmitBlock eb;
  unsigned long long active_edges;
  unsigned long long result;  // the reducer's bits (R of the call)
  unsigned mode;              // 0 = sparse (push), 1 = dense (pull)
  unsigned n_heavy;           // sparse messages whose lists exceed kWaveList
  unsigned long long pack_off[kEmitSegs + 1];  // segment offsets of the packed messages
};

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// One atomic per wave into one of kSlots spread slots (the slots are zero
// between calls: the reduce kernel clears what it reads).
template <typename R>
__device__ inline void slot_add_wave(R *slots, R v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) || v == R(0)) return;
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

// A trivially copyable value from lane `src` (4-byte pieces).
template <typename T>
__device__ inline T shfl_any(const T &x, int src) {
  constexpr int n = (sizeof(T) + 3) / 4;
  unsigned w[n] = {};
  __builtin_memcpy(w, &x, sizeof(T));
#pragma unroll
  for (int i = 0; i < n; i++) w[i] = (unsigned)__shfl((int)w[i], src);
  T y;
  __builtin_memcpy(&y, w, sizeof(T));
  return y;
}

// A wave loads 64 bitmap words at once (one per lane), then visits only the
// non-empty ones, each as 64 lanes over its 64 consecutive vertices (so
// callbacks that set bits of a word in lane order still meet on one word).
// Calls `g(v)` for every set bit v < V.
template <typename G>
__device__ inline void for_each_set_bit(const unsigned long *active, VertexId V, G g) {
  const int lane = threadIdx.x & 63;
  const unsigned long long nwords = ((unsigned long long)V + 63) / 64;
  const unsigned long long W = (unsigned long long)gridDim.x * (blockDim.x >> 6);
  const unsigned long long gw =
      (unsigned long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (unsigned long long w0 = gw * 64; w0 < nwords; w0 += W * 64) {
    const unsigned long long wi = w0 + lane;
    const unsigned long word = wi < nwords ? active[wi] : 0ul;
    for (unsigned long long m = __ballot(word != 0ul); m; m &= m - 1) {
      const int k = __builtin_ctzll(m);
      const unsigned long wk = shfl_any(word, k);
      const unsigned long long v = (w0 + k) * 64 + lane;
      if (v < V && (wk >> lane) & 1ul) g((VertexId)v);
    }
  }
}

template <typename R, typename F>
__global__ void process_vertices_kernel(F f, const unsigned long *active, VertexId V, R *slots) {
  R x = R(0);
  for_each_set_bit(active, V, [&](VertexId v) { x += (R)f(v); });
  slot_add_wave<R>(slots, x);
}

// One block: the sum of the slots as R into *out, the slots back to zero.
template <typename R>
__global__ void reduce_slots_kernel(R *slots, unsigned long long *out) {
  __shared__ R part[4];
  R x = slots[(size_t)threadIdx.x * 16];
  slots[(size_t)threadIdx.x * 16] = R(0);
  x = wave_sum(x);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    R t = part[0] + part[1] + part[2] + part[3];
    unsigned long long bits = 0;
    __builtin_memcpy(&bits, &t, sizeof(R));
    *out = bits;
  }
}

// One block: the switch (core/graph.hpp:1503-1509) from the active out-degree
// sum in the slots, which it clears for the call's reducer; resets the
// message count for the mode chosen.
template <int = 0>
__global__ void decide_kernel(Ctl *c, EdgeId *slots, EdgeId threshold) {
  __shared__ EdgeId part[4];
  EdgeId x = slots[(size_t)threadIdx.x * 16];
  slots[(size_t)threadIdx.x * 16] = 0;
  x = wave_sum(x);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    const EdgeId active_edges = part[0] + part[1] + part[2] + part[3];
    const unsigned mode = active_edges < threshold ? 0u : 1u;
    c->active_edges = active_edges;
    c->mode = mode;
    c->n_heavy = 0;
    c->eb.count = 0;
    c->eb.ovf_count = 0;
    c->eb.dropped = 0;
  }
  for (int k = threadIdx.x; k < kEmitSegs; k += blockDim.x) c->eb.seg_cnt[(size_t)k * 32] = 0;
}

// One block of kEmitSegs threads: offsets of the segments' messages in the
// packed array, and the packed count.
template <int = 0>
__global__ void pack_offsets_kernel(Ctl *c) {
  __shared__ unsigned long long wsum[kEmitSegs / 64];
  const int t = threadIdx.x, lane = t & 63;
  const unsigned long long cnt = c->eb.seg_cnt[(size_t)t * 32];
  const unsigned long long n = cnt < c->eb.seg_cap ? cnt : c->eb.seg_cap;
  unsigned long long inc = n;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  if (lane == 63) wsum[t >> 6] = inc;
  __syncthreads();
  unsigned long long base = 0;
  for (int w = 0; w < (t >> 6); w++) base += wsum[w];
  c->pack_off[t] = base + inc - n;
  if (t == kEmitSegs - 1) {
    c->pack_off[kEmitSegs] = base + inc;
    const unsigned long long o = c->eb.ovf_count < c->eb.ovf_cap ? c->eb.ovf_count : c->eb.ovf_cap;
    c->eb.count = base + inc + o;
  }
}

// Block s copies segment s (s == kEmitSegs: the overflow region) into place.
template <typename M>
__global__ void pack_copy_kernel(Ctl *c) {
  const MsgUnit<M> *src = (const MsgUnit<M> *)c->eb.seg_buf;
  MsgUnit<M> *dst = (MsgUnit<M> *)c->eb.buf;
  for (unsigned s = blockIdx.x; s <= (unsigned)kEmitSegs; s += gridDim.x) {
    const unsigned long long b = c->pack_off[s < kEmitSegs ? s : kEmitSegs];
    const unsigned long long n = s < kEmitSegs ? c->pack_off[s + 1] - b : c->eb.count - b;
    const MsgUnit<M> *from = src + (size_t)s * c->eb.seg_cap;
    for (unsigned long long i = threadIdx.x; i < n; i += blockDim.x) dst[b + i] = from[i];
  }
}

template <typename F>
__global__ void sparse_signal_kernel(F f, const Ctl *c, const unsigned long *active, VertexId V) {
  if (c->mode != 0) return;
  for_each_set_bit(active, V, [&](VertexId v) { f(v); });
}

// Sparse slots, 64 messages per wave step: lists up to kSubList entries go
// to their own lane whole; longer ones (up to kWaveList) are cut into
// kSubList-entry sub-lists spread over the wave's lanes; the longest go to
// the heavy kernel.  Sub-lists are disjoint, as the reference's per-socket
// ones are (core/graph.hpp:1600-1633), and the reducer is their sum.
template <typename R, typename M, typename E, typename F>
__global__ void sparse_slot_kernel(F f, Ctl *c, const EdgeId *index, AdjUnit<E> *adj, R *slots,
                                   unsigned *heavy, unsigned heavy_cap) {
  if (c->mode != 0) return;
  const unsigned long long n = c->eb.count;
  const MsgUnit<M> *msgs = (const MsgUnit<M> *)c->eb.buf;
  const int lane = threadIdx.x & 63;
  const unsigned long long W = (unsigned long long)gridDim.x * (blockDim.x >> 6);
  const unsigned long long gw = (unsigned long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (unsigned long long base = gw * 64; base < n; base += W * 64) {
    const unsigned long long i = base + lane;
    VertexId v = 0;
    M msg{};
    EdgeId b = 0, e = 0;
    if (i < n) {
      v = msgs[i].vertex;
      msg = msgs[i].msg_data;
      b = index[v];
      e = index[v + 1];
    }
    const EdgeId d = e - b;
    R acc = R(0);
    unsigned chunks = 0;
    if (d > (EdgeId)kWaveList) {
      const unsigned k = atomicAdd(&c->n_heavy, 1u);
      if (k < heavy_cap)
        heavy[k] = (unsigned)i;
      else
        atomicAdd(&c->eb.dropped, 1ull);  // more than one per vertex: the host aborts
    } else if (d > (EdgeId)kSubList) {
      chunks = (unsigned)((d + kSubList - 1) / kSubList);
    } else if (d) {
      acc += (R)f(v, msg, VertexAdjList<E>(adj + b, adj + e));
    }
    // exclusive scan of the chunk counts over the wave
    unsigned inc = chunks;
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned y = (unsigned)__shfl_up((int)inc, off);
      if (lane >= off) inc += y;
    }
    const unsigned total = (unsigned)__shfl((int)inc, 63), excl = inc - chunks;
    for (unsigned j0 = 0; j0 < total; j0 += 64) {
      const unsigned j = j0 + lane;
      // owner: the last lane whose exclusive start is <= j (a lane without
      // chunks shares its start with the next lane, so it is never the
      // last one); six steps for every lane, so the shuffles stay uniform
      int lo = 0;
      for (int step = 32; step > 0; step >>= 1) {
        const unsigned x = (unsigned)__shfl((int)excl, lo + step < 64 ? lo + step : 63);
        if (lo + step < 64 && x <= j) lo += step;
      }
      // all lanes take part in the shuffles above; only j < total calls f
      const VertexId vk = (VertexId)__shfl((int)v, lo);
      const M mk = shfl_any(msg, lo);
      const EdgeId bk = shfl_any(b, lo), ek = shfl_any(e, lo);
      const unsigned ex = (unsigned)__shfl((int)excl, lo);
      if (j < total) {
        const EdgeId lb = bk + (EdgeId)(j - ex) * kSubList;
        const EdgeId le = lb + kSubList < ek ? lb + kSubList : ek;
        acc += (R)f(vk, mk, VertexAdjList<E>(adj + lb, adj + le));
      }
    }
    slot_add_wave<R>(slots, acc);
  }
}

// Sparse slots of the longest lists.  One block lays out the kSubList-entry
// sub-lists of every heavy message in one sequence (pre[h] = the first
// sub-list of heavy message h, pre[nh] = their number); the slot kernel then
// spreads that sequence over the whole grid, so a few hub messages (a BFS
// level just after the root) still fill the chip.
template <typename M>
__global__ void heavy_prefix_kernel(const Ctl *c, const unsigned *heavy, unsigned heavy_cap,
                                    const EdgeId *index, unsigned long long *pre) {
  if (c->mode != 0) return;
  __shared__ unsigned long long wsum[16];
  const unsigned nh = c->n_heavy < heavy_cap ? c->n_heavy : heavy_cap;
  const MsgUnit<M> *msgs = (const MsgUnit<M> *)c->eb.buf;
  const unsigned lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long carry = 0;
  for (unsigned base = 0; base < nh; base += blockDim.x) {
    const unsigned h = base + threadIdx.x;
    unsigned long long n = 0;
    if (h < nh) {
      const VertexId v = msgs[heavy[h]].vertex;
      n = (index[v + 1] - index[v] + kSubList - 1) / kSubList;
    }
    unsigned long long inc = n;  // inclusive scan over the wave
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned long long y = shfl_any(inc, lane >= (unsigned)off ? (int)(lane - off) : (int)lane);
      if (lane >= (unsigned)off) inc += y;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    unsigned long long before = carry, all = 0;
    for (unsigned k = 0; k < nw; k++) {
      if (k < wv) before += wsum[k];
      all += wsum[k];
    }
    if (h < nh) pre[h] = before + inc - n;
    carry += all;
    __syncthreads();  // wsum is rewritten by the next round
  }
  if (threadIdx.x == 0) pre[nh] = carry;
}
template <typename R, typename M, typename E, typename F>
__global__ void sparse_slot_heavy_kernel(F f, const Ctl *c, const unsigned *heavy,
                                         unsigned heavy_cap, const unsigned long long *pre,
                                         const EdgeId *index, AdjUnit<E> *adj, R *slots) {
  if (c->mode != 0) return;
  const unsigned nh = c->n_heavy < heavy_cap ? c->n_heavy : heavy_cap;
  if (!nh) return;
  const MsgUnit<M> *msgs = (const MsgUnit<M> *)c->eb.buf;
  const unsigned long long total = pre[nh];
  const unsigned long long T = (unsigned long long)gridDim.x * blockDim.x;
  R acc = R(0);
  for (unsigned long long s = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; s < total;
       s += T) {
    unsigned lo = 0, hi = nh;  // the last heavy message whose first sub-list is <= s
    while (hi - lo > 1) {
      const unsigned mid = (lo + hi) >> 1;
      if (pre[mid] <= s)
        lo = mid;
      else
        hi = mid;
    }
    const MsgUnit<M> &m = msgs[heavy[lo]];
    const EdgeId e = index[m.vertex + 1];
    const EdgeId lb = index[m.vertex] + (EdgeId)(s - pre[lo]) * kSubList;
    const EdgeId le = lb + kSubList < e ? lb + kSubList : e;
    acc += (R)f(m.vertex, m.msg_data, VertexAdjList<E>(adj + lb, adj + le));
  }
  slot_add_wave<R>(slots, acc);
}

// Dense signals: one thread per vertex over its whole in-list (up to
// kLightMax entries), as one partition does.  The vertices come in id order
// (`order`, built at load), so neighbouring lanes read neighbouring lists.
template <typename E, typename F>
__global__ void dense_signal_kernel(F f, const Ctl *c, const VertexId *order, VertexId n,
                                    const EdgeId *index, AdjUnit<E> *adj) {
  if (c->mode != 1) return;
  for (VertexId i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const VertexId v = order[i];
    const EdgeId b = index[v], e = index[v + 1];
    f(v, VertexAdjList<E>(adj + b, adj + e));
  }
}

// Longer in-lists: parts of at most kPart entries, one thread each, each
// part its own dense_signal call, as the reference's per-socket parts are
// (core/graph.hpp:1767-1770); dense_slot combines them.  A list's first part
// holds its highest-degree in-neighbours (lists are hub-first).
template <typename E, typename F>
__global__ void dense_signal_parts_kernel(F f, const Ctl *c, const VertexId *part_v,
                                          const EdgeId *part_b, unsigned n_parts,
                                          const EdgeId *index, AdjUnit<E> *adj) {
  if (c->mode != 1) return;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n_parts;
       i += gridDim.x * blockDim.x) {
    const VertexId v = part_v[i];
    const EdgeId b = part_b[i], e = index[v + 1];
    f(v, VertexAdjList<E>(adj + b, adj + (b + kPart < e ? b + kPart : e)));
  }
}

template <typename R, typename M, typename F>
__global__ void dense_slot_kernel(F f, const Ctl *c, R *slots) {
  if (c->mode != 1) return;
  const unsigned long long n = c->eb.count;
  const MsgUnit<M> *msgs = (const MsgUnit<M> *)c->eb.buf;
  const unsigned long long S = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long base = (unsigned long long)blockIdx.x * blockDim.x; base < n;
       base += S) {
    const unsigned long long i = base + threadIdx.x;
    R x = R(0);
    if (i < n) x = (R)f(msgs[i].vertex, msgs[i].msg_data);
    slot_add_wave<R>(slots, x);
  }
}

template <typename T>
__global__ void fill_kernel(T *p, T v, VertexId n) {
  for (VertexId i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T>
__global__ void set_value_kernel(T *p, T v) {
  *p = v;
}

template <int = 0>
__global__ void set_buf_kernel(Ctl *c, void *buf, void *seg_buf, unsigned *seg_cnt,
                               unsigned long long seg_cap, unsigned long long ovf_cap) {
  c->eb.buf = buf;
  c->eb.seg_buf = seg_buf;
  c->eb.seg_cnt = seg_cnt;
  c->eb.seg_cap = seg_cap;
  c->eb.ovf_cap = ovf_cap;
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

// Blocks of 256 threads for n items, at most one per four wave slots of
// the chip (256 CUs x 4 SIMDs x 8 waves): every kernel here loops over its
// items, and a launch whose mode was not chosen (it returns at once) then
// costs a few microseconds instead of dispatching 65,536 blocks.
inline unsigned grid_for(unsigned long long n) {
  unsigned long long g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 2048 ? 2048 : g));
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
  // emit a message to a vertex (core/graph.hpp:1479-1490); legal inside
  // signals.  The lanes of a wave that emit at the same instant reserve
  // their slots in the wave's segment with one atomic (the reference fills
  // thread-local buffers for the same reason, core/graph.hpp:1482-1489).
  template <typename M>
  __device__ void emit(VertexId vtx, M msg) const {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long lanes = __ballot(1);  // the lanes emitting now
    const int leader = __ffsll((long long)lanes) - 1;
    const unsigned seg =
        (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % gg_facade::kEmitSegs;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(&eb->seg_cnt[(size_t)seg * 32], (unsigned)__popcll(lanes));
    base = (unsigned)__shfl((int)base, leader);
    const unsigned long long i = (unsigned long long)base + __popcll(lanes & below);
    gg_facade::MsgUnit<M> *m = (gg_facade::MsgUnit<M> *)eb->seg_buf;
    if (i < eb->seg_cap) {
      m[(size_t)seg * eb->seg_cap + i] = gg_facade::MsgUnit<M>{vtx, msg};
      return;
    }
    // the segment is full: the overflow region
    const unsigned long long over = __ballot(1);
    const int l2 = __ffsll((long long)over) - 1;
    unsigned long long ob = 0;
    if (lane == l2) ob = atomicAdd(&eb->ovf_count, (unsigned long long)__popcll(over));
    ob = __shfl(ob, l2);
    const unsigned long long j = ob + __popcll(over & below);
    if (j < eb->ovf_cap)
      m[(size_t)gg_facade::kEmitSegs * eb->seg_cap + j] = gg_facade::MsgUnit<M>{vtx, msg};
    else
      atomicAdd(&eb->dropped, 1ull);
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
    gg_facade::hip_check(hipMalloc((void **)&ctl_, sizeof(gg_facade::Ctl)), "control block");
    gg_facade::hip_check(hipMemset(ctl_, 0, sizeof(gg_facade::Ctl)), "control block");
    gg_facade::hip_check(hipMalloc((void **)&slots_, kSlotBytes), "slots");
    gg_facade::hip_check(hipMemset(slots_, 0, kSlotBytes), "slots");  // zero between calls
    gg_facade::hip_check(hipMalloc((void **)&seg_cnt_, sizeof(unsigned) * 32 * gg_facade::kEmitSegs),
                         "segment counters");
  }
  ~Graph() {
    for (void *p : owned_) (void)hipFree(p);
    (void)hipFree(ctl_);
    (void)hipFree(slots_);
    (void)hipFree(msg_buf_);
    (void)hipFree(seg_buf_);
    (void)hipFree(seg_cnt_);
    (void)hipFree(heavy_msgs_);
    (void)hipFree(heavy_pre_);
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
  T *alloc_vertex_array() {  // gg_malloc: freed arrays are kept for the next call
    T *p = nullptr;
    if (gg_malloc((void **)&p, sizeof(T) * (size_t)vertices) != GG_OK) {
      fprintf(stderr, "gg: alloc_vertex_array: %s\n", gg_last_error());
      abort();
    }
    return p;
  }
  template <typename T>
  T *alloc_interleaved_vertex_array() {
    return alloc_vertex_array<T>();
  }
  // queued on the null stream: later launches and copies see it in order
  template <typename T>
  void fill_vertex_array(T *array, T value) {
    hipLaunchKernelGGL(gg_facade::fill_kernel<T>, dim3(gg_facade::grid_for(vertices)), dim3(256), 0,
                       0, array, value, vertices);
  }
  template <typename T>
  void dealloc_vertex_array(T *array) {
    (void)gg_free(array);
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
    hipLaunchKernelGGL(gg_facade::set_value_kernel<T>, dim3(1), dim3(1), 0, 0, array + v, value);
  }
  template <typename T>
  T get_value(const T *array, VertexId v) const {
    T x;
    gg_facade::hip_check(hipMemcpy(&x, array + v, sizeof(T), hipMemcpyDeviceToHost), "get_value");
    return x;
  }
  VertexSubset *alloc_vertex_subset() { return new VertexSubset(vertices); }

  GraphView<EdgeData> view() const {
    return GraphView<EdgeData>{vertices, edges, out_degree, in_degree, partition_id, partitions,
                               &ctl_->eb};
  }

  // ---- engine (core/graph.hpp:1406-1867) ----
  template <typename R, typename F>
  R process_vertices(F process, Bitmap *active) {
    static_assert(sizeof(R) <= 8, "reducer of at most 8 bytes");
    hipLaunchKernelGGL((gg_facade::process_vertices_kernel<R, F>), dim3(gg_facade::grid_for(vertices)),
                       dim3(256), 0, 0, process, (const unsigned long *)active->data, vertices,
                       (R *)slots_);
    hipLaunchKernelGGL((gg_facade::reduce_slots_kernel<R>), dim3(1), dim3(gg_facade::kSlots), 0, 0,
                       (R *)slots_, &ctl_->result);
    return read_result<R>();
  }

  // Every launch is queued at once: the device decides the mode, the other
  // mode's kernels return at once, message counts stay on the device; the
  // host waits once, for the reducer.
  template <typename R, typename M, typename SS, typename SL, typename DS, typename DL>
  R process_edges(SS sparse_signal, SL sparse_slot, DS dense_signal, DL dense_slot, Bitmap *active,
                  Bitmap *dense_selective = nullptr) {
    (void)dense_selective;  // only synchronises replicas across partitions
    static_assert(sizeof(R) <= 8, "reducer of at most 8 bytes");
    ensure_messages<M>();
    const VertexId *od = out_degree;
    hipLaunchKernelGGL((gg_facade::process_vertices_kernel<EdgeId, gg_facade_out_degree>),
                       dim3(gg_facade::grid_for(vertices)), dim3(256), 0, 0,
                       gg_facade_out_degree{od}, (const unsigned long *)active->data, vertices,
                       (EdgeId *)slots_);
    hipLaunchKernelGGL(gg_facade::decide_kernel<0>, dim3(1), dim3(gg_facade::kSlots), 0, 0, ctl_,
                       (EdgeId *)slots_, (EdgeId)(edges / 20));  // core/graph.hpp:1509
    const dim3 grid(gg_facade::kPhaseGrid), block(256);
    R *slots = (R *)slots_;
    // signals of the mode chosen (BSP: all signals and emits complete before
    // any slot, core/graph.hpp:1541-1556 then :1585; :1755-1793 then :1801)
    hipLaunchKernelGGL((gg_facade::sparse_signal_kernel<SS>), dim3(gg_facade::grid_for(vertices)),
                       block, 0, 0, sparse_signal, ctl_, (const unsigned long *)active->data,
                       vertices);
    hipLaunchKernelGGL((gg_facade::dense_signal_kernel<EdgeData, DS>),
                       dim3(gg_facade::grid_for(in_.n_light)), block, 0, 0, dense_signal, ctl_,
                       in_.light, in_.n_light, in_.index, in_.adj);
    if (in_.n_parts)
      hipLaunchKernelGGL((gg_facade::dense_signal_parts_kernel<EdgeData, DS>),
                         dim3(gg_facade::grid_for(in_.n_parts)), block, 0, 0, dense_signal, ctl_,
                         in_.part_v, in_.part_b, in_.n_parts, in_.index, in_.adj);
    // the messages, packed
    hipLaunchKernelGGL(gg_facade::pack_offsets_kernel<0>, dim3(1), dim3(gg_facade::kEmitSegs), 0, 0,
                       ctl_);
    hipLaunchKernelGGL((gg_facade::pack_copy_kernel<M>), dim3(gg_facade::kEmitSegs + 1), block, 0,
                       0, ctl_);
    // slots of the mode chosen
    hipLaunchKernelGGL((gg_facade::sparse_slot_kernel<R, M, EdgeData, SL>), grid, block, 0, 0,
                       sparse_slot, ctl_, out_.index, out_.adj, slots, heavy_msgs_,
                       vertices + 1);
    hipLaunchKernelGGL(gg_facade::heavy_prefix_kernel<M>, dim3(1), dim3(1024), 0, 0, ctl_,
                       heavy_msgs_, vertices + 1, out_.index, heavy_pre_);
    hipLaunchKernelGGL((gg_facade::sparse_slot_heavy_kernel<R, M, EdgeData, SL>), grid, block, 0,
                       0, sparse_slot, ctl_, heavy_msgs_, vertices + 1, heavy_pre_, out_.index,
                       out_.adj, slots);
    hipLaunchKernelGGL((gg_facade::dense_slot_kernel<R, M, DL>), grid, block, 0, 0, dense_slot,
                       ctl_, slots);
    hipLaunchKernelGGL((gg_facade::reduce_slots_kernel<R>), dim3(1), dim3(gg_facade::kSlots), 0, 0,
                       slots, &ctl_->result);
    gg_facade::Ctl h;
    gg_facade::hip_check(hipMemcpy(&h, ctl_, sizeof(h), hipMemcpyDeviceToHost), "process_edges");
    if (h.eb.dropped) {
      fprintf(stderr, "gg: message buffer overflow (%llu messages dropped)\n", h.eb.dropped);
      abort();
    }
    last_mode_sparse = h.mode == 0;
    R r;
    __builtin_memcpy(&r, &h.result, sizeof(R));
    return r;
  }

  bool last_mode_sparse = false;

 private:
  static constexpr size_t kSlotBytes = 16 * 8 * gg_facade::kSlots;
  struct gg_facade_out_degree {
    const VertexId *od;
    __device__ EdgeId operator()(VertexId v) const { return (EdgeId)od[v]; }
  };
  struct DirView {
    EdgeId *index = nullptr;
    AdjUnit<EdgeData> *adj = nullptr;
    VertexId *light = nullptr;  // vertices with 1..kLightMax entries, in id order
    VertexId n_light = 0;
    VertexId *part_v = nullptr; // longer lists: parts of kPart entries (vertex, first entry)
    EdgeId *part_b = nullptr;
    unsigned n_parts = 0;
  };
  gg_graph *g_ = nullptr;
  DirView out_, in_;
  gg_facade::Ctl *ctl_ = nullptr;
  void *msg_buf_ = nullptr;   // packed messages
  void *seg_buf_ = nullptr;   // segments and overflow region
  unsigned *seg_cnt_ = nullptr;
  size_t msg_cap_bytes_ = 0;
  void *slots_ = nullptr;
  unsigned *heavy_msgs_ = nullptr;
  unsigned long long *heavy_pre_ = nullptr;  // sub-list offsets of the heavy messages
  std::vector<void *> owned_;

  void load(const std::string &path, VertexId v, bool undirected) {
    int weighted = sizeof(EdgeData) == 0 || __is_same(EdgeData, Empty) ? 0 : 1;
    static_assert(__is_same(EdgeData, Empty) || __is_same(EdgeData, float),
                  "edge data: Empty or float");
    // file ids, as the callbacks index vertex arrays by them; lists hubs first
    gg_facade::check(gg_graph_load_ex(path.c_str(), v, weighted, undirected ? 1 : 0,
                                      GG_BUILD_HUB_FIRST, &g_),
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
    gg_facade::hip_check(
        hipMalloc((void **)&heavy_pre_, sizeof(unsigned long long) * ((size_t)vertices + 2)),
        "heavy sub-list offsets");
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
    std::vector<VertexId> hd(vertices);
    std::vector<EdgeId> hidx((size_t)vertices + 1);
    gg_facade::hip_check(hipMemcpy(hd.data(), deg, sizeof(VertexId) * vertices, hipMemcpyDeviceToHost),
                         "degrees");
    gg_facade::hip_check(hipMemcpy(hidx.data(), index, sizeof(EdgeId) * ((size_t)vertices + 1),
                                   hipMemcpyDeviceToHost),
                         "offsets");
    // the light vertices in id order (neighbouring lists and destinations:
    // a wave's loads share lines and its messages reach the slots in
    // destination order; sorting them by degree for even lanes was slower,
    // BFS 3.43 vs 3.06 ms and PageRank 152 vs 122 ms on RMAT-24); parts of
    // the heavy lists
    std::vector<VertexId> lv, pv;
    std::vector<EdgeId> pb;
    for (VertexId v = 0; v < vertices; v++) {
      if (hd[v] > (VertexId)gg_facade::kLightMax) {
        for (EdgeId b = hidx[v]; b < hidx[v + 1]; b += gg_facade::kPart) {
          pv.push_back(v);
          pb.push_back(b);
        }
      } else if (hd[v]) {
        lv.push_back(v);
      }
    }
    d.n_light = (VertexId)lv.size();
    d.n_parts = (unsigned)pv.size();
    auto upload = [&](void **dst, const void *src, size_t bytes, const char *what) {
      gg_facade::hip_check(hipMalloc(dst, bytes ? bytes : 8), what);
      owned_.push_back(*dst);
      if (bytes) gg_facade::hip_check(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
    };
    upload((void **)&d.light, lv.data(), sizeof(VertexId) * lv.size(), "light vertices");
    upload((void **)&d.part_v, pv.data(), sizeof(VertexId) * pv.size(), "parts");
    upload((void **)&d.part_b, pb.data(), sizeof(EdgeId) * pb.size(), "parts");
    gg_facade::hip_check(hipDeviceSynchronize(), "make_dir");
    return d;
  }

  // Message buffers for a call with messages of type M: at most one message
  // per vertex (sparse) or per light vertex and heavy-list part (dense);
  // segments hold 1.5 times that spread over kEmitSegs, the overflow region
  // that again.  Grown on the host only when M needs more bytes than any
  // earlier call.
  template <typename M>
  void ensure_messages() {
    const unsigned long long cap = (unsigned long long)vertices + in_.n_parts + 1;
    const unsigned long long seg_cap = (cap + cap / 2) / gg_facade::kEmitSegs + 64;
    const unsigned long long total = seg_cap * gg_facade::kEmitSegs + cap;
    const size_t bytes = sizeof(gg_facade::MsgUnit<M>) * (size_t)total;
    if (bytes > msg_cap_bytes_) {
      gg_facade::hip_check(hipDeviceSynchronize(), "message buffer");
      (void)hipFree(msg_buf_);
      (void)hipFree(seg_buf_);
      gg_facade::hip_check(hipMalloc(&msg_buf_, bytes), "message buffer");
      gg_facade::hip_check(hipMalloc(&seg_buf_, bytes), "message segments");
      msg_cap_bytes_ = bytes;
      hipLaunchKernelGGL(gg_facade::set_buf_kernel<0>, dim3(1), dim3(1), 0, 0, ctl_, msg_buf_,
                         seg_buf_, seg_cnt_, seg_cap, cap);
    }
  }
  template <typename R>
  R read_result() {
    unsigned long long bits = 0;
    gg_facade::hip_check(hipMemcpy(&bits, &ctl_->result, 8, hipMemcpyDeviceToHost), "reducer");
    R r;
    __builtin_memcpy(&r, &bits, sizeof(R));
    return r;
  }
};

#endif

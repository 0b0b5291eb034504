// The five toolkits on the device engine: PageRank, BFS, CC, SSSP, BC.
//
// Each driver mirrors the compute() of its reference toolkit: the same
// initialisation, the same per-iteration process_edges / process_vertices
// sequence, the same push/pull switch (core/graph.hpp:1503-1509) and the
// same BSP visibility: a signal reads state as it stood when the iteration
// began.  CC and SSSP read and write one array, so their pull runs in two
// phases (gather into a message array, then apply) and their push reads a
// snapshot of the frontier's values; BFS, BC and PageRank are hazard-free and
// run fused.
#include <string.h>

#include <algorithm>

#include "kernels.hpp"

namespace gg {

// --------------------------------------------------------------- utilities
template <class T>
__global__ void fill_kernel(T *p, T v, uint64_t n) {
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s) p[i] = v;
}
__global__ void iota_kernel(uint32_t *p, uint32_t n) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s) p[i] = i;
}
__global__ void popcount_kernel(const uint64_t *bits, size_t words, unsigned long long *out) {
  size_t s = (size_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += s)
    c += (uint64_t)__popcll(bits[i]);
  c = wave_sum_u64(c);
  if (lane_id() == 0 && c) atomicAdd(out, (unsigned long long)c);
}

__global__ __launch_bounds__(kBlock) void compact_kernel(const uint64_t *bits, size_t words,
                                                         const uint32_t *bin_deg,
                                                         uint32_t small_max, uint32_t big_max,
                                                         uint32_t *q0, uint32_t *q1, uint32_t *q2,
                                                         Counters *cnt, uint64_t *visited,
                                                         uint64_t *clear) {
  compact_words(bits, words, bin_deg, small_max, big_max, q0, q1, q2, cnt, visited, clear);
}

// The counts of a frontier without its queues: one lane per vertex (one
// wave per bitmap word, degrees read coalesced), per-block sums added into
// one of kCountSlots copies (atomics on one copy from every block would
// serialise), summed on the host; the same side effects as the compaction
// (visited OR, the other bitmap cleared).
constexpr uint32_t kCountSlots = 256;
__global__ __launch_bounds__(kBlock) void count_kernel(const uint64_t *bits, size_t words,
                                                       uint32_t V, const uint32_t *bin_deg,
                                                       uint32_t small_max, uint32_t big_max,
                                                       Counters *slots, uint64_t *visited,
                                                       uint64_t *clear) {
  Counters *cnt = slots + blockIdx.x % kCountSlots;
  __shared__ uint32_t s_n[3][kWavesPerBlock];
  __shared__ unsigned long long s_e[kWavesPerBlock];
  const int lane = lane_id(), wv = wave_in_block();
  const size_t W = (size_t)gridDim.x * kWavesPerBlock;
  uint32_t c0 = 0, c1 = 0, c2 = 0;
  uint64_t es = 0;
  for (size_t wi = global_wave(); wi < words; wi += W) {
    const uint64_t word = bits[wi];
    if (lane == 0) {
      if (visited && word) visited[wi] |= word;
      if (clear) clear[wi] = 0;
    }
    const uint32_t v = (uint32_t)(wi * 64 + lane);
    if (((word >> lane) & 1ULL) && v < V) {
      const uint32_t d = bin_deg[v];
      es += d;
      c0 += d <= small_max;
      c1 += d > small_max && d <= big_max;
      c2 += d > big_max;
    }
  }
  const uint64_t w_e = wave_sum_u64(es);
  const uint32_t w0 = (uint32_t)wave_sum_u64(c0), w1 = (uint32_t)wave_sum_u64(c1),
                 w2 = (uint32_t)wave_sum_u64(c2);
  if (lane == 0) {
    s_n[0][wv] = w0;
    s_n[1][wv] = w1;
    s_n[2][wv] = w2;
    s_e[wv] = w_e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t b0 = 0, b1 = 0, b2 = 0;
    unsigned long long be = 0;
    for (int i = 0; i < kWavesPerBlock; i++) {
      b0 += s_n[0][i];
      b1 += s_n[1][i];
      b2 += s_n[2][i];
      be += s_e[i];
    }
    if (b0 | b1 | b2) {
      if (b0) atomicAdd(&cnt->n[0], b0);
      if (b1) atomicAdd(&cnt->n[1], b1);
      if (b2) atomicAdd(&cnt->n[2], b2);
      atomicAdd(&cnt->total, b0 + b1 + b2);
      atomicAdd((unsigned long long *)&cnt->edges, be);
    }
  }
}

// Blocks for kernels that walk vertices in 256-aligned rounds.
static unsigned grid_vertex_rounds(uint32_t V) { return grid1(V, 1u << 14); }

template <class T>
static void fill(T *p, T v, uint64_t n) {
  if (n) hipLaunchKernelGGL(fill_kernel<T>, dim3(grid1(n)), dim3(kBlock), 0, stream(), p, v, n);
}

// Sets bits for one vertex per lane in 256-vertex rounds; each wave owns one
// word per round, so one OR per wave suffices.
__device__ __forceinline__ void set_owned_bit(uint64_t *bits, uint32_t v, bool on) {
  uint64_t m = __ballot(on);
  if (lane_id() == 0 && m) atomicOr((unsigned long long *)&bits[v >> 6], (unsigned long long)m);
}

__global__ void set_root_kernel(uint64_t *bits, uint32_t root) {
  bits[root >> 6] |= 1ULL << (root & 63);
}
template <class A, class B>
__global__ void put2_kernel(A *pa, A va, B *pb, B vb, uint32_t i) {
  pa[i] = va;
  if (pb) pb[i] = vb;
}
// Bitmap::fill (core/bitmap.hpp:42-52): full words, exact tail, rest zero.
__global__ void bitmap_fill_kernel(uint64_t *bits, size_t words, uint32_t V) {
  size_t full = (size_t)V >> 6;
  size_t s = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += s) {
    uint64_t x = 0;
    if (i < full)
      x = ~0ULL;
    else if (i == full && (V & 63))
      x = (1ULL << (V & 63)) - 1ULL;
    bits[i] = x;
  }
}

// Boundary of reordered graphs: results computed in internal order are
// scattered back to original ids (vertex-valued results mapped through iperm).
template <class T>
__global__ void unpermute_kernel(const T *in, T *out, const uint32_t *iperm, uint32_t V) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < V; n += s) out[iperm[n]] = in[n];
}
__global__ void unpermute_ids_kernel(const uint32_t *in, uint32_t *out, const uint32_t *iperm,
                                     uint32_t V, uint32_t sentinel) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < V; n += s) {
    uint32_t p = in[n];
    out[iperm[n]] = p == sentinel ? sentinel : iperm[p];
  }
}
__global__ void copy_u32_kernel(const uint32_t *in, uint32_t *out, uint32_t V) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < V; n += s) out[n] = in[n];
}

// --------------------------------------------------------------- work area
int work_get(gg_graph *g, Work **out) {
  const int slot = context_slot();
  if (g->works[slot]) {
    *out = g->works[slot];
    return GG_OK;
  }
  Work *w = new Work;
  w->V = g->vertices;
  w->words = ((size_t)g->vertices >> 6) + 1;
  size_t V = (size_t)g->vertices + 1;
  auto al = [&](void **p, size_t bytes) -> int {
    int r = device_alloc(p, bytes);
    if (r == GG_OK) {
      w->allocations.push_back(*p);
      w->bytes += bytes;
    }
    return r;
  };
  int r = GG_OK;
  for (int i = 0; i < 2 && r == GG_OK; i++) {
    Frontier &f = w->f[i];
    r = al((void **)&f.bits, w->words * 8);
    for (int b = 0; b < 3 && r == GG_OK; b++) r = al((void **)&f.q[b], V * 4);
    for (int b = 0; b < 3 && r == GG_OK; b++) r = al((void **)&f.snap_f[b], V * 4);
    if (r == GG_OK) r = al((void **)&f.cnt, sizeof(Counters));
  }
  if (r == GG_OK) r = al((void **)&w->visited, w->words * 8);
  if (r == GG_OK) r = al((void **)&w->aux_bits, w->words * 8);
  if (r == GG_OK) r = al(&w->msg, V * 8);
  if (r == GG_OK && g->reordered) r = al(&w->res_a, V * 8);
  if (r == GG_OK && g->reordered) r = al(&w->res_b, V * 8);
  size_t nh = std::max<size_t>(std::max(g->out.n_heavy, g->in.n_heavy), 1);
  if (r == GG_OK) r = al((void **)&w->hacc, nh * 8);
  if (r == GG_OK) r = al((void **)&w->exam, 1024 * 8);
  if (r == GG_OK) r = al((void **)&w->cnt_slots, kCountSlots * sizeof(Counters));
  if (r == GG_OK && hipHostMalloc((void **)&w->h_cnt, sizeof(Counters) * 2) != hipSuccess)
    r = fail(GG_ENOMEM, "pinned counters");
  if (r == GG_OK &&
      hipHostMalloc((void **)&w->h_cnt_slots, kCountSlots * sizeof(Counters)) != hipSuccess)
    r = fail(GG_ENOMEM, "pinned counters");
  if (r == GG_OK && hipHostMalloc((void **)&w->h_exam, 1024 * 8) != hipSuccess)
    r = fail(GG_ENOMEM, "pinned counters");
  if (r != GG_OK) {
    work_free(w);
    return r;
  }
  hipMemsetAsync(w->hacc, 0, nh * 8, stream());
  hipMemsetAsync(w->exam, 0, 1024 * 8, stream());
  g->works[slot] = w;
  *out = w;
  return GG_OK;
}

void work_free(Work *w) {
  if (!w) return;
  for (void *p : w->allocations) device_free(p);
  device_free(w->dacc);
  device_free(w->dsum);
  device_free(w->ms_vis);
  device_free(w->ms_lmask);
  device_free(w->pr_vals);
  device_free(w->pr_gsum);
  device_free(w->hpart);
  bfs_state_free(w->bfs);
  device_free(w->ms_f[0]);
  device_free(w->ms_f[1]);
  device_free(w->ms_roots);
  device_free(w->ms_tmp);
  device_free(w->mb_a);
  device_free(w->mb_b);
  device_free(w->ms_levels);
  msd_state_free(w->msd);
  if (w->h_cnt) hipHostFree(w->h_cnt);
  if (w->h_cnt_slots) hipHostFree(w->h_cnt_slots);
  if (w->h_exam) hipHostFree(w->h_exam);
  delete w;
}

// Host-side engine state for one traversal.
struct Engine {
  gg_graph *g;
  Work *w;
  int cur = 0;    // w->f[cur] is the input frontier
  Counters in{};  // host copy of the input frontier's counters
  gg_run_stats st{};
  uint32_t small_max, big_max;
  unsigned long long *exam;
  // true: a frontier is counted first and compacted into queues only when
  // the next iteration pushes (the loop reads the queues in push mode only)
  bool lazy_queues = false;
  uint64_t divisor;  // the switch's divisor (option switch_divisor unless set)

  Engine(gg_graph *g_, Work *w_) : g(g_), w(w_) {
    small_max = (uint32_t)options().push_small_max;
    big_max = (uint32_t)options().push_big_max;
    divisor = (uint64_t)options().switch_divisor;
    exam = options().count_examined ? w->exam : nullptr;
  }
  Frontier &fin() { return w->f[cur]; }
  Frontier &fout() { return w->f[cur ^ 1]; }
  int clear_bits(Frontier &f) {
    GG_HIP(hipMemsetAsync(f.bits, 0, w->words * 8, stream()));
    return GG_OK;
  }
  // Bitmap -> binned queues and counts; reads the counts back.  The other
  // frontier's bitmap is cleared in the same pass: it is the next output.
  int compact(Frontier &f, const uint32_t *bin_deg, uint64_t *visited_or, Counters *outc) {
    Frontier &other = (&f == &w->f[0]) ? w->f[1] : w->f[0];
    if (lazy_queues) {
      GG_HIP(hipMemsetAsync(w->cnt_slots, 0, kCountSlots * sizeof(Counters), stream()));
      {
        ProfScope p("count");
        hipLaunchKernelGGL(count_kernel, dim3(grid_waves(std::min<size_t>(w->words, 1u << 16))),
                           dim3(kBlock), 0, stream(), f.bits, w->words, g->vertices, bin_deg,
                           small_max, big_max, w->cnt_slots, visited_or, other.bits);
      }
      GG_HIP(hipMemcpyAsync(w->h_cnt_slots, w->cnt_slots, kCountSlots * sizeof(Counters),
                            hipMemcpyDeviceToHost, stream()));
      GG_HIP(hipStreamSynchronize(stream()));
      GG_HIP(hipGetLastError());
      Counters c{};
      for (uint32_t k = 0; k < kCountSlots; k++) {
        const Counters &x = w->h_cnt_slots[k];
        c.n[0] += x.n[0];
        c.n[1] += x.n[1];
        c.n[2] += x.n[2];
        c.total += x.total;
        c.edges += x.edges;
      }
      *outc = c;
      if (!sparse(outc->edges)) return GG_OK;  // a pull comes next: no queues
      visited_or = nullptr;                    // (done by the count pass)
    }
    GG_HIP(hipMemsetAsync(f.cnt, 0, sizeof(Counters), stream()));
    {
      ProfScope p("compact");
      hipLaunchKernelGGL(compact_kernel, dim3(grid1(w->words, 2048)), dim3(kBlock), 0, stream(),
                         f.bits, w->words, bin_deg, small_max, big_max, f.q[0], f.q[1], f.q[2],
                         f.cnt, visited_or, other.bits);
    }
    GG_HIP(hipMemcpyAsync(w->h_cnt, f.cnt, sizeof(Counters), hipMemcpyDeviceToHost, stream()));
    GG_HIP(hipStreamSynchronize(stream()));
    GG_HIP(hipGetLastError());
    *outc = *w->h_cnt;
    return GG_OK;
  }
  int compact_in(const uint32_t *bin_deg, uint64_t *visited_or = nullptr) {
    return compact(fin(), bin_deg, visited_or, &in);
  }
  int seed(uint32_t root, const uint32_t *bin_deg, uint64_t *visited_or) {
    GG_TRY(clear_bits(fin()));
    hipLaunchKernelGGL(set_root_kernel, dim3(1), dim3(1), 0, stream(), fin().bits, root);
    return compact_in(bin_deg, visited_or);
  }
  // The reference switch: sparse iff active out-edges < |E| / 20.
  bool sparse(uint64_t active_edges) const {
    int64_t fm = options().force_mode;
    if (fm == 1) return true;
    if (fm == 2) return false;
    return active_edges < g->edges / divisor;
  }
  int note_iteration(bool was_sparse) {
    int i = st.iterations;
    if (i < GG_STATS_RECORDS) {
      st.modes[i] = was_sparse ? 0u : 1u;
      st.it_active[i] = in.total;
    }
    st.iterations++;
    if (was_sparse) st.push_iterations++;
    if (exam) {
      GG_HIP(hipMemcpyAsync(w->h_exam, exam, 1024 * 8, hipMemcpyDeviceToHost, stream()));
      GG_HIP(hipMemsetAsync(exam, 0, 1024 * 8, stream()));
      GG_HIP(hipStreamSynchronize(stream()));
      uint64_t x = 0;
      for (int k = 0; k < 64; k++) x += w->h_exam[k * 16];
      st.edges_examined += x;
      if (i < GG_STATS_RECORDS) st.it_examined[i] = x;
    }
    return GG_OK;
  }
  // After the edge kernels: compact the output frontier and swap.
  int finish_iteration(bool was_sparse, const uint32_t *bin_deg, uint64_t *visited_or) {
    GG_TRY(note_iteration(was_sparse));
    Counters c;
    GG_TRY(compact(fout(), bin_deg, visited_or, &c));
    if (st.iterations - 1 < GG_STATS_RECORDS) st.it_new[st.iterations - 1] = c.total;
    cur ^= 1;
    in = c;
    return GG_OK;
  }
  PushArgs push_args(const Dir &d, bool snaps) {
    Frontier &f = fin();
    PushArgs a;
    a.index = d.index;
    a.nbr = d.nbr;
    a.w = d.w;
    a.q0 = f.q[0];
    a.q1 = f.q[1];
    a.q2 = f.q[2];
    a.s0 = snaps ? (const uint32_t *)f.snap_f[0] : nullptr;
    a.s1 = snaps ? (const uint32_t *)f.snap_f[1] : nullptr;
    a.s2 = snaps ? (const uint32_t *)f.snap_f[2] : nullptr;
    a.n0 = in.n[0];
    a.n1 = in.n[1];
    a.n2 = in.n[2];
    a.waves0 = (in.n[0] + 63) / 64;
    a.waves1 = in.n[1];
    a.waves2 = in.n[2] ? (uint32_t)options().huge_grid_waves : 0;
    a.out = fout().bits;
    a.exam = exam;
    return a;
  }
  template <class Op>
  int launch_push(const Dir &d, const Op &op, bool snaps) {
    PushArgs a = push_args(d, snaps);
    uint64_t waves = (uint64_t)a.waves0 + a.waves1 + a.waves2;
    if (!waves) return GG_OK;
    ProfScope p("push");
    hipLaunchKernelGGL(push_kernel<Op>, dim3(grid_waves(waves)), dim3(kBlock), 0, stream(), a, op);
    return GG_OK;
  }
  template <class Op>
  int launch_pull_group(const Dir &d, const Op &op, double *dacc, const char *name = "pull",
                        const uint32_t *count = nullptr) {
    PullArgs a;
    a.count = count;
    a.index = d.index;
    a.nbr = d.nbr;
    a.w = d.w;
    a.V = g->vertices;
    a.light_max = d.light_max;
    a.ngroups = (g->vertices + 63) / 64;
    a.out = fout().bits;
    a.exam = exam;
    a.dacc = dacc;
    ProfScope p(name);
    hipLaunchKernelGGL(pull_group_kernel<Op>, dim3(grid_waves(a.ngroups)), dim3(kBlock), 0,
                       stream(), a, op);
    return GG_OK;
  }
  template <class Op>
  int launch_pull_heavy(const Dir &d, const Op &op, const char *name = "pull_heavy") {
    if (!d.n_chunks) return GG_OK;
    if (w->hpart_cap < d.n_chunks) {
      device_free(w->hpart);
      w->hpart = nullptr;
      w->hpart_cap = 0;
      GG_TRY(device_alloc(&w->hpart, (size_t)d.n_chunks * 8));
      w->hpart_cap = d.n_chunks;
    }
    HeavyArgs a{d.nbr, d.w, d.chunk_v, d.chunk_h, d.chunk_beg, d.chunk_end, d.n_chunks,
                fout().bits, exam, w->hpart};
    ProfScope p(name);
    const uint64_t hw = options().heavy_waves;
    hipLaunchKernelGGL(pull_heavy_kernel<Op>,
                       dim3(grid_waves(hw ? std::min<uint64_t>(d.n_chunks, hw) : d.n_chunks)),
                       dim3(kBlock), 0, stream(), a, op);
    hipLaunchKernelGGL(heavy_reduce_kernel<Op>, dim3(grid_waves(d.n_heavy)), dim3(kBlock), 0,
                       stream(), d.heavy_v, d.heavy_first, d.n_heavy, w->hpart, fout().bits, op);
    return GG_OK;
  }
};

struct PrHeavy : NoAfter {
  using T = double;
  const double *curr;
  double *hacc;
  __device__ double identity() const { return 0.0; }
  __device__ bool dst_active(uint32_t) const { return true; }
  __device__ double value(uint32_t src, float) const { return curr[src]; }
  __device__ double combine(double a, double b) const { return a + b; }
  __device__ bool commit(uint32_t, uint32_t h, double x) const {
    unsafeAtomicAdd(&hacc[h], x);
    return false;
  }
};
__global__ void pr_heavy_finish_kernel(const uint32_t *heavy_v, uint32_t n, double *hacc, PrPull op,
                                       double *dacc) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double dsum = 0;
  if (i < n) {
    uint32_t v = heavy_v[i];
    double s = hacc[i];
    hacc[i] = 0.0;
    op.finish(v, s, 0, dsum);
  }
  double t = wave_sum_f64(dsum);
  if (lane_id() == 0 && t != 0.0) unsafeAtomicAdd(&dacc[((i >> 6) % kScalarSlots) * 16], t);
}
// One block per spread accumulator: sums its kScalarSlots partials.
__global__ void scalar_reduce_kernel(const double *dacc, double *out) {
  __shared__ double part[kWavesPerBlock];
  const double *a = dacc + (size_t)blockIdx.x * kScalarStride;
  double x = 0;
  for (uint32_t s = threadIdx.x; s < kScalarSlots; s += blockDim.x) x += a[s * 16];
  x = wave_sum_f64(x);
  if (lane_id() == 0) part[wave_in_block()] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < kWavesPerBlock; i++) t += part[i];
    out[blockIdx.x] = t;
  }
}
__global__ void pr_init_kernel(double *curr, const uint32_t *outdeg, uint32_t V) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s) {
    double x = 1.0;  // pagerank.cpp:36-42
    if (outdeg[v] > 0) x /= outdeg[v];
    curr[v] = x;
  }
}

// BFS (toolkits/bfs.cpp) has a device-driven loop of its own: bfs.hip.


// ====================================================================== CC
// toolkits/cc.cpp:22-109
struct CcPush : NoAfter {
  using Msg = uint32_t;
  uint32_t *label;
  __device__ uint32_t msg(uint32_t, uint32_t snap) const { return snap; }
  __device__ bool edge(uint32_t, uint32_t m, uint32_t dst, float) const {
    // cc.cpp:52-56: a plain compare, then write_min; the bit is set whenever
    // the compare succeeds.
    if (m < label[dst]) {
      atomicMin(&label[dst], m);
      return true;
    }
    return false;
  }
};
struct CcPull : NoAfter {  // phase 1 of the dense step: gather the minimum in-label
  using T = uint32_t;
  const uint32_t *label;
  uint32_t *msg;
  __device__ uint32_t identity() const { return 0xFFFFFFFFu; }
  __device__ bool dst_active(uint32_t) const { return true; }
  __device__ uint32_t value(uint32_t src, float) const { return label[src]; }
  __device__ uint32_t combine(uint32_t a, uint32_t b) const { return a < b ? a : b; }
  __device__ bool finish(uint32_t v, uint32_t acc, uint32_t deg, double &) const {
    if (deg) msg[v] = acc;
    return false;
  }
  __device__ bool commit(uint32_t v, uint32_t, uint32_t x) const {
    atomicMin(&msg[v], x);
    return false;
  }
};
// The distances at the start of a dense SSSP step on the plan (f32 bits).
__global__ void sssp_snapshot_kernel(const float *dist, uint32_t *snap, uint32_t V) {
  const uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s)
    snap[v] = __float_as_uint(dist[v]);
}
// The labels at the start of a dense CC step on the plan, complemented.
__global__ void cc_snapshot_kernel(const uint32_t *label, uint32_t *nlab, uint32_t V) {
  const uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s)
    nlab[v] = ~label[v];
}
// phase 2: dense_slot (cc.cpp:72-79), one lane per vertex.
__global__ void cc_apply_kernel(uint32_t *label, uint32_t *msg, uint32_t V, uint64_t *out) {
  uint32_t n = (V + 255) & ~255u;
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += s) {
    uint32_t v = base + threadIdx.x;
    bool cand = false;
    if (v < V) {
      uint32_t m = msg[v];
      if (m != 0xFFFFFFFFu) {
        msg[v] = 0xFFFFFFFFu;
        if (m < label[v]) {
          label[v] = m;
          cand = true;
        }
      }
    }
    set_owned_bit(out, v, cand);
  }
}

// ==================================================================== SSSP
// toolkits/sssp.cpp:22-105.  Distances are non-negative, so the IEEE bit
// patterns order like the values and write_min is an integer atomicMin.
constexpr uint32_t kInfBits = 0x4E6E6B28u;  // 1e9f
struct SsspPush : NoAfter {
  using Msg = float;
  float *dist;
  __device__ float msg(uint32_t, uint32_t snap) const { return __uint_as_float(snap); }
  __device__ bool edge(uint32_t, float m, uint32_t dst, float w) const {
    float relax = m + w;
    if (relax < dist[dst]) {
      uint32_t old = atomicMin((uint32_t *)&dist[dst], __float_as_uint(relax));
      return relax < __uint_as_float(old);
    }
    return false;
  }
};
struct SsspPull : NoAfter {
  using T = float;
  const float *dist;
  uint32_t *msg;
  __device__ float identity() const { return 1e9f; }
  __device__ bool dst_active(uint32_t) const { return true; }
  __device__ float value(uint32_t src, float w) const { return dist[src] + w; }
  __device__ float combine(float a, float b) const { return b < a ? b : a; }
  __device__ bool finish(uint32_t v, float acc, uint32_t deg, double &) const {
    if (deg) msg[v] = __float_as_uint(acc);
    return false;
  }
  __device__ bool commit(uint32_t v, uint32_t, float x) const {
    atomicMin(&msg[v], __float_as_uint(x));
    return false;
  }
};
__global__ void sssp_apply_kernel(float *dist, uint32_t *msg, uint32_t V, uint64_t *out) {
  uint32_t n = (V + 255) & ~255u;
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += s) {
    uint32_t v = base + threadIdx.x;
    bool cand = false;
    if (v < V) {
      uint32_t mb = msg[v];
      if (mb != kInfBits) {
        msg[v] = kInfBits;
        float m = __uint_as_float(mb);
        if (m < 1e9f && m < dist[v]) {  // sssp.cpp:71, :74-77
          dist[v] = m;
          cand = true;
        }
      }
    }
    set_owned_bit(out, v, cand);
  }
}
// Message snapshot of the input frontier's queues (the signal phase).
__global__ void snapshot_kernel(const uint32_t *vals, Frontier f, uint32_t n0, uint32_t n1,
                                uint32_t n2) {
  uint64_t n = (uint64_t)n0 + n1 + n2;
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s) {
    int b = i < n0 ? 0 : (i < (uint64_t)n0 + n1 ? 1 : 2);
    uint64_t j = b == 0 ? i : (b == 1 ? i - n0 : i - n0 - n1);
    ((uint32_t *)f.snap_f[b])[j] = vals[f.q[b][j]];
  }
}

// ====================================================================== BC
// toolkits/bc.cpp:24-196
struct BcFwdPush {
  using Msg = double;
  double *sigma;
  const uint64_t *visited;
  uint32_t *depth;
  uint32_t level;
  __device__ double msg(uint32_t v, uint32_t) const { return sigma[v]; }
  __device__ bool edge(uint32_t, double m, uint32_t dst, float) const {
    if (!test_bit(visited, dst)) {  // bc.cpp:61-66
      unsafeAtomicAdd(&sigma[dst], m);
      return true;
    }
    return false;
  }
  __device__ void after(bool fresh, uint32_t v) const {
    if (fresh) depth[v] = level;
  }
};
struct BcFwdPull {
  using T = double;
  double *sigma;
  const uint64_t *visited, *fin;
  uint32_t *depth;
  uint32_t level;
  __device__ double identity() const { return 0.0; }
  __device__ bool dst_active(uint32_t v) const { return !test_bit(visited, v); }
  __device__ double value(uint32_t src, float) const {
    return test_bit(fin, src) ? sigma[src] : 0.0;
  }
  __device__ double combine(double a, double b) const { return a + b; }
  __device__ bool finish(uint32_t v, double sum, uint32_t, double &) const {
    if (sum > 0) {  // bc.cpp:77-86
      sigma[v] += sum;
      return true;
    }
    return false;
  }
  __device__ bool commit(uint32_t v, uint32_t, double x) const {
    if (x > 0) {
      unsafeAtomicAdd(&sigma[v], x);
      return true;
    }
    return false;
  }
  __device__ void after(bool fresh, uint32_t v) const {
    if (fresh) depth[v] = level;
  }
};
// backward step from level L into level L-1 over the transposed graph
// (bc.cpp:125-157).  Only vertices at depth L-1 can receive a non-zero sum;
// the reference also adds into unreached vertices, which end as NaN either way.
struct BcBwdPush : NoAfter {
  using Msg = double;
  double *dep;
  const uint32_t *depth;
  uint32_t target;  // L-1
  __device__ double msg(uint32_t v, uint32_t) const { return dep[v]; }
  __device__ bool edge(uint32_t, double m, uint32_t dst, float) const {
    if (depth[dst] == target) unsafeAtomicAdd(&dep[dst], m);
    return false;
  }
};
struct BcBwdPull : NoAfter {
  using T = double;
  double *dep;
  const uint32_t *depth;
  const uint64_t *lvl;  // bitmap of level L
  uint32_t target;
  __device__ double identity() const { return 0.0; }
  __device__ bool dst_active(uint32_t v) const { return depth[v] == target; }
  __device__ double value(uint32_t src, float) const { return test_bit(lvl, src) ? dep[src] : 0.0; }
  __device__ double combine(double a, double b) const { return a + b; }
  __device__ bool finish(uint32_t v, double sum, uint32_t, double &) const {
    dep[v] += sum;
    return false;
  }
  __device__ bool commit(uint32_t v, uint32_t, double x) const {
    unsafeAtomicAdd(&dep[v], x);
    return false;
  }
};
// Level-L subset by depth, with dependencies[v] += 1/num_paths[v]
// (bc.cpp:158-167 applied to levels.back()).
__global__ void bc_level_kernel(const uint32_t *depth, double *dep, const double *inv, uint32_t V,
                                uint32_t L, uint64_t *out) {
  uint32_t n = (V + 255) & ~255u;
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += s) {
    uint32_t v = base + threadIdx.x;
    bool on = v < V && depth[v] == L;
    if (on) dep[v] += inv[v];
    set_owned_bit(out, v, on);
  }
}
__global__ void bc_inv_kernel(double *paths, double *dep, uint32_t V) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s) {
    paths[v] = 1 / paths[v];  // bc.cpp:103-111, +inf when unreached
    dep[v] = 0;
  }
}
__global__ void bc_final_kernel(double *dep, const double *inv, uint32_t V) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s)
    dep[v] = (dep[v] - inv[v]) / inv[v];  // bc.cpp:170-176
}

// ========================================================= batched BFS (64)
__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, int j) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, j);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), j);
  return ((uint64_t)hi << 32) | lo;
}

// Bit-parallel multi-source BFS: vis[v] and f[v] are 64-bit masks, bit i for
// root i.  A level is the plain BFS level of every root at once: push ORs a
// source's frontier mask into its out-neighbours, pull ORs the in-neighbours'
// masks; `apply` then keeps the bits not yet visited and records depths.
struct MsPush : NoAfter {
  using Msg = uint64_t;
  const uint64_t *fcur, *vis;
  uint64_t *fnext;
  __device__ uint64_t msg(uint32_t v, uint32_t) const { return fcur[v]; }
  __device__ bool edge(uint32_t, uint64_t m, uint32_t dst, float) const {
    uint64_t f = m & ~vis[dst];
    if (!f) return false;
    if ((fnext[dst] & f) != f) atomicOr((unsigned long long *)&fnext[dst], (unsigned long long)f);
    return true;
  }
};
struct MsPull : NoAfter {
  using T = uint64_t;
  const uint64_t *fcur, *vis;
  uint64_t *fnext;
  uint64_t full;
  __device__ uint64_t identity() const { return 0ULL; }
  __device__ bool dst_active(uint32_t v) const { return (vis[v] & full) != full; }
  __device__ uint64_t value(uint32_t src, float) const { return fcur[src]; }
  __device__ uint64_t combine(uint64_t a, uint64_t b) const { return a | b; }
  __device__ bool finish(uint32_t v, uint64_t acc, uint32_t, double &) const {
    uint64_t nw = acc & ~vis[v] & full;
    if (nw) fnext[v] = nw;
    return nw != 0;
  }
  __device__ bool commit(uint32_t v, uint32_t, uint64_t x) const {
    x &= ~vis[v] & full;
    if (!x) return false;
    atomicOr((unsigned long long *)&fnext[v], (unsigned long long)x);
    return true;
  }
};
// For every vertex activated this level: keep the new bits, mark them
// visited, write their depths (vertex-major, 64 entries per vertex).
// Dense multi-root step with early exit: a vertex stops scanning its in-list
// once every root still missing from it has been found (the bottom-up BFS
// break of bfs.cpp:66-67, per root mask).  Short lists: one lane each; longer
// ones: the whole wave, 64 entries at a time.  Heavy lists go to the chunked
// kernel.
struct MsPullArgs {
  const uint64_t *index;
  const uint32_t *nbr;
  const uint64_t *fcur, *vis;
  uint64_t *fnext, *out;
  uint64_t full;
  uint32_t V, light_max, serial_max, ngroups;
  const uint32_t *gate;  // null, or: run only when *gate == kGateDense
};
__global__ __launch_bounds__(kBlock) void ms_pull_group_kernel(MsPullArgs a) {
  const uint32_t g = global_wave();
  const int lane = lane_id();
  if (g >= a.ngroups) return;
  if (a.gate && *a.gate != kGateDense) return;
  uint32_t v = g * 64 + lane;
  uint64_t need = 0, b = 0, e1 = 0;
  if (v < a.V) {
    need = a.full & ~a.vis[v];
    if (need) {
      b = a.index[v];
      e1 = a.index[v + 1];
    }
  }
  uint64_t deg = e1 - b;
  bool cand = need && deg > 0 && deg <= a.light_max;
  bool serial = cand && deg <= a.serial_max;
  uint64_t acc = 0;
  if (serial) {
    // four neighbours and their masks in flight per step
    for (uint64_t e = b; e < e1; e += 4) {
      uint32_t u[4];
#pragma unroll
      for (int t = 0; t < 4; t++) u[t] = e + t < e1 ? stream_load(&a.nbr[e + t]) : 0;
#pragma unroll
      for (int t = 0; t < 4; t++) acc |= e + t < e1 ? a.fcur[u[t]] : 0ULL;
      if ((acc & need) == need) break;
    }
  }
  uint64_t mmask = __ballot(cand && !serial);
  while (mmask) {
    int k = __builtin_ctzll(mmask);
    mmask &= mmask - 1;
    uint64_t bk = shfl_u64(b, k), ek = shfl_u64(e1, k), nk = shfl_u64(need, k);
    uint64_t ak = 0;
    for (uint64_t e0 = bk; e0 < ek; e0 += 64) {
      uint64_t e = e0 + lane;
      uint64_t x = e < ek ? a.fcur[stream_load(&a.nbr[e])] : 0ULL;
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) x |= shfl_u64(x, lane ^ m);
      ak |= x;
      if ((ak & nk) == nk) break;
    }
    if (lane == k) acc = ak;
  }
  uint64_t nw = acc & need;
  if (nw) a.fnext[v] = nw;
  uint64_t m = __ballot(nw != 0);
  if (lane == 0 && m) atomicOr((unsigned long long *)&a.out[g], (unsigned long long)m);
}

// One wave per 64-vertex bitmap word: lane k settles vertex 64w+k's masks,
// then the wave writes each activated vertex's depth row with lane = root
// (one coalesced 256-byte masked store per vertex).
__global__ __launch_bounds__(kBlock) void ms_apply_kernel(const uint64_t *active, size_t words,
                                                          uint64_t *vis, uint64_t *fnext,
                                                          uint32_t *depth, uint32_t level) {
  const int lane = lane_id();
  const size_t nw_waves = (size_t)gridDim.x * kWavesPerBlock;
  for (size_t wi = global_wave(); wi < words; wi += nw_waves) {
    uint64_t act = active[wi];
    if (!act) continue;
    uint32_t v = (uint32_t)(wi * 64 + lane);
    uint64_t nw = 0;
    if ((act >> lane) & 1ULL) {
      nw = fnext[v] & ~vis[v];
      fnext[v] = nw;
      vis[v] |= nw;
    }
    if (!depth) continue;  // depths come from the per-level masks
    for (uint64_t x = act; x; x &= x - 1) {
      int k = __builtin_ctzll(x);
      uint64_t m = shfl_u64(nw, k);
      if ((m >> lane) & 1ULL) depth[(size_t)(wi * 64 + k) * 64 + lane] = level;
    }
  }
}
__global__ void ms_seed_kernel(const uint32_t *roots, int n, uint64_t *vis, uint64_t *f,
                               uint64_t *bits, uint32_t *depth) {
  const int i = threadIdx.x;  // one thread per root; roots may repeat
  if (i >= n) return;
  const uint32_t r = roots[i];
  atomicOr((unsigned long long *)&vis[r], 1ULL << i);
  atomicOr((unsigned long long *)&f[r], 1ULL << i);
  atomicOr((unsigned long long *)&bits[r >> 6], 1ULL << (r & 63));
  if (depth) depth[(size_t)r * 64 + i] = 0;
}

// Depth rows from the per-level root masks, one wave per 64 vertices: the
// lanes load the 64 vertices' mask of each level (coalesced), then for each
// vertex lane i reads the vertex's masks and keeps the level whose mask has
// bit i (kUnset if none).  Row of vertex v goes to row_of[v] (null: v).
__global__ __launch_bounds__(kBlock) void ms_rows_from_masks_kernel(const uint64_t *lmask,
                                                                    uint32_t V, uint32_t levels,
                                                                    const uint32_t *row_of,
                                                                    uint32_t *depth) {
  const int lane = lane_id();
  const uint32_t W = gridDim.x * kWavesPerBlock;
  for (uint32_t g = global_wave(); (size_t)g * 64 < V; g += W) {
    const uint32_t v = g * 64 + lane;
    uint32_t d[64];
#pragma unroll
    for (int j = 0; j < 64; j++) d[j] = kUnset;
    uint64_t next = v < V && levels ? lmask[v] : 0;
    for (uint32_t L = 0; L < levels; L++) {
      const uint64_t m = next;
      next = v < V && L + 1 < levels ? lmask[(size_t)(L + 1) * V + v] : 0;
      if (!__ballot(m != 0)) continue;
#pragma unroll
      for (int j = 0; j < 64; j++)
        if ((readlane_u64(m, j) >> lane) & 1ULL) d[j] = L;
    }
    const uint32_t r = v < V ? (row_of ? row_of[v] : v) : 0;
#pragma unroll
    for (int j = 0; j < 64; j++) {
      const uint32_t rj = (uint32_t)__builtin_amdgcn_readlane((int)r, j);
      if (g * 64 + j < V) depth[(size_t)rj * 64 + lane] = d[j];
    }
  }
}
// out-degrees of the vertices an SSSP reached (distance below the 1e9 sentinel)
__global__ void sssp_reach_kernel(const float *dist, const uint32_t *outdeg, uint32_t V,
                                  unsigned long long *acc) {
  uint32_t s = gridDim.x * blockDim.x;
  uint64_t x = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s)
    if (dist[v] < 1e9f) x += outdeg[v];
  x = wave_sum_u64(x);
  if (lane_id() == 0 && x) atomicAdd(&acc[(global_wave() & 63) * 16], (unsigned long long)x);
}
// sum over roots of the out-degrees of reached vertices
__global__ void ms_reach_kernel(const uint64_t *vis, const uint32_t *outdeg, uint32_t V,
                                unsigned long long *acc) {
  uint32_t s = gridDim.x * blockDim.x;
  uint64_t x = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s)
    x += (uint64_t)__popcll(vis[v]) * outdeg[v];
  x = wave_sum_u64(x);
  if (lane_id() == 0 && x) atomicAdd(&acc[(global_wave() & 63) * 16], (unsigned long long)x);
}
// Roots of a batch by value, and the order of their lanes (order_lanes).
struct Roots64 {
  uint32_t r[64];
};
struct LaneMap {
  uint8_t to_caller[64];  // internal lane -> the caller's root index
  uint8_t to_lane[64];    // the caller's root index -> internal lane
};
static LaneMap identity_lanes() {
  LaneMap m;
  for (int i = 0; i < 64; i++) m.to_caller[i] = m.to_lane[i] = (uint8_t)i;
  return m;
}
template <class T>
__global__ void unpermute_rows_kernel(const T *in, T *out, const uint32_t *iperm, uint32_t V) {
  uint64_t n = (uint64_t)V * 64, s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += s)
    out[(uint64_t)iperm[k >> 6] * 64 + (k & 63)] = in[k];
}

// ------------------------------------------- lane = root kernels (64 roots)
// One wave per vertex, lane i carries root i; per-(vertex, root) state is
// vertex-major, so every edge step moves one coalesced 256/512-byte row.
struct LaneArgs {
  const uint64_t *index;
  const uint32_t *nbr;
  const float *w;
  const uint32_t *list;  // vertices to visit (null: 0..n-1)
  uint32_t n;
  uint32_t light_max;   // longer lists go to lane_heavy_kernel
};

// One adjacency range of vertex v, 64 roots per wave (lane = root).  The
// lanes first load 64 arcs at once -- neighbour, weight and the neighbour's
// selection mask (the roots for which the arc matters, e.g. the roots whose
// distance at the neighbour changed) -- then walk only the selected arcs, in
// arc order, Op::kRows row loads in flight at a time.
// Op: T, L, kRows; uint64_t select(u); L load(u, lane); void combine(T&, L, w).
template <class Op>
__device__ __forceinline__ void lane_range(const Op &op, const uint32_t *nbr, const float *wt,
                                           uint64_t b, uint64_t e, uint64_t vm, int lane,
                                           typename Op::T &acc) {
  constexpr int kRows = Op::kRows;
  for (uint64_t base = b; base < e; base += 64) {
    const uint64_t k = base + lane;
    uint32_t u = 0;
    float w = 0.f;
    uint64_t sel = 0;
    if (k < e) {
      u = stream_load(nbr + k);
      if (wt) w = stream_load(wt + k);
      sel = op.select(u) & vm;
    }
    uint64_t any = __ballot(sel != 0);
    while (any) {
      uint32_t uj[kRows];
      float wj[kRows];
      bool on[kRows];
#pragma unroll
      for (int t = 0; t < kRows; t++) {
        uj[t] = 0;
        wj[t] = 0.f;
        on[t] = false;
        if (any) {
          const int j = __builtin_ctzll(any);
          any &= any - 1;
          uj[t] = (uint32_t)__builtin_amdgcn_readlane((int)u, j);
          wj[t] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w), j));
          on[t] = (readlane_u64(sel, j) >> lane) & 1ULL;
        }
      }
      typename Op::L x[kRows];
#pragma unroll
      for (int t = 0; t < kRows; t++) x[t] = on[t] ? op.load(uj[t], lane) : typename Op::L{};
#pragma unroll
      for (int t = 0; t < kRows; t++)
        if (on[t]) op.combine(acc, x[t], wj[t]);
    }
  }
}

// One wave per listed vertex.  Op adds: bool enabled() (the whole launch
// does nothing when false); bool active(v) (wave-uniform); T init(v, lane);
// bool on(v, lane, T) (root takes part at v); void finish(v, lane, T).  A
// grid of fewer than n waves loops.
template <class Op>
__global__ __launch_bounds__(kBlock) void lane_vertex_kernel(LaneArgs a, Op op) {
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const int lane = lane_id();
  if (!op.enabled()) return;
  for (uint32_t gw = global_wave(); gw < a.n; gw += W) {
    uint32_t v = a.list ? a.list[gw] : gw;
    if (!op.active(v)) continue;
    uint64_t b = a.index[v], e = a.index[v + 1];
    if (e - b > a.light_max || e == b) continue;
    typename Op::T acc = op.init(v, lane);
    const uint64_t vm = __ballot(op.on(v, lane, acc));
    if (!vm) continue;
    lane_range(op, a.nbr, a.w, b, e, vm, lane, acc);
    op.finish(v, lane, acc);
  }
}
// Heavy lists: one wave per chunk; partials through op.commit (atomics).
template <class Op>
__global__ __launch_bounds__(kBlock) void lane_heavy_kernel(HeavyArgs a, Op op) {
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const int lane = lane_id();
  if (!op.enabled()) return;
  for (uint32_t c = global_wave(); c < a.n_chunks; c += W) {
    uint32_t v = a.chunk_v[c];
    if (!op.active(v)) continue;
    typename Op::T acc = op.init(v, lane);
    const uint64_t vm = __ballot(op.on(v, lane, acc));
    if (!vm) continue;
    lane_range(op, a.nbr, a.w, a.chunk_beg[c], a.chunk_end[c], vm, lane, acc);
    op.commit(v, lane, acc);
  }
}

// Batched SSSP, dense step: dist row of v = min over in-edges of (row of u + w).
struct MssPull {
  using T = float;
  using L = float;
  static constexpr int kRows = 8;  // f32 rows: 8 loads in flight beat 4 by 4% (RMAT-24, 64 roots)
  float *dist;
  const uint64_t *cm_cur;  // roots whose distance at a vertex changed last step
  uint64_t *cm_next, *out_bits;
  const uint32_t *gate;    // null, or: run only when *gate == kGateDense
  __device__ bool enabled() const { return !gate || *gate == kGateDense; }
  __device__ bool active(uint32_t) const { return true; }
  __device__ float init(uint32_t, int) const { return 1e9f; }
  __device__ bool on(uint32_t, int, float) const { return true; }
  // only neighbours whose distance changed can lower v's: the others were
  // relaxed into v when they last changed (same fixed point, fewer rows read)
  __device__ uint64_t select(uint32_t u) const { return cm_cur[u]; }
  __device__ float load(uint32_t u, int lane) const { return dist[(size_t)u * 64 + lane]; }
  __device__ void combine(float &acc, float du, float w) const {
    float r = du + w;
    if (r < acc) acc = r;
  }
  __device__ void finish(uint32_t v, int lane, float acc) const {
    float *p = &dist[(size_t)v * 64 + lane];
    bool imp = acc < *p;
    if (imp) *p = acc;
    uint64_t m = __ballot(imp);
    if (m && lane == 0) {
      cm_next[v] = m;
      activate(out_bits, v);
    }
  }
  __device__ void commit(uint32_t v, int lane, float acc) const {
    float *p = &dist[(size_t)v * 64 + lane];
    bool imp = false;
    if (acc < *p) imp = acc < __uint_as_float(atomicMin((uint32_t *)p, __float_as_uint(acc)));
    uint64_t m = __ballot(imp);
    if (m && lane == 0) {
      atomicOr((unsigned long long *)&cm_next[v], (unsigned long long)m);
      activate(out_bits, v);
    }
  }
};
// Batched SSSP, sparse step: each changed (vertex, root) relaxes its out-edges.
// Arcs [b, e) of a changed vertex u, lane = root: roots whose distance at u
// changed last step relax each arc (sssp.cpp:44-56); improvements go to
// the target's mask of changed roots and its frontier bit.
__device__ __forceinline__ void mss_relax_range(const uint32_t *nbr, const float *wt, float *dist,
                                                const uint64_t *cm_cur, uint64_t *cm_next,
                                                uint64_t *out_bits, uint32_t u, uint64_t b,
                                                uint64_t e) {
  const int lane = lane_id();
  uint64_t m = cm_cur[u];
  bool act = (m >> lane) & 1ULL;
  float du = dist[(size_t)u * 64 + lane];
  for (uint64_t k = b; k < e; k++) {
    uint32_t v = nbr[k];
    bool imp = false;
    if (act) {
      float c = du + wt[k];
      float *p = &dist[(size_t)v * 64 + lane];
      if (c < *p) imp = c < __uint_as_float(atomicMin((uint32_t *)p, __float_as_uint(c)));
    }
    uint64_t im = __ballot(imp);
    if (im && lane == 0) {
      atomicOr((unsigned long long *)&cm_next[v], (unsigned long long)im);
      activate(out_bits, v);
    }
  }
}
__global__ __launch_bounds__(kBlock) void mss_push_kernel(const uint64_t *index, const uint32_t *nbr,
                                                           const float *wt, const uint32_t *q0,
                                                           uint32_t n0, const uint32_t *q1,
                                                           uint32_t n1, const uint32_t *q2,
                                                           uint32_t n2, uint32_t waves2,
                                                           float *dist, const uint64_t *cm_cur,
                                                           uint64_t *cm_next, uint64_t *out_bits) {
  const uint32_t gw = global_wave();
  auto relax_range = [&](uint32_t u, uint64_t b, uint64_t e) {
    mss_relax_range(nbr, wt, dist, cm_cur, cm_next, out_bits, u, b, e);
  };
  if (gw < n0 + n1) {
    uint32_t u = gw < n0 ? q0[gw] : q1[gw - n0];
    relax_range(u, index[u], index[u + 1]);
  } else {
    // the grid is rounded up to whole blocks: waves past waves2 have no share
    uint32_t hw = gw - n0 - n1;
    if (hw >= waves2) return;
    for (uint32_t h = 0; h < n2; h++) {
      uint32_t u = q2[h];
      uint64_t b = index[u], e = index[u + 1];
      uint64_t nch = (e - b + kChunkEdges - 1) / kChunkEdges;
      // chunk c goes to wave (c + rot) % waves2, rotated per vertex
      const uint32_t rot = (uint32_t)(((uint64_t)h * 2654435761u) % waves2);
      for (uint64_t c = (hw + waves2 - rot) % waves2; c < nch; c += waves2) {
        uint64_t cb = b + c * kChunkEdges;
        relax_range(u, cb, cb + kChunkEdges < e ? cb + kChunkEdges : e);
      }
    }
  }
}
__global__ void mss_seed_kernel(const uint32_t *roots, int n, float *dist, uint64_t *cm,
                                uint64_t *bits) {
  for (int i = 0; i < n; i++) {
    uint32_t r = roots[i];
    dist[(size_t)r * 64 + i] = 0.f;
    cm[r] |= 1ULL << i;
    bits[r >> 6] |= 1ULL << (r & 63);
  }
}

// Batched BC, forward: sigma row of each newly reached vertex = sum of the
// sigma of its in-neighbours in the previous level, per root (bc.cpp:58-89).
struct MsbcSigma {
  using T = double;
  using L = double;
  static constexpr int kRows = 4;
  __device__ bool enabled() const { return true; }
  double *sigma;
  const uint64_t *fprev, *fnew;
  __device__ bool active(uint32_t v) const { return fnew[v] != 0; }
  __device__ double init(uint32_t, int) const { return 0.0; }
  __device__ bool on(uint32_t v, int lane, double) const { return (fnew[v] >> lane) & 1ULL; }
  __device__ uint64_t select(uint32_t u) const { return fprev[u]; }
  __device__ double load(uint32_t u, int lane) const { return sigma[(size_t)u * 64 + lane]; }
  __device__ void combine(double &acc, double x, float) const { acc += x; }
  __device__ void finish(uint32_t v, int lane, double acc) const {
    if ((fnew[v] >> lane) & 1ULL) sigma[(size_t)v * 64 + lane] = acc;
  }
  __device__ void commit(uint32_t v, int lane, double acc) const {
    if (((fnew[v] >> lane) & 1ULL) && acc != 0.0) unsafeAtomicAdd(&sigma[(size_t)v * 64 + lane], acc);
  }
};
// Batched BC, backward: for v at depth L-1 of root i,
// D(v) += sum of D(w) over out-edges v->w with w at depth L (bc.cpp:125-167).
struct MsbcBack {
  using T = double;
  static constexpr int kRows = 4;
  __device__ bool enabled() const { return true; }
  struct L {
    uint32_t d;
    double dd;
  };
  double *dd;
  const uint32_t *depth;
  const uint64_t *level_bits;  // vertices at depth L-1 for some root
  const uint64_t *next_bits;   // vertices at depth L for some root
  // per-vertex root masks of levels L-1 and L (null: tested on the depth rows)
  const uint64_t *mask_prev, *mask_next;
  uint32_t lvl;
  __device__ bool active(uint32_t v) const { return test_bit(level_bits, v); }
  __device__ double init(uint32_t, int) const { return 0.0; }
  __device__ bool on(uint32_t v, int lane, double = 0.0) const {
    return mask_prev ? (mask_prev[v] >> lane) & 1ULL : depth[(size_t)v * 64 + lane] == lvl - 1;
  }
  __device__ uint64_t select(uint32_t u) const {
    return mask_next ? mask_next[u] : test_bit(next_bits, u) ? ~0ULL : 0ULL;
  }
  __device__ L load(uint32_t u, int lane) const {
    return L{mask_next ? lvl : depth[(size_t)u * 64 + lane], dd[(size_t)u * 64 + lane]};
  }
  __device__ void combine(double &acc, L x, float) const {
    if (x.d == lvl) acc += x.dd;
  }
  __device__ void finish(uint32_t v, int lane, double acc) const {
    if (on(v, lane)) dd[(size_t)v * 64 + lane] += acc;
  }
  __device__ void commit(uint32_t v, int lane, double acc) const {
    if (acc != 0.0 && on(v, lane)) unsafeAtomicAdd(&dd[(size_t)v * 64 + lane], acc);
  }
};
__global__ void msbc_seed_sigma_kernel(const uint32_t *roots, int n, double *sigma) {
  for (int i = 0; i < n; i++) sigma[(size_t)roots[i] * 64 + i] = 1.0;
}
// inv = 1/sigma in place (+inf where unreached), D = inv
__global__ void msbc_inv_kernel(double *sigma_inv, double *dd, uint64_t n) {
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += s) {
    double inv = 1 / sigma_inv[k];
    sigma_inv[k] = inv;
    dd[k] = inv;
  }
}
// Entries k < live come from the traversal; the rows after them belong to
// vertices no root can reach (no in-arcs, not a root): sigma 0, so
// inv = 1/0 = +inf and D = (inf - inf) / inf = NaN, written without reading.
// In place; a wave holds one 64-lane row per step (n and live are whole
// rows, blocks whole waves), so the lanes go back to the caller's order by a
// shuffle: caller lane c takes internal lane m.to_lane[c].
__global__ void msbc_final_kernel(double *dd, double *inv, uint64_t live, uint64_t n, LaneMap m) {
  uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  const int lane = lane_id();
  const int from = m.to_lane[lane];
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += s) {
    if (k < live) {
      const double x = inv[k];
      const double dep = (dd[k] - x) / x;  // bc.cpp:170-176
      const double xo = shfl_f64(x, from), depo = shfl_f64(dep, from);
      inv[k] = xo;
      dd[k] = depo;
    } else {
      inv[k] = __builtin_inf();
      dd[k] = __builtin_nan("");
    }
  }
}
// The same on a relabelled graph, with both rows written in original order
// in the same pass (one read of each row instead of three passes).
__global__ void msbc_final_unpermute_kernel(const double *dd, const double *inv,
                                            const uint32_t *iperm, uint32_t V, uint64_t live,
                                            double *dep_out, double *inv_out, LaneMap m) {
  const uint64_t n = (uint64_t)V * 64, s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += s) {
    const uint64_t o = (uint64_t)iperm[k >> 6] * 64 + m.to_caller[k & 63];
    if (k < live) {
      const double x = inv[k];
      dep_out[o] = (dd[k] - x) / x;  // bc.cpp:170-176
      inv_out[o] = x;
    } else {
      dep_out[o] = __builtin_nan("");
      inv_out[o] = __builtin_inf();
    }
  }
}

// ------------------------------------------------- batched-query plumbing
static int batch_roots(const gg_graph *g, const uint32_t *h_roots, int n_roots, uint32_t *out) {
  if (!h_roots || n_roots < 1 || n_roots > 64)
    return fail(GG_EINVAL, "batched query: 1 to 64 roots");
  for (int i = 0; i < n_roots; i++) {
    if (h_roots[i] >= g->vertices) return fail(GG_EINVAL, "batched query: root out of range");
    out[i] = g->to_internal(h_roots[i]);
  }
  return GG_OK;
}

// Lane order of a batch.  A row gather loads only the lanes (roots) an arc
// selects, and a 64-lane row of doubles spans four 128-byte lines, so roots
// whose searches reach each level at the same pace are put side by side:
// their lanes tend to be selected together and fall into fewer lines.  The
// pace is estimated by the root's two-hop reach (the out-degree sum of its
// out-neighbours), lanes in descending order of it.  Results do not depend
// on the order; the outputs are written back in the caller's order.
__global__ void root_reach_kernel(const uint64_t *index, const uint32_t *nbr, const uint32_t *deg,
                                  Roots64 roots, uint64_t *out) {
  const uint32_t r = roots.r[blockIdx.x];
  const int lane = lane_id();
  const uint64_t b = index[r], e = index[r + 1];
  uint64_t s = 0;
  for (uint64_t k = b + lane; k < e; k += 64) s += deg[nbr[k]];
  for (int d = 32; d >= 1; d >>= 1) s += shfl_u64(s, lane ^ d);
  if (lane == 0) out[blockIdx.x] = s;
}
// roots (internal ids, caller order) -> roots in lane order, and the map;
// d_tmp: device scratch of tmp_bytes (n x 8 needed, else the order is kept).
static int order_lanes(const Dir &out, uint32_t *roots, int n, uint64_t *d_tmp, size_t tmp_bytes,
                       LaneMap *m) {
  *m = identity_lanes();
  if (n < 2 || tmp_bytes < 8 * (size_t)n) return GG_OK;
  Roots64 r;
  for (int i = 0; i < n; i++) r.r[i] = roots[i];
  hipLaunchKernelGGL(root_reach_kernel, dim3(n), dim3(64), 0, stream(), out.index, out.nbr,
                     out.degree, r, d_tmp);
  uint64_t reach[64];
  GG_HIP(hipMemcpyAsync(reach, d_tmp, 8 * (size_t)n, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  uint8_t ord[64];
  for (int i = 0; i < n; i++) ord[i] = (uint8_t)i;
  std::stable_sort(ord, ord + n, [&](uint8_t a, uint8_t b) { return reach[a] > reach[b]; });
  for (int j = 0; j < n; j++) {
    m->to_caller[j] = ord[j];
    m->to_lane[ord[j]] = (uint8_t)j;
    roots[j] = r.r[ord[j]];
  }
  return GG_OK;
}

// Store of per-level root masks (batched BFS/BC): initial levels.
constexpr size_t kMaskLevels0 = 16;

size_t work_footprint(const gg_graph *g, int kind) {
  // mirrors work_get and ms_alloc with results kept in the context (+ the
  // first ms_store_level growth for BC)
  size_t V = (size_t)g->vertices + 1, words = ((size_t)g->vertices >> 6) + 1;
  size_t nh = std::max<size_t>(std::max(g->out.n_heavy, g->in.n_heavy), 1);
  size_t b = 2 * (words * 8 + 6 * V * 4 + sizeof(Counters)) + 2 * words * 8 + V * 8 +
             (g->reordered ? 2 * V * 8 : 0) + nh * 8 + 1024 * 8;
  b += 3 * (V * 8 + 8) + 64 * 4 + V * 64 * 4;
  b += kMaskLevels0 * (V - 1) * 8;  // per-level root masks
  if (kind == GG_QUERY_BC) b += 2 * V * 64 * 8 + 64 * words * 8;
  return b;
}

static int ms_alloc(gg_graph *g, Work *w, bool bc, bool keep = false) {
  size_t V = g->vertices;
  if (!w->ms_vis) {
    GG_TRY(device_alloc((void **)&w->ms_vis, V * 8 + 8));
    GG_TRY(device_alloc((void **)&w->ms_f[0], V * 8 + 8));
    GG_TRY(device_alloc((void **)&w->ms_f[1], V * 8 + 8));
    GG_TRY(device_alloc((void **)&w->ms_roots, 64 * 4));
    GG_TRY(device_alloc(&w->ms_tmp, V * 64 * 4));
  }
  if (bc && (g->reordered || keep) && !w->mb_a) {
    GG_TRY(device_alloc(&w->mb_a, V * 64 * 8));
    GG_TRY(device_alloc(&w->mb_b, V * 64 * 8));
  }
  return GG_OK;
}

// Keeps a copy of one level's frontier bitmap (BC's `levels`, bc.cpp:33, :99).
static int ms_store_level(Work *w, const uint64_t *bits, uint32_t level) {
  if (level >= w->ms_levels_cap) {
    size_t cap = std::max<size_t>(64, 2 * (size_t)level + 2);
    uint64_t *p = nullptr;
    GG_TRY(device_alloc((void **)&p, cap * w->words * 8));
    if (w->ms_levels)
      GG_HIP(hipMemcpyAsync(p, w->ms_levels, w->ms_levels_cap * w->words * 8,
                            hipMemcpyDeviceToDevice, stream()));
    GG_HIP(hipStreamSynchronize(stream()));
    device_free(w->ms_levels);
    w->ms_levels = p;
    w->ms_levels_cap = cap;
  }
  GG_HIP(hipMemcpyAsync(w->ms_levels + (size_t)level * w->words, bits, w->words * 8,
                        hipMemcpyDeviceToDevice, stream()));
  return GG_OK;
}

// Room for level `level` in the store of per-level root masks (grown by
// doubling).  False when the store would exceed kMaskStoreMax (e.g. a
// long-diameter graph): the caller then falls back to depth rows.
static int ms_mask_room(Work *w, uint32_t level, bool *ok) {
  const size_t row = (size_t)w->V * 8;
  const size_t limit = (size_t)options().level_mask_bytes;
  const size_t need = (size_t)level + 1;
  *ok = need * row <= limit;
  if (!*ok || level < w->ms_lmask_cap) return GG_OK;
  // grow (never shrink): the old capacity is <= level < the new one
  const size_t cap = std::min(std::max<size_t>(kMaskLevels0, 2 * need), limit / row);
  uint64_t *p = nullptr;
  GG_TRY(device_alloc((void **)&p, cap * row));
  if (w->ms_lmask)
    GG_HIP(hipMemcpyAsync(p, w->ms_lmask, w->ms_lmask_cap * row, hipMemcpyDeviceToDevice,
                          stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  device_free(w->ms_lmask);
  w->ms_lmask = p;
  w->ms_lmask_cap = cap;
  return GG_OK;
}

static int ms_rows_from_masks(Work *w, uint32_t levels, const uint32_t *row_of, uint32_t *depth) {
  const uint32_t V = w->V;
  hipLaunchKernelGGL(ms_rows_from_masks_kernel, dim3(grid_waves(std::min<uint64_t>((V + 63) / 64, 16384))),
                     dim3(kBlock), 0, stream(), w->ms_lmask, V, levels, row_of, depth);
  GG_HIP(hipGetLastError());
  return GG_OK;
}

// What a multi-root BFS run leaves: the number of levels and whether their
// root masks are in the store (else the depth rows were written).
struct MsLevels {
  uint32_t levels = 0;
  bool masks = false;
};

// Multi-root BFS levels (bit-parallel masks); hook(level, fprev, fnew) runs
// after every non-empty level with the level's frontier compacted in en.fin()
// (and for level 0 when store_levels).  With level masks on, each level's
// root masks go to the store and `depth` (rows in internal order) is left
// alone; if the store runs out of room, the rows of the levels so far are
// built from it and the remaining levels write rows directly.
template <class Hook>
static int ms_bfs_run(gg_graph *g, Work *w, Engine &en, const uint32_t *roots, int n_roots,
                      uint32_t *depth, Hook hook, bool store_levels, MsLevels *res) {
  const Dir &out = out_dir(g), &in = in_dir(g);
  uint32_t V = g->vertices;
  bool masks = options().level_masks != 0;
  if (masks) GG_TRY(ms_mask_room(w, 0, &masks));
  if (!masks) GG_HIP(hipMemsetAsync(depth, 0xFF, (size_t)V * 64 * 4, stream()));
  GG_HIP(hipMemsetAsync(w->ms_vis, 0, (size_t)V * 8, stream()));
  GG_HIP(hipMemsetAsync(w->ms_f[0], 0, (size_t)V * 8, stream()));
  GG_HIP(hipMemcpyAsync(w->ms_roots, roots, 4 * (size_t)n_roots, hipMemcpyHostToDevice, stream()));
  GG_TRY(en.clear_bits(en.fin()));
  hipLaunchKernelGGL(ms_seed_kernel, dim3(1), dim3(64), 0, stream(), w->ms_roots, n_roots,
                     w->ms_vis, w->ms_f[0], en.fin().bits, masks ? nullptr : depth);
  GG_TRY(en.compact_in(out.degree));
  if (masks)
    GG_HIP(hipMemcpyAsync(w->ms_lmask, w->ms_f[0], (size_t)V * 8, hipMemcpyDeviceToDevice,
                          stream()));
  if (store_levels) GG_TRY(hook(0, nullptr, w->ms_f[0]));
  const uint64_t full = n_roots == 64 ? ~0ULL : ((1ULL << n_roots) - 1ULL);
  PlanRef plan_ref;  // held until this call returns
  GG_TRY(pr_plan_get(g, &plan_ref));
  const PrPlan *plan = plan_ref.get();
  int fc = 0;
  uint32_t level = 1;
  for (; en.in.total > 0; level++) {
    uint64_t *fcur = w->ms_f[fc], *fnext = w->ms_f[fc ^ 1];
    GG_HIP(hipMemsetAsync(fnext, 0, (size_t)V * 8, stream()));
    bool sp = en.sparse(en.in.edges);
    if (sp) {
      MsPush op;
      op.fcur = fcur;
      op.vis = w->ms_vis;
      op.fnext = fnext;
      GG_TRY(en.launch_push(out, op, false));
    } else if (plan) {
      // relabelled graph: the dense step is an OR over every in-arc, done on
      // the destination-bin plan (LDS masks per bin, prplan.hip)
      GG_TRY(ms_plan_iteration(g, plan, w, fcur, w->ms_vis, full, fnext, en.fout().bits));
    } else {
      MsPull op;
      op.fcur = fcur;
      op.vis = w->ms_vis;
      op.fnext = fnext;
      op.full = full;
      MsPullArgs pa{in.index, in.nbr, fcur, w->ms_vis, fnext, en.fout().bits, full, V,
                    in.light_max, (uint32_t)options().pull_serial_max, (V + 63) / 64, nullptr};
      {
        ProfScope p("ms_pull");
        hipLaunchKernelGGL(ms_pull_group_kernel, dim3(grid_waves(pa.ngroups)), dim3(kBlock), 0,
                           stream(), pa);
      }
      GG_TRY(en.launch_pull_heavy(in, op, "ms_pull_heavy"));
    }
    if (masks) {
      bool room;
      GG_TRY(ms_mask_room(w, level, &room));
      if (!room) {  // rows of levels [0, level) from the store, then rows directly
        GG_TRY(ms_rows_from_masks(w, level, nullptr, depth));
        masks = false;
      }
    }
    {
      ProfScope p("ms_apply");
      hipLaunchKernelGGL(ms_apply_kernel, dim3(grid_waves(std::min<size_t>(w->words, 1u << 16))),
                         dim3(kBlock), 0, stream(), en.fout().bits, w->words, w->ms_vis, fnext,
                         masks ? nullptr : depth, level);
    }
    if (masks)
      GG_HIP(hipMemcpyAsync(w->ms_lmask + (size_t)level * V, fnext, (size_t)V * 8,
                            hipMemcpyDeviceToDevice, stream()));
    GG_TRY(en.finish_iteration(sp, out.degree, nullptr));
    if (en.in.total > 0) GG_TRY(hook(level, fcur, fnext));
    fc ^= 1;
  }
  res->levels = level;  // levels 0 .. level-1 hold vertices (the last one may be empty)
  res->masks = masks;
  return GG_OK;
}

// ------------------------------------------- batched BFS, device-driven loop
// The level loop of the batched BFS without a host round trip per level.
// Every level is the same fixed sequence of launches, issued ahead on one
// stream; each kernel reads the level's mode from the control block and
// returns at once when the level is not its kind or the traversal is over:
//   bin    -- push only: frontier bitmap -> three degree-binned queues;
//   push   -- push only: a persistent grid over the queues (MsPush);
//   dense  -- pull only: the destination-bin plan (relabelled graphs), or
//             the 64-vertex group pull plus the heavy in-list chunks;
//   apply  -- every level: the new masks join visited, the level's root
//             masks (or depth rows) are stored, the new frontier is counted,
//             and the spent frontier's bitmap and masks are cleared, so each
//             buffer of the two pairs is clean when it is next written;
//   decide -- one block: the push/pull switch on the new frontier's
//             out-arcs (core/graph.hpp:1509), the level's records, the end.
// The host stays one level ahead and polls the decisions in coherent pinned
// memory, as the single-root BFS loop does.
enum : uint32_t { kMsPush = 0, kMsPull = kGateDense, kMsDone = 2 };
static_assert(kMsPull != kMsPush && kMsPull != kMsDone, "distinct modes");
static_assert(kBlock == (int)kScalarSlots, "one count slot per decide thread");

struct MsCtl {
  uint32_t done, iter;   // iter = iterations decided
  uint32_t mode, level;  // level = the one the current iteration produces
  uint32_t push_iters;
  int32_t force;
  uint64_t threshold;
  uint32_t modes[GG_STATS_RECORDS];
  uint64_t active[GG_STATS_RECORDS], fresh[GG_STATS_RECORDS];
};
struct MsHostFlag {
  uint32_t decided;  // iterations decided so far
  uint32_t done;
};
// One iteration's launches captured as a graph, with everything the
// launches bake in (buffers, shapes, options): replayed while that stays.
struct MsdGraph {
  hipGraphExec_t exec = nullptr;
  uint64_t key[16] = {};
};
struct MsdState {
  MsdGraph graphs[2];  // by iteration parity (the buffer pairs alternate)
  MsCtl *ctl = nullptr;
  uint64_t *set = nullptr;  // kScalarSlots count slots of 16 u64 (one 128-byte line)
  Counters *qc = nullptr;
  uint64_t *settled = nullptr;  // bitmap: every root has reached the vertex
  MsCtl *h_ctl = nullptr;         // pinned
  MsHostFlag *h_flag = nullptr;   // pinned, coherent: written by the device
};

void msd_state_free(void *p) {
  MsdState *s = (MsdState *)p;
  if (!s) return;
  device_free(s->ctl);
  device_free(s->set);
  device_free(s->qc);
  device_free(s->settled);
  for (MsdGraph &gr : s->graphs)
    if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
  if (s->h_ctl) (void)hipHostFree(s->h_ctl);
  if (s->h_flag) (void)hipHostFree(s->h_flag);
  delete s;
}

static int msd_state(Work *w, MsdState **out) {
  MsdState *s = (MsdState *)w->msd;
  if (!s) {
    s = new MsdState;
    w->msd = s;
    GG_TRY(device_alloc((void **)&s->ctl, sizeof(MsCtl)));
    GG_TRY(device_alloc((void **)&s->set, (size_t)kScalarSlots * 16 * 8));
    GG_TRY(device_alloc((void **)&s->qc, sizeof(Counters)));
    GG_TRY(device_alloc((void **)&s->settled, w->words * 8));
    GG_HIP(hipHostMalloc((void **)&s->h_ctl, sizeof(MsCtl)));
    GG_HIP(hipHostMalloc((void **)&s->h_flag, sizeof(MsHostFlag), hipHostMallocCoherent));
  }
  *out = s;
  return GG_OK;
}

// Reduces the new frontier's counts (the spread slots are cleared for the
// next level) and decides the next iteration.  All kBlock threads of one
// block.
__device__ void msd_decide(MsCtl *c, uint64_t *set, Counters *qc, MsHostFlag *hf) {
  // every load is issued before any result is used (the slots by all
  // threads, the control block's scalars by thread 0): one round of memory
  // latency instead of three dependent ones
  __shared__ uint64_t s_part[2][kWavesPerBlock];
  const int t = threadIdx.x;
  const int lane = lane_id(), wv = t >> 6;
  uint32_t done = 0, i = 0, push_iters = 0;
  int32_t force = 0;
  uint64_t threshold = 0;
  if (t == 0) {
    done = c->done;
    i = c->iter;
    push_iters = c->push_iters;
    force = c->force;
    threshold = c->threshold;
  }
  uint64_t *p = set + (size_t)t * 16;
  uint64_t n = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t e = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (n | e) {  // a slot nobody added to is already 0
    __hip_atomic_store(p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  n = wave_sum_u64(n);
  e = wave_sum_u64(e);
  if (lane == 0) {
    s_part[0][wv] = n;
    s_part[1][wv] = e;
  }
  __syncthreads();
  if (t != 0 || done) return;
  uint64_t total = 0, edges = 0;
  for (int k = 0; k < kWavesPerBlock; k++) {
    total += s_part[0][k];
    edges += s_part[1][k];
  }
  if (i > 0 && i - 1 < GG_STATS_RECORDS) c->fresh[i - 1] = total;
  if (total == 0) {
    c->done = 1;
    c->mode = kMsDone;
    __threadfence_system();
    __hip_atomic_store(&hf->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  // core/graph.hpp:1509: sparse iff the frontier's out-arcs < |E| / divisor
  const uint32_t mode = force == 1          ? kMsPush
                        : force == 2        ? kMsPull
                        : edges < threshold ? kMsPush
                                            : kMsPull;
  if (i < GG_STATS_RECORDS) {
    c->modes[i] = mode == kMsPush ? 0u : 1u;
    c->active[i] = total;
  }
  c->mode = mode;
  c->level = i + 1;
  c->iter = i + 1;
  if (mode == kMsPush) c->push_iters = push_iters + 1;
  qc->n[0] = qc->n[1] = qc->n[2] = 0;
  qc->total = 0;
  qc->edges = 0;
  __threadfence_system();
  __hip_atomic_store(&hf->decided, i + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// One block: the roots (root i gets bit i of its vertex's masks; roots may
// repeat), the control block, the count of distinct roots and their
// out-arcs, and the decision for the first iteration.
__global__ __launch_bounds__(kBlock) void msd_seed_kernel(MsCtl *c, uint64_t *set, Counters *qc,
                                                          MsHostFlag *hf, const uint32_t *roots,
                                                          int n, uint64_t *vis, uint64_t *f,
                                                          uint64_t *bits, uint32_t *depth,
                                                          const uint32_t *outdeg, int32_t force,
                                                          uint64_t threshold) {
  const int t = threadIdx.x;
  if (t < n) {
    const uint32_t r = roots[t];
    atomicOr((unsigned long long *)&vis[r], 1ULL << t);
    atomicOr((unsigned long long *)&f[r], 1ULL << t);
    atomicOr((unsigned long long *)&bits[r >> 6], 1ULL << (r & 63));
    if (depth) depth[(size_t)r * 64 + t] = 0;
    bool first = true;
    for (int j = 0; j < t; j++) first = first && roots[j] != r;
    if (first) {
      atomicAdd((unsigned long long *)&set[0], 1ull);
      atomicAdd((unsigned long long *)&set[1], (unsigned long long)outdeg[r]);
    }
  }
  if (t == 0) {
    c->done = 0;
    c->iter = 0;
    c->mode = kMsDone;
    c->level = 0;
    c->push_iters = 0;
    c->force = force;
    c->threshold = threshold;
  }
  __threadfence();
  __syncthreads();
  msd_decide(c, set, qc, hf);
}

__global__ __launch_bounds__(kBlock) void msd_decide_kernel(MsCtl *c, uint64_t *set, Counters *qc,
                                                            MsHostFlag *hf) {
  msd_decide(c, set, qc, hf);
}

__global__ __launch_bounds__(kCompactBlock) void msd_bin_kernel(const MsCtl *c, const uint64_t *fin,
                                                                size_t words, const uint32_t *deg,
                                                                uint32_t small_max,
                                                                uint32_t big_max, uint32_t *q0,
                                                                uint32_t *q1, uint32_t *q2,
                                                                Counters *qc) {
  if (c->mode != kMsPush) return;
  compact_words_spread(fin, words, deg, small_max, big_max, q0, q1, q2, qc, nullptr, nullptr,
                       gridDim.x);
}

struct MsdPushArgs {
  const uint64_t *index;
  const uint32_t *nbr;
  const uint32_t *q0, *q1, *q2;
  const Counters *qc;
  uint64_t *fout;
};

// Persistent grid over the three bins (sizes on the device): 64 small
// vertices per wave step (in-wave load balancing), one mid vertex per wave
// step, then chunks of every huge vertex shared by all waves, rotated per
// vertex so a vertex with few chunks does not load the first waves only.
__global__ __launch_bounds__(kBlock) void msd_push_kernel(const MsCtl *c, MsdPushArgs a,
                                                          MsPush op) {
  if (c->mode != kMsPush) return;
  const uint32_t W = gridDim.x * kWavesPerBlock, gw = global_wave();
  const int lane = lane_id();
  const uint32_t n0 = a.qc->n[0], n1 = a.qc->n[1], n2 = a.qc->n[2];
  const uint32_t waves0 = (n0 + 63) / 64;
  for (uint32_t vw = gw; vw < waves0 + n1; vw += W) {
    if (vw < waves0) {
      const uint32_t i = vw * 64 + lane;
      const bool valid = i < n0;
      const uint32_t v = valid ? a.q0[i] : 0;
      const uint64_t b = valid ? a.index[v] : 0;
      const uint32_t deg = valid ? (uint32_t)(a.index[v + 1] - b) : 0;
      const uint64_t m = valid ? op.fcur[v] : 0;
      uint32_t total;
      const uint32_t excl = wave_excl_scan(deg, &total);
      for (uint32_t j0 = 0; j0 < total; j0 += 64) {
        const uint32_t j = j0 + lane;
        const int k = wave_owner(excl, j);
        const uint64_t bk = shfl_u64(b, k), mk = shfl_u64(m, k);
        const uint32_t ek = __shfl(excl, k);
        if (j < total) {
          const uint32_t dst = a.nbr[bk + (j - ek)];
          if (op.edge(0, mk, dst, 0.f)) activate(a.fout, dst);
        }
      }
    } else {
      const uint32_t v = a.q1[vw - waves0];
      const uint64_t b = a.index[v], e1 = a.index[v + 1];
      const uint64_t m = op.fcur[v];
      for (uint64_t e = b + lane; e < e1; e += 64) {
        const uint32_t dst = a.nbr[e];
        if (op.edge(0, m, dst, 0.f)) activate(a.fout, dst);
      }
    }
  }
  for (uint32_t h = 0; h < n2; h++) {
    const uint32_t v = a.q2[h];
    const uint64_t b = a.index[v], e1 = a.index[v + 1];
    const uint64_t m = op.fcur[v];
    const uint64_t nch = (e1 - b + kChunkEdges - 1) / kChunkEdges;
    const uint32_t rot = (uint32_t)(((uint64_t)h * 2654435761u) % W);
    for (uint64_t ch = (gw + W - rot) % W; ch < nch; ch += W) {
      const uint64_t cb = b + ch * kChunkEdges;
      const uint64_t ce = cb + kChunkEdges < e1 ? cb + kChunkEdges : e1;
      for (uint64_t e = cb + lane; e < ce; e += 64) {
        const uint32_t dst = a.nbr[e];
        if (op.edge(0, m, dst, 0.f)) activate(a.fout, dst);
      }
    }
  }
}

struct MsdHeavyArgs {
  const uint32_t *nbr;
  const uint32_t *chunk_v;
  const uint64_t *chunk_beg, *chunk_end;
  uint32_t n_chunks;
  const uint64_t *fcur, *vis;
  uint64_t *fnext, *fout;
  uint64_t full;
};

// The heavy in-lists of the dense step (graphs without the bin plan): one
// wave per chunk ORs the in-neighbours' masks, 64 entries per step, until
// every root still missing from the vertex is found; the new bits go to the
// vertex's mask and bitmap word with atomics (chunks of a vertex combine).
__global__ __launch_bounds__(kBlock) void msd_heavy_kernel(const MsCtl *c, MsdHeavyArgs a) {
  if (c->mode != kMsPull) return;
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const int lane = lane_id();
  for (uint32_t ch = global_wave(); ch < a.n_chunks; ch += W) {
    const uint32_t v = a.chunk_v[ch];
    const uint64_t need = a.full & ~a.vis[v];
    if (!need) continue;
    const uint64_t e1 = a.chunk_end[ch];
    uint64_t acc = 0;
    for (uint64_t e0 = a.chunk_beg[ch]; e0 < e1; e0 += 64) {
      const uint64_t e = e0 + lane;
      uint64_t x = e < e1 ? a.fcur[stream_load(&a.nbr[e])] : 0ULL;
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) x |= shfl_u64(x, lane ^ m);
      acc |= x;
      if ((acc & need) == need) break;
    }
    const uint64_t nw = acc & need;
    if (lane == 0 && nw) {
      atomicOr((unsigned long long *)&a.fnext[v], (unsigned long long)nw);
      atomicOr((unsigned long long *)&a.fout[v >> 6], 1ULL << (v & 63));
    }
  }
}

struct MsdApplyArgs {
  uint64_t *fin, *fout;    // spent and new frontier bitmaps
  uint64_t *fcur, *fnext;  // spent and new masks
  uint64_t *vis;
  uint64_t *settled;       // bit v set once vis[v] covers full
  uint64_t full;
  uint64_t *lmask;         // per-level root masks (V per level), or null
  uint32_t *depth;         // depth rows when lmask is null (internal order), or null
  const uint32_t *outdeg;
  uint64_t *set;
  size_t words;
  uint32_t V;
};

// One wave per 64-vertex bitmap word (lane = vertex): see the loop above.
__global__ __launch_bounds__(kBlock) void msd_apply_kernel(const MsCtl *c, MsdApplyArgs a) {
  if (c->mode == kMsDone) return;
  const uint32_t level = c->level;
  const int lane = lane_id();
  const size_t W = (size_t)gridDim.x * kWavesPerBlock;
  uint64_t cnt = 0, edg = 0;
  for (size_t wi = global_wave(); wi < a.words; wi += W) {
    const uint64_t act = a.fout[wi], old = a.fin[wi];
    const uint32_t v = (uint32_t)(wi * 64 + lane);
    uint64_t nw = 0;
    bool now_settled = false;
    if ((act >> lane) & 1ULL) {
      // every step writes only roots not yet in vis, so fnext[v] stays as it is
      nw = a.fnext[v];
      const uint64_t vs = a.vis[v] | nw;
      a.vis[v] = vs;
      now_settled = (vs & a.full) == a.full;
      cnt++;
      edg += a.outdeg[v];
    }
    const uint64_t sm = __ballot(now_settled);
    if (lane == 0 && sm) a.settled[wi] |= sm;  // this wave owns word wi
    if ((old >> lane) & 1ULL) a.fcur[v] = 0;
    if (lane == 0 && old) a.fin[wi] = 0;
    if (a.lmask && v < a.V) a.lmask[(size_t)level * a.V + v] = nw;
    if (a.depth) {
      for (uint64_t x = act; x; x &= x - 1) {
        const int k = __builtin_ctzll(x);
        const uint64_t m = shfl_u64(nw, k);
        if ((m >> lane) & 1ULL) a.depth[(wi * 64 + (size_t)k) * 64 + lane] = level;
      }
    }
  }
  cnt = wave_sum_u64(cnt);
  edg = wave_sum_u64(edg);
  if (lane == 0 && cnt) {
    uint64_t *p = a.set + (size_t)(global_wave() % kScalarSlots) * 16;
    atomicAdd((unsigned long long *)p, (unsigned long long)cnt);
    atomicAdd((unsigned long long *)(p + 1), (unsigned long long)edg);
  }
}

// Multi-root BFS levels on the device-driven loop.  Same results and the
// same level records as ms_bfs_run without a hook: with level masks on, each
// level's root masks go to the store (grown, or given up for depth rows,
// before the level is issued); `depth` holds rows in internal order.  With
// the batch_graphs option (and no per-kernel profiling), iterations after
// the first replay one captured graph per parity instead of their launches;
// the first one also sizes the plan's buffers, so no capture allocates.
static int ms_bfs_device(gg_graph *g, Work *w, const uint32_t *roots, int n_roots,
                         uint32_t *depth, gg_run_stats *st, MsLevels *res) {
  const Dir &out = out_dir(g), &in = in_dir(g);
  const uint32_t V = g->vertices;
  const size_t words = w->words;
  const Options &opt = options();
  MsdState *s;
  GG_TRY(msd_state(w, &s));
  hipStream_t sm = stream();
  bool masks = opt.level_masks != 0;
  if (masks) GG_TRY(ms_mask_room(w, 0, &masks));
  if (!masks) GG_HIP(hipMemsetAsync(depth, 0xFF, (size_t)V * 64 * 4, sm));
  GG_HIP(hipMemsetAsync(w->ms_vis, 0, (size_t)V * 8, sm));
  GG_HIP(hipMemsetAsync(w->ms_f[0], 0, (size_t)V * 8, sm));
  GG_HIP(hipMemsetAsync(w->ms_f[1], 0, (size_t)V * 8, sm));
  GG_HIP(hipMemsetAsync(w->f[0].bits, 0, words * 8, sm));
  GG_HIP(hipMemsetAsync(w->f[1].bits, 0, words * 8, sm));
  GG_HIP(hipMemsetAsync(s->settled, 0, words * 8, sm));
  GG_HIP(hipMemsetAsync(s->set, 0, (size_t)kScalarSlots * 16 * 8, sm));
  GG_HIP(hipMemcpyAsync(w->ms_roots, roots, 4 * (size_t)n_roots, hipMemcpyHostToDevice, sm));
  volatile MsHostFlag *hf = s->h_flag;
  hf->decided = 0;
  hf->done = 0;
  const uint64_t threshold = g->edges / (uint64_t)opt.switch_divisor;
  hipLaunchKernelGGL(msd_seed_kernel, dim3(1), dim3(kBlock), 0, sm, s->ctl, s->set, s->qc,
                     s->h_flag, w->ms_roots, n_roots, w->ms_vis, w->ms_f[0], w->f[0].bits,
                     masks ? nullptr : depth, out.degree, (int32_t)opt.force_mode, threshold);
  if (masks)
    GG_HIP(hipMemcpyAsync(w->ms_lmask, w->ms_f[0], (size_t)V * 8, hipMemcpyDeviceToDevice, sm));
  const uint64_t full = n_roots == 64 ? ~0ULL : ((1ULL << n_roots) - 1ULL);
  PlanRef plan_ref;  // held until this call returns
  GG_TRY(pr_plan_get(g, &plan_ref));
  const PrPlan *plan = plan_ref.get();
  // fixed launch shapes: the frontier sizes live on the device
  const unsigned bin_grid = compact_grid(words);
  const unsigned push_grid = grid_waves((uint64_t)opt.huge_grid_waves);
  const uint32_t ngroups = (V + 63) / 64;
  const unsigned heavy_grid =
      grid_waves(std::min<uint64_t>(std::max(in.n_chunks, 1u), (uint64_t)opt.huge_grid_waves));
  // several words per wave: one word each left the sparse levels launch-bound
  const unsigned apply_grid = grid_waves(std::min<size_t>(words, 1u << 13));
  const bool use_graphs = opt.batch_graphs && !prof_on();
  for (uint32_t n = 0;; n++) {
    // keep one iteration queued: before issuing iteration n, wait until
    // iteration n-1 has been decided (so iteration n-2 has finished)
    if (n >= 1) {
      uint64_t spins = 0;
      while (!hf->done && hf->decided < n) {
        if ((++spins & 0xFFF) == 0 && hipStreamQuery(sm) == hipSuccess && !hf->done &&
            hf->decided < n)
          return fail(GG_EHIP, "batched BFS: device loop stopped without a decision");
      }
      if (hf->done) break;
    }
    if (n > V + 2) return fail(GG_EHIP, "batched BFS: traversal did not terminate");
    const uint32_t level = n + 1;
    if (masks) {
      bool room;
      GG_TRY(ms_mask_room(w, level, &room));
      if (!room) {  // rows of levels [0, level) from the store, then rows directly
        GG_TRY(ms_rows_from_masks(w, level, nullptr, depth));
        masks = false;
      }
    }
    uint64_t *fin = w->f[n & 1].bits, *fout = w->f[(n + 1) & 1].bits;
    uint64_t *fcur = w->ms_f[n & 1], *fnext = w->ms_f[(n + 1) & 1];
    uint32_t **q = w->f[0].q;
    auto issue = [&]() -> int {
    {
      ProfScope p("ms_bin");
      hipLaunchKernelGGL(msd_bin_kernel, dim3(bin_grid), dim3(kCompactBlock), 0, sm, s->ctl, fin,
                         words, out.degree, (uint32_t)opt.push_small_max,
                         (uint32_t)opt.push_big_max, q[0], q[1], q[2], s->qc);
    }
    {
      ProfScope p("ms_push");
      MsPush op;
      op.fcur = fcur;
      op.vis = w->ms_vis;
      op.fnext = fnext;
      MsdPushArgs a{out.index, out.nbr, q[0], q[1], q[2], s->qc, fout};
      hipLaunchKernelGGL(msd_push_kernel, dim3(push_grid), dim3(kBlock), 0, sm, s->ctl, a, op);
    }
    if (plan) {
      GG_TRY(ms_plan_iteration(g, plan, w, fcur, w->ms_vis, full, fnext, fout, &s->ctl->mode,
                               s->settled));
    } else {
      MsPullArgs pa{in.index, in.nbr, fcur, w->ms_vis, fnext, fout, full, V,
                    in.light_max, (uint32_t)opt.pull_serial_max, ngroups, &s->ctl->mode};
      {
        ProfScope p("ms_pull");
        hipLaunchKernelGGL(ms_pull_group_kernel, dim3(grid_waves(ngroups)), dim3(kBlock), 0, sm,
                           pa);
      }
      if (in.n_chunks) {
        ProfScope p("ms_pull_heavy");
        MsdHeavyArgs h{in.nbr, in.chunk_v, in.chunk_beg, in.chunk_end, in.n_chunks,
                       fcur, w->ms_vis, fnext, fout, full};
        hipLaunchKernelGGL(msd_heavy_kernel, dim3(heavy_grid), dim3(kBlock), 0, sm, s->ctl, h);
      }
    }
    {
      ProfScope p("ms_apply");
      MsdApplyArgs a{fin,     fout, fcur, fnext, w->ms_vis, s->settled, full,
                     masks ? w->ms_lmask : nullptr, masks ? nullptr : depth, out.degree, s->set,
                     words, V};
      hipLaunchKernelGGL(msd_apply_kernel, dim3(apply_grid), dim3(kBlock), 0, sm, s->ctl, a);
    }
    hipLaunchKernelGGL(msd_decide_kernel, dim3(1), dim3(kBlock), 0, sm, s->ctl, s->set, s->qc,
                       s->h_flag);
    GG_HIP(hipGetLastError());
    return GG_OK;
    };
    if (!use_graphs || n == 0) {
      GG_TRY(issue());
      continue;
    }
    MsdGraph &gr = s->graphs[n & 1];
    const uint64_t key[16] = {(uint64_t)fin, (uint64_t)fcur, (uint64_t)(masks ? w->ms_lmask : nullptr),
                              (uint64_t)(masks ? nullptr : depth), full,
                              plan ? plan->serial : 0,
                              (uint64_t)s->ctl, (uint64_t)w->ms_vis, (uint64_t)q[0],
                              (uint64_t)opt.push_small_max, (uint64_t)opt.push_big_max,
                              (uint64_t)opt.huge_grid_waves, (uint64_t)opt.pull_serial_max,
                              (uint64_t)g, (uint64_t)s->settled, (uint64_t)w->pr_vals};
    if (!gr.exec || memcmp(gr.key, key, sizeof(key))) {
      if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
      gr.exec = nullptr;
      hipGraph_t graph = nullptr;
      GG_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
      const int rc = issue();
      const hipError_t e = hipStreamEndCapture(sm, &graph);
      if (rc != GG_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != GG_OK ? rc : fail(GG_EHIP, "batched BFS: graph capture failed");
      }
      const hipError_t ie = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        gr.exec = nullptr;
        return fail(GG_EHIP, "batched BFS: graph instantiation failed");
      }
      memcpy(gr.key, key, sizeof(key));
    }
    GG_HIP(hipGraphLaunch(gr.exec, sm));
  }
  GG_HIP(hipMemcpyAsync(s->h_ctl, s->ctl, sizeof(MsCtl), hipMemcpyDeviceToHost, sm));
  GG_HIP(hipStreamSynchronize(sm));
  const MsCtl &c = *s->h_ctl;
  st->iterations = c.iter;
  st->push_iterations = c.push_iters;
  for (uint32_t i = 0; i < std::min<uint32_t>(c.iter, GG_STATS_RECORDS); i++) {
    st->modes[i] = c.modes[i];
    st->it_active[i] = c.active[i];
    st->it_new[i] = c.fresh[i];
  }
  res->levels = c.level + 1;  // levels 0 .. c.level hold vertices (the last one is empty)
  res->masks = masks;
  return GG_OK;
}

// ------------------------------------------ batched SSSP, device-driven loop
// The batched SSSP's iterations on the same control block and decide as the
// batched BFS: bin (push only), a persistent push over the queues, the pull
// (lane kernels gated through MssPull), then an apply that counts the new
// frontier and clears the spent frontier's bitmap and changed-root masks,
// and the decide.  The host stays one iteration ahead.

// One block: root i's distance 0 and its bit in the root vertex's changed
// mask (mss_seed_kernel), the control block, the count of distinct roots
// and their out-arcs, and the first decision.
__global__ __launch_bounds__(kBlock) void mssd_seed_kernel(MsCtl *c, uint64_t *set, Counters *qc,
                                                           MsHostFlag *hf, const uint32_t *roots,
                                                           int n, float *dist, uint64_t *cm,
                                                           uint64_t *bits, const uint32_t *outdeg,
                                                           int32_t force, uint64_t threshold) {
  const int t = threadIdx.x;
  if (t < n) {
    const uint32_t r = roots[t];
    dist[(size_t)r * 64 + t] = 0.f;
    atomicOr((unsigned long long *)&cm[r], 1ULL << t);
    atomicOr((unsigned long long *)&bits[r >> 6], 1ULL << (r & 63));
    bool first = true;
    for (int j = 0; j < t; j++) first = first && roots[j] != r;
    if (first) {
      atomicAdd((unsigned long long *)&set[0], 1ull);
      atomicAdd((unsigned long long *)&set[1], (unsigned long long)outdeg[r]);
    }
  }
  if (t == 0) {
    c->done = 0;
    c->iter = 0;
    c->mode = kMsDone;
    c->level = 0;
    c->push_iters = 0;
    c->force = force;
    c->threshold = threshold;
  }
  __threadfence();
  __syncthreads();
  msd_decide(c, set, qc, hf);
}

struct MssdPushArgs {
  const uint64_t *index;
  const uint32_t *nbr;
  const float *wt;
  const uint32_t *q0, *q1, *q2;
  const Counters *qc;
  float *dist;
  const uint64_t *cm_cur;
  uint64_t *cm_next, *out_bits;
};

// Persistent grid over the binned queues (sizes on the device): one wave
// per small or mid vertex, then chunks of every huge vertex shared by all
// waves, rotated per vertex.
__global__ __launch_bounds__(kBlock) void mssd_push_kernel(const MsCtl *c, MssdPushArgs a) {
  if (c->mode != kMsPush) return;
  const uint32_t W = gridDim.x * kWavesPerBlock, gw = global_wave();
  const uint32_t n0 = a.qc->n[0], n1 = a.qc->n[1], n2 = a.qc->n[2];
  for (uint32_t i = gw; i < n0 + n1; i += W) {
    const uint32_t u = i < n0 ? a.q0[i] : a.q1[i - n0];
    mss_relax_range(a.nbr, a.wt, a.dist, a.cm_cur, a.cm_next, a.out_bits, u, a.index[u],
                    a.index[u + 1]);
  }
  for (uint32_t h = 0; h < n2; h++) {
    const uint32_t u = a.q2[h];
    const uint64_t b = a.index[u], e = a.index[u + 1];
    const uint64_t nch = (e - b + kChunkEdges - 1) / kChunkEdges;
    const uint32_t rot = (uint32_t)(((uint64_t)h * 2654435761u) % W);
    for (uint64_t ch = (gw + W - rot) % W; ch < nch; ch += W) {
      const uint64_t cb = b + ch * kChunkEdges;
      mss_relax_range(a.nbr, a.wt, a.dist, a.cm_cur, a.cm_next, a.out_bits, u, cb,
                      cb + kChunkEdges < e ? cb + kChunkEdges : e);
    }
  }
}

// One wave per 64-vertex bitmap word: counts the new frontier (vertices and
// out-arcs, for the decide) and clears the spent frontier's word and masks.
__global__ __launch_bounds__(kBlock) void mssd_apply_kernel(const MsCtl *c, uint64_t *fin,
                                                            const uint64_t *fout,
                                                            uint64_t *cm_spent,
                                                            const uint32_t *outdeg, uint64_t *set,
                                                            size_t words) {
  if (c->mode == kMsDone) return;
  const int lane = lane_id();
  const size_t W = (size_t)gridDim.x * kWavesPerBlock;
  uint64_t cnt = 0, edg = 0;
  for (size_t wi = global_wave(); wi < words; wi += W) {
    const uint64_t act = fout[wi], old = fin[wi];
    const uint32_t v = (uint32_t)(wi * 64 + lane);
    if ((act >> lane) & 1ULL) {
      cnt++;
      edg += outdeg[v];
    }
    if ((old >> lane) & 1ULL) cm_spent[v] = 0;
    if (lane == 0 && old) fin[wi] = 0;
  }
  cnt = wave_sum_u64(cnt);
  edg = wave_sum_u64(edg);
  if (lane == 0 && cnt) {
    uint64_t *p = set + (size_t)(global_wave() % kScalarSlots) * 16;
    atomicAdd((unsigned long long *)p, (unsigned long long)cnt);
    atomicAdd((unsigned long long *)(p + 1), (unsigned long long)edg);
  }
}

// Batched SSSP iterations on the device-driven loop: `dist` (rows in
// internal order, 1e9 everywhere) gets the distances; the same iterations,
// modes and frontier sizes as the host-driven loop.
static int mss_device(gg_graph *g, Work *w, const uint32_t *roots, int n_roots, float *dist,
                      gg_run_stats *st) {
  const Dir &out = out_dir(g), &in = in_dir(g);
  const uint32_t V = g->vertices;
  const size_t words = w->words;
  const Options &opt = options();
  MsdState *s;
  GG_TRY(msd_state(w, &s));
  hipStream_t sm = stream();
  GG_HIP(hipMemsetAsync(w->ms_f[0], 0, (size_t)V * 8, sm));
  GG_HIP(hipMemsetAsync(w->ms_f[1], 0, (size_t)V * 8, sm));
  GG_HIP(hipMemsetAsync(w->f[0].bits, 0, words * 8, sm));
  GG_HIP(hipMemsetAsync(w->f[1].bits, 0, words * 8, sm));
  GG_HIP(hipMemsetAsync(s->set, 0, (size_t)kScalarSlots * 16 * 8, sm));
  GG_HIP(hipMemcpyAsync(w->ms_roots, roots, 4 * (size_t)n_roots, hipMemcpyHostToDevice, sm));
  volatile MsHostFlag *hf = s->h_flag;
  hf->decided = 0;
  hf->done = 0;
  const uint64_t threshold = g->edges / (uint64_t)opt.switch_divisor;
  hipLaunchKernelGGL(mssd_seed_kernel, dim3(1), dim3(kBlock), 0, sm, s->ctl, s->set, s->qc,
                     s->h_flag, w->ms_roots, n_roots, dist, w->ms_f[0], w->f[0].bits, out.degree,
                     (int32_t)opt.force_mode, threshold);
  const unsigned bin_grid = compact_grid(words);
  const unsigned push_grid = grid_waves((uint64_t)opt.huge_grid_waves);
  // looping grids: a gated launch of one wave per vertex would cost its
  // dispatch on every push iteration
  const unsigned pull_grid = grid_waves(std::min<uint64_t>(V, 1u << 16));
  const uint64_t hw = std::max<uint64_t>((uint64_t)opt.heavy_waves, 1);
  const unsigned heavy_grid = grid_waves(std::min<uint64_t>(std::max(in.n_chunks, 1u), hw));
  const unsigned apply_grid = grid_waves(std::min<size_t>(words, 1u << 13));
  for (uint32_t n = 0;; n++) {
    // one iteration queued ahead, as in the batched BFS loop
    if (n >= 1) {
      uint64_t spins = 0;
      while (!hf->done && hf->decided < n) {
        if ((++spins & 0xFFF) == 0 && hipStreamQuery(sm) == hipSuccess && !hf->done &&
            hf->decided < n)
          return fail(GG_EHIP, "batched SSSP: device loop stopped without a decision");
      }
      if (hf->done) break;
    }
    if ((uint64_t)n > 64ull * V + 2) return fail(GG_EHIP, "batched SSSP: did not converge");
    uint64_t *fin = w->f[n & 1].bits, *fout = w->f[(n + 1) & 1].bits;
    uint64_t *cm_cur = w->ms_f[n & 1], *cm_next = w->ms_f[(n + 1) & 1];
    uint32_t **q = w->f[0].q;
    {
      ProfScope p("mss_bin");
      hipLaunchKernelGGL(msd_bin_kernel, dim3(bin_grid), dim3(kCompactBlock), 0, sm, s->ctl, fin,
                         words, out.degree, (uint32_t)opt.push_small_max,
                         (uint32_t)opt.push_big_max, q[0], q[1], q[2], s->qc);
    }
    {
      ProfScope p("mss_push");
      MssdPushArgs a{out.index, out.nbr, out.w, q[0], q[1], q[2], s->qc, dist, cm_cur, cm_next,
                     fout};
      hipLaunchKernelGGL(mssd_push_kernel, dim3(push_grid), dim3(kBlock), 0, sm, s->ctl, a);
    }
    {
      ProfScope p("mss_pull");
      MssPull op{dist, cm_cur, cm_next, fout, &s->ctl->mode};
      LaneArgs a{in.index, in.nbr, in.w, nullptr, V, in.light_max};
      hipLaunchKernelGGL(lane_vertex_kernel<MssPull>, dim3(pull_grid), dim3(kBlock), 0, sm, a, op);
      if (in.n_chunks) {
        HeavyArgs h{in.nbr, in.w, in.chunk_v, in.chunk_h, in.chunk_beg, in.chunk_end, in.n_chunks,
                    fout, nullptr, nullptr};
        hipLaunchKernelGGL(lane_heavy_kernel<MssPull>, dim3(heavy_grid), dim3(kBlock), 0, sm, h,
                           op);
      }
    }
    hipLaunchKernelGGL(mssd_apply_kernel, dim3(apply_grid), dim3(kBlock), 0, sm, s->ctl, fin, fout,
                       cm_cur, out.degree, s->set, words);
    hipLaunchKernelGGL(msd_decide_kernel, dim3(1), dim3(kBlock), 0, sm, s->ctl, s->set, s->qc,
                       s->h_flag);
    GG_HIP(hipGetLastError());
  }
  GG_HIP(hipMemcpyAsync(s->h_ctl, s->ctl, sizeof(MsCtl), hipMemcpyDeviceToHost, sm));
  GG_HIP(hipStreamSynchronize(sm));
  const MsCtl &c = *s->h_ctl;
  st->iterations = c.iter;
  st->push_iterations = c.push_iters;
  for (uint32_t i = 0; i < std::min<uint32_t>(c.iter, GG_STATS_RECORDS); i++) {
    st->modes[i] = c.modes[i];
    st->it_active[i] = c.active[i];
    st->it_new[i] = c.fresh[i];
  }
  return GG_OK;
}

// Sum over roots of the out-degrees of the vertices each reached.
static int ms_reach(gg_graph *g, Work *w, uint64_t *m) {
  GG_HIP(hipMemsetAsync(w->exam, 0, 1024 * 8, stream()));
  hipLaunchKernelGGL(ms_reach_kernel, dim3(grid1(g->vertices, 4096)), dim3(kBlock), 0, stream(),
                     w->ms_vis, out_dir(g).degree, g->vertices, w->exam);
  GG_HIP(hipMemcpyAsync(w->h_exam, w->exam, 1024 * 8, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipMemsetAsync(w->exam, 0, 1024 * 8, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  uint64_t s = 0;
  for (int k = 0; k < 64; k++) s += w->h_exam[k * 16];
  *m = s;
  return GG_OK;
}

}  // namespace gg

namespace gg {
__global__ void gg_copy_kernel(uint4 *dst, const uint4 *src, size_t n) {
  size_t s = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s) dst[i] = src[i];
}
}  // namespace gg

using namespace gg;

extern "C" {

int gg_device_copy(void *d_dst, const void *d_src, size_t bytes) {
  GG_TRY(ensure_init());
  if (!d_dst || !d_src || (bytes & 15) || (((uintptr_t)d_dst | (uintptr_t)d_src) & 15))
    return fail(GG_EINVAL, "gg_device_copy: 16-byte aligned pointers and size required");
  size_t n = bytes / 16;
  if (n) {
    ProfScope p("copy");
    hipLaunchKernelGGL(gg_copy_kernel, dim3(grid1(n, 8192)), dim3(kBlock), 0, stream(),
                       (uint4 *)d_dst, (const uint4 *)d_src, n);
  }
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

int gg_subset_count(const gg_subset *s, uint64_t *count) {
  if (!s || !count) return fail(GG_EINVAL, "gg_subset_count: null argument");
  GG_TRY(ensure_init());
  unsigned long long *d = nullptr;
  GG_TRY(device_alloc((void **)&d, 8));
  hipMemsetAsync(d, 0, 8, stream());
  hipLaunchKernelGGL(popcount_kernel, dim3(grid1(s->words, 4096)), dim3(kBlock), 0, stream(), s->d,
                     s->words, d);
  unsigned long long h = 0;
  hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, stream());
  hipError_t e = hipStreamSynchronize(stream());
  device_free(d);
  if (e != hipSuccess) return fail(GG_EHIP, hipGetErrorString(e));
  *count = h;
  return GG_OK;
}

int gg_pagerank(gg_graph *g, int iterations, double *d_curr_out, double *h_deltas,
                gg_run_stats *stats) {
  GG_TRY(check_graph(g));
  if (iterations < 0 || iterations > 4000) return fail(GG_EINVAL, "iterations out of range");
  if (!d_curr_out) return fail(GG_EINVAL, "null output");
  if (options().force_mode == 1)
    return fail(GG_EINVAL, "PageRank runs dense: its active set is every vertex");
  Work *w;
  GG_TRY(work_get(g, &w));
  Engine en(g, w);
  const Dir &out = out_dir(g), &in = in_dir(g);
  uint32_t V = g->vertices;
  DeviceTimer tm;
  double *curr = g->reordered ? (double *)w->res_a : d_curr_out;
  double *next = (double *)w->msg;
  if (w->dacc_cap < iterations + 1) {
    int cap = std::max(iterations + 1, 64);
    double *a = nullptr, *s = nullptr;
    GG_TRY(device_alloc((void **)&a, sizeof(double) * kScalarStride * cap));
    int r = device_alloc((void **)&s, sizeof(double) * cap);
    if (r != GG_OK) {
      device_free(a);
      return r;
    }
    device_free(w->dacc);
    device_free(w->dsum);
    w->dacc = a;
    w->dsum = s;
    w->dacc_cap = cap;
  }
  // relabelled graphs: destination bins summed in LDS (prplan.hip)
  PlanRef plan_ref;  // held until this call returns
  GG_TRY(pr_plan_get(g, &plan_ref));
  const PrPlan *plan = plan_ref.get();
  hipLaunchKernelGGL(pr_init_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), curr, out.degree, V);
  GG_HIP(hipMemsetAsync(w->dacc, 0, sizeof(double) * kScalarStride * (iterations + 1), stream()));
  for (int it = 0; it < iterations; it++) {
    // active = every vertex, so active_edges = |E| and the step is dense
    // (core/graph.hpp:1509).
    double *acc = w->dacc + (size_t)(it + 1) * kScalarStride;
    PrPull op{{}, curr, next, out.degree, it == iterations - 1};
    if (plan) {
      GG_TRY(pr_plan_iteration(g, plan, w, op, acc));
    } else {
      GG_TRY(en.launch_pull_group(in, op, acc, "pr_pull"));
      PrHeavy hop{{}, curr, w->hacc};
      GG_TRY(en.launch_pull_heavy(in, hop, "pr_heavy"));
      if (in.n_heavy)
        hipLaunchKernelGGL(pr_heavy_finish_kernel, dim3((in.n_heavy + kBlock - 1) / kBlock),
                           dim3(kBlock), 0, stream(), in.heavy_v, in.n_heavy, w->hacc, op, acc);
    }
    if (en.st.iterations < GG_STATS_RECORDS) en.st.modes[en.st.iterations] = 1;
    en.st.iterations++;
    en.st.edges_examined += g->arcs;
    std::swap(curr, next);
  }
  if (g->reordered)
    hipLaunchKernelGGL(unpermute_kernel<double>, dim3(grid1(V)), dim3(kBlock), 0, stream(), curr,
                       d_curr_out, g->iperm, V);
  else if (curr != d_curr_out)
    GG_HIP(hipMemcpyAsync(d_curr_out, curr, sizeof(double) * V, hipMemcpyDeviceToDevice, stream()));
  en.st.device_ms = tm.stop();
  GG_HIP(hipGetLastError());
  if (h_deltas) {
    std::vector<double> acc(iterations + 1);
    hipLaunchKernelGGL(scalar_reduce_kernel, dim3(iterations + 1), dim3(kBlock), 0, stream(),
                       w->dacc, w->dsum);
    GG_HIP(hipMemcpyAsync(acc.data(), w->dsum, sizeof(double) * (iterations + 1),
                          hipMemcpyDeviceToHost, stream()));
    GG_HIP(hipStreamSynchronize(stream()));
    h_deltas[0] = 1.0;  // every vertex contributes 1, divided by |V|
    for (int it = 1; it <= iterations; it++) h_deltas[it] = acc[it] / V;
  }
  en.st.m_reach = g->arcs;
  en.st.reached = V;
  put_stats(stats, en.st);
  return GG_OK;
}

int gg_cc(gg_graph *g, uint32_t *d_label, uint32_t *h_frontier, int cap, gg_run_stats *stats) {
  GG_TRY(check_graph(g));
  if (!d_label) return fail(GG_EINVAL, "null output");
  Work *w;
  GG_TRY(work_get(g, &w));
  Engine en(g, w);
  en.lazy_queues = true;  // queues only before a push
  const Dir &out = out_dir(g), &in = in_dir(g);
  uint32_t V = g->vertices;
  DeviceTimer tm;
  uint32_t *out_label = d_label;
  if (g->reordered) {
    // labels are original ids, so the propagated minimum is the reference's
    d_label = (uint32_t *)w->res_a;
    hipLaunchKernelGGL(copy_u32_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), g->iperm,
                       d_label, V);
  } else {
    hipLaunchKernelGGL(iota_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), d_label, V);
  }
  uint32_t *msg = (uint32_t *)w->msg;
  fill<uint32_t>(msg, 0xFFFFFFFFu, V);
  // active_in->fill() (cc.cpp:28), exact tail as in core/bitmap.hpp:42-52
  hipLaunchKernelGGL(bitmap_fill_kernel, dim3(grid1(w->words)), dim3(kBlock), 0, stream(),
                     en.fin().bits, w->words, V);
  // relabelled graph: dense steps on the destination-bin plan (min in LDS),
  // which cost a few ps per arc: push longer (set before the first count,
  // which decides whether to build queues with the loop's switch)
  PlanRef plan_ref;  // held until this call returns
  GG_TRY(pr_plan_get(g, &plan_ref));
  const PrPlan *plan = plan_ref.get();
  if (plan) en.divisor = (uint64_t)options().plan_switch_divisor;
  GG_TRY(en.compact_in(out.degree));
  int lv = 0;
  while (en.in.total > 0) {
    if (h_frontier && lv < cap) h_frontier[lv] = en.in.total;
    lv++;
    bool sp = en.sparse(en.in.edges);
    if (sp) {
      uint64_t n = (uint64_t)en.in.n[0] + en.in.n[1] + en.in.n[2];
      hipLaunchKernelGGL(snapshot_kernel, dim3(grid1(n)), dim3(kBlock), 0, stream(), d_label,
                         en.fin(), en.in.n[0], en.in.n[1], en.in.n[2]);
      CcPush op;
      op.label = d_label;
      GG_TRY(en.launch_push(out, op, true));
    } else if (plan) {
      uint32_t *nlab = (uint32_t *)w->msg;
      hipLaunchKernelGGL(cc_snapshot_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), d_label,
                         nlab, V);
      GG_TRY(cc_plan_iteration(g, plan, w, nlab, d_label, en.fout().bits));
    } else {
      CcPull op{{}, d_label, msg};
      GG_TRY(en.launch_pull_group(in, op, nullptr));
      GG_TRY(en.launch_pull_heavy(in, op));
      hipLaunchKernelGGL(cc_apply_kernel, dim3(grid_vertex_rounds(V)), dim3(kBlock), 0, stream(),
                         d_label, msg, V, en.fout().bits);
    }
    GG_TRY(en.finish_iteration(sp, out.degree, nullptr));
  }
  if (g->reordered)
    hipLaunchKernelGGL(unpermute_kernel<uint32_t>, dim3(grid1(V)), dim3(kBlock), 0, stream(),
                       d_label, out_label, g->iperm, V);
  en.st.device_ms = tm.stop();
  put_stats(stats, en.st);
  return GG_OK;
}

int gg_sssp(gg_graph *g, uint32_t root, float *d_dist, uint32_t *h_frontier, int cap,
            gg_run_stats *stats) {
  GG_TRY(check_graph(g));
  if (root >= g->vertices) return fail(GG_EINVAL, "root out of range");
  if (!g->weighted) return fail(GG_EINVAL, "SSSP needs a weighted graph");
  if (!d_dist) return fail(GG_EINVAL, "null output");
  Work *w;
  GG_TRY(work_get(g, &w));
  Engine en(g, w);
  en.lazy_queues = true;  // queues only before a push
  const Dir &out = out_dir(g), &in = in_dir(g);
  uint32_t V = g->vertices;
  DeviceTimer tm;
  float *out_dist = d_dist;
  if (g->reordered) {
    d_dist = (float *)w->res_a;
    root = g->to_internal(root);
  }
  fill<float>(d_dist, 1e9f, V);  // sssp.cpp:163-164
  hipLaunchKernelGGL((put2_kernel<float, float>), dim3(1), dim3(1), 0, stream(), d_dist, 0.f,
                     (float *)nullptr, 0.f, root);
  uint32_t *msg = (uint32_t *)w->msg;
  fill<uint32_t>(msg, kInfBits, V);
  // relabelled graph: dense steps on the destination-bin plan (its arcs carry
  // their weights), which cost a few ps per arc: push longer.  (Set before
  // the first frontier is counted: the count decides whether to build queues
  // with the same switch as the loop.)
  PlanRef plan_ref;  // held until this call returns
  GG_TRY(pr_plan_get(g, &plan_ref));
  const PrPlan *plan = plan_ref.get();
  if (plan && !plan->hot_w) plan = nullptr;
  if (plan) en.divisor = (uint64_t)options().plan_switch_divisor;
  GG_TRY(en.seed(root, out.degree, nullptr));
  int lv = 0;
  while (en.in.total > 0) {
    if (h_frontier && lv < cap) h_frontier[lv] = en.in.total;
    lv++;
    bool sp = en.sparse(en.in.edges);
    if (sp) {
      uint64_t n = (uint64_t)en.in.n[0] + en.in.n[1] + en.in.n[2];
      hipLaunchKernelGGL(snapshot_kernel, dim3(grid1(n)), dim3(kBlock), 0, stream(),
                         (const uint32_t *)d_dist, en.fin(), en.in.n[0], en.in.n[1], en.in.n[2]);
      SsspPush op;
      op.dist = d_dist;
      GG_TRY(en.launch_push(out, op, true));
    } else if (plan) {
      uint32_t *snap = (uint32_t *)w->msg;
      hipLaunchKernelGGL(sssp_snapshot_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), d_dist,
                         snap, V);
      GG_TRY(sssp_plan_iteration(g, plan, w, snap, d_dist, en.fout().bits));
    } else {
      SsspPull op{{}, d_dist, msg};
      GG_TRY(en.launch_pull_group(in, op, nullptr));
      GG_TRY(en.launch_pull_heavy(in, op));
      hipLaunchKernelGGL(sssp_apply_kernel, dim3(grid_vertex_rounds(V)), dim3(kBlock), 0, stream(),
                         d_dist, msg, V, en.fout().bits);
    }
    GG_TRY(en.finish_iteration(sp, out.degree, nullptr));
  }
  if (g->reordered)
    hipLaunchKernelGGL(unpermute_kernel<float>, dim3(grid1(V)), dim3(kBlock), 0, stream(), d_dist,
                       out_dist, g->iperm, V);
  en.st.device_ms = tm.stop();
  // out-degree sum of the reached vertices (outside the timed compute, as
  // the reference's SSSP has no such count)
  GG_HIP(hipMemsetAsync(w->exam, 0, 1024 * 8, stream()));
  hipLaunchKernelGGL(sssp_reach_kernel, dim3(grid1(V, 4096)), dim3(kBlock), 0, stream(), d_dist,
                     out.degree, V, w->exam);
  GG_HIP(hipMemcpyAsync(w->h_exam, w->exam, 1024 * 8, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipMemsetAsync(w->exam, 0, 1024 * 8, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  uint64_t reached_edges = 0;
  for (int k = 0; k < 64; k++) reached_edges += w->h_exam[k * 16];
  en.st.m_reach = reached_edges;
  put_stats(stats, en.st);
  return GG_OK;
}

int gg_bc(gg_graph *g, uint32_t root, double *d_dep, double *d_inv_paths, uint32_t *h_frontier,
          int cap, gg_run_stats *stats) {
  GG_TRY(check_graph(g));
  if (root >= g->vertices) return fail(GG_EINVAL, "root out of range");
  if (!d_dep || !d_inv_paths) return fail(GG_EINVAL, "null output");
  Work *w;
  GG_TRY(work_get(g, &w));
  Engine en(g, w);
  const Dir &out = out_dir(g), &in = in_dir(g);
  uint32_t V = g->vertices;
  DeviceTimer tm;
  double *out_dep = d_dep, *out_inv = d_inv_paths;
  if (g->reordered) {
    d_dep = (double *)w->res_a;
    d_inv_paths = (double *)w->res_b;
    root = g->to_internal(root);
  }
  double *sigma = d_inv_paths;  // bc.cpp:103: inv_num_paths aliases num_paths
  uint32_t *depth = (uint32_t *)w->msg;
  fill<double>(sigma, 0.0, V);
  fill<uint32_t>(depth, kUnset, V);
  hipLaunchKernelGGL((put2_kernel<double, uint32_t>), dim3(1), dim3(1), 0, stream(), sigma, 1.0,
                     depth, 0u, root);
  GG_HIP(hipMemsetAsync(w->visited, 0, w->words * 8, stream()));
  GG_TRY(en.seed(root, out.degree, w->visited));
  int lv = 0;
  uint64_t reached = 1;
  // forward (bc.cpp:48-101)
  uint32_t level = 1;
  for (; en.in.total > 0; level++) {
    if (h_frontier && lv < cap) h_frontier[lv] = en.in.total;
    lv++;
    bool sp = en.sparse(en.in.edges);
    if (sp) {
      BcFwdPush op{sigma, w->visited, depth, level};
      GG_TRY(en.launch_push(out, op, false));
    } else {
      BcFwdPull op{sigma, w->visited, en.fin().bits, depth, level};
      GG_TRY(en.launch_pull_group(in, op, nullptr));
      GG_TRY(en.launch_pull_heavy(in, op));
    }
    GG_TRY(en.finish_iteration(sp, out.degree, w->visited));
    reached += en.in.total;
  }
  // levels 0 .. max_level are non-empty; the reference's trailing empty
  // level costs one no-op process_edges in the backward pass.
  uint32_t max_level = level - 2;
  hipLaunchKernelGGL(bc_inv_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), sigma, d_dep, V);
  GG_TRY(en.note_iteration(true));
  // backward over the transposed graph (bc.cpp:121-168)
  for (uint32_t L = max_level;; L--) {
    GG_TRY(en.clear_bits(en.fin()));
    hipLaunchKernelGGL(bc_level_kernel, dim3(grid_vertex_rounds(V)), dim3(kBlock), 0, stream(),
                       depth, d_dep, sigma, V, L, en.fin().bits);
    if (L == 0) break;
    GG_TRY(en.compact_in(in.degree));  // transposed out-degree = in_degree
    bool sp = en.sparse(en.in.edges);
    if (sp) {
      BcBwdPush op{{}, d_dep, depth, L - 1};
      GG_TRY(en.launch_push(in, op, false));
    } else {
      BcBwdPull op{{}, d_dep, depth, en.fin().bits, L - 1};
      GG_TRY(en.launch_pull_group(out, op, nullptr));
      GG_TRY(en.launch_pull_heavy(out, op));
    }
    GG_TRY(en.note_iteration(sp));
  }
  hipLaunchKernelGGL(bc_final_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), d_dep, sigma, V);
  if (g->reordered) {
    hipLaunchKernelGGL(unpermute_kernel<double>, dim3(grid1(V)), dim3(kBlock), 0, stream(), d_dep,
                       out_dep, g->iperm, V);
    hipLaunchKernelGGL(unpermute_kernel<double>, dim3(grid1(V)), dim3(kBlock), 0, stream(), sigma,
                       out_inv, g->iperm, V);
  }
  en.st.device_ms = tm.stop();
  GG_HIP(hipGetLastError());
  en.st.reached = reached;
  put_stats(stats, en.st);
  return GG_OK;
}

}  // extern "C"

namespace gg {

int bfs_batch_run(gg_graph *g, const uint32_t *h_roots, int n_roots, uint32_t *d_depth,
                  gg_run_stats *stats, const void **rows) {
  GG_TRY(check_graph(g));
  const bool keep = d_depth == nullptr;
  uint32_t roots[64];
  GG_TRY(batch_roots(g, h_roots, n_roots, roots));
  Work *w;
  GG_TRY(work_get(g, &w));
  GG_TRY(ms_alloc(g, w, false));
  Engine en(g, w);
  uint32_t V = g->vertices;
  DeviceTimer tm;
  uint32_t *depth = g->reordered || keep ? (uint32_t *)w->ms_tmp : d_depth;
  MsLevels lv;
  if (options().batch_device_loop)
    GG_TRY(ms_bfs_device(g, w, roots, n_roots, depth, &en.st, &lv));
  else
    GG_TRY(ms_bfs_run(
        g, w, en, roots, n_roots, depth,
        [](uint32_t, const uint64_t *, const uint64_t *) { return GG_OK; }, false, &lv));
  uint64_t m = 0;
  GG_TRY(ms_reach(g, w, &m));
  if (rows) *rows = depth;
  if (lv.masks)  // every row once, in internal order (kept) or in original ids
    GG_TRY(ms_rows_from_masks(w, lv.levels, keep || !g->reordered ? nullptr : g->iperm,
                              keep ? depth : d_depth));
  else if (g->reordered && !keep)
    hipLaunchKernelGGL(unpermute_rows_kernel<uint32_t>, dim3(grid1((uint64_t)V * 64)), dim3(kBlock),
                       0, stream(), depth, d_depth, g->iperm, V);
  en.st.device_ms = tm.stop();
  GG_HIP(hipGetLastError());
  en.st.m_reach = m;
  en.st.reached = (uint64_t)n_roots;
  put_stats(stats, en.st);
  return GG_OK;
}

int sssp_batch_run(gg_graph *g, const uint32_t *h_roots, int n_roots, float *d_dist,
                   gg_run_stats *stats, const void **rows) {
  GG_TRY(check_graph(g));
  if (!g->weighted) return fail(GG_EINVAL, "SSSP needs a weighted graph");
  const bool keep = d_dist == nullptr;
  uint32_t roots[64];
  GG_TRY(batch_roots(g, h_roots, n_roots, roots));
  Work *w;
  GG_TRY(work_get(g, &w));
  GG_TRY(ms_alloc(g, w, false));
  Engine en(g, w);
  const Dir &out = out_dir(g), &in = in_dir(g);
  uint32_t V = g->vertices;
  DeviceTimer tm;
  float *dist = g->reordered || keep ? (float *)w->ms_tmp : d_dist;
  if (rows) *rows = dist;
  fill<uint32_t>((uint32_t *)dist, kInfBits, (uint64_t)V * 64);  // 1e9 (sssp.cpp:163)
  if (options().batch_device_loop) {
    GG_TRY(mss_device(g, w, roots, n_roots, dist, &en.st));
  } else {
    GG_HIP(hipMemsetAsync(w->ms_f[0], 0, (size_t)V * 8, stream()));
    GG_HIP(hipMemcpyAsync(w->ms_roots, roots, 4 * (size_t)n_roots, hipMemcpyHostToDevice,
                          stream()));
    GG_TRY(en.clear_bits(en.fin()));
    hipLaunchKernelGGL(mss_seed_kernel, dim3(1), dim3(1), 0, stream(), w->ms_roots, n_roots, dist,
                       w->ms_f[0], en.fin().bits);
    GG_TRY(en.compact_in(out.degree));
    int fc = 0;
    while (en.in.total > 0) {
      uint64_t *cm_cur = w->ms_f[fc], *cm_next = w->ms_f[fc ^ 1];
      GG_HIP(hipMemsetAsync(cm_next, 0, (size_t)V * 8, stream()));
      bool sp = en.sparse(en.in.edges);
      if (sp) {
        uint32_t waves2 = en.in.n[2] ? (uint32_t)options().huge_grid_waves : 0;
        uint64_t waves = (uint64_t)en.in.n[0] + en.in.n[1] + waves2;
        ProfScope p("mss_push");
        hipLaunchKernelGGL(mss_push_kernel, dim3(grid_waves(waves)), dim3(kBlock), 0, stream(),
                           out.index, out.nbr, out.w, en.fin().q[0], en.in.n[0], en.fin().q[1],
                           en.in.n[1], en.fin().q[2], en.in.n[2], waves2, dist, cm_cur, cm_next,
                           en.fout().bits);
      } else {
        MssPull op{dist, cm_cur, cm_next, en.fout().bits, nullptr};
        ProfScope p("mss_pull");
        LaneArgs a{in.index, in.nbr, in.w, nullptr, V, in.light_max};
        hipLaunchKernelGGL(lane_vertex_kernel<MssPull>, dim3(grid_waves(V)), dim3(kBlock), 0,
                           stream(), a, op);
        if (in.n_chunks) {
          HeavyArgs h{in.nbr,       in.w,           in.chunk_v, in.chunk_h, in.chunk_beg,
                      in.chunk_end, in.n_chunks, en.fout().bits, nullptr};
          hipLaunchKernelGGL(lane_heavy_kernel<MssPull>, dim3(grid_waves(in.n_chunks)),
                             dim3(kBlock), 0, stream(), h, op);
        }
      }
      GG_TRY(en.finish_iteration(sp, out.degree, nullptr));
      fc ^= 1;
    }
  }
  if (g->reordered && !keep)
    hipLaunchKernelGGL(unpermute_rows_kernel<float>, dim3(grid1((uint64_t)V * 64)), dim3(kBlock), 0,
                       stream(), dist, d_dist, g->iperm, V);
  en.st.device_ms = tm.stop();
  GG_HIP(hipGetLastError());
  en.st.reached = (uint64_t)n_roots;
  put_stats(stats, en.st);
  return GG_OK;
}

int bc_batch_run(gg_graph *g, const uint32_t *h_roots, int n_roots, double *d_dep,
                 double *d_inv_paths, gg_run_stats *stats, const void **rows) {
  GG_TRY(check_graph(g));
  const bool keep = d_dep == nullptr;
  uint32_t roots[64];
  GG_TRY(batch_roots(g, h_roots, n_roots, roots));
  Work *w;
  GG_TRY(work_get(g, &w));
  GG_TRY(ms_alloc(g, w, true, keep));
  Engine en(g, w);
  const Dir &out = out_dir(g), &in = in_dir(g);
  uint32_t V = g->vertices;
  const uint64_t n = (uint64_t)V * 64;
  DeviceTimer tm;
  uint32_t *depth = (uint32_t *)w->ms_tmp;
  // (kept rows are read back per root index: those keep the caller's order)
  LaneMap lanes = identity_lanes();
  if (!keep)  // the depth rows' buffer is free until the forward pass
    GG_TRY(order_lanes(out_dir(g), roots, n_roots, (uint64_t *)w->ms_tmp, (size_t)V * 64 * 4,
                       &lanes));
  double *sigma = g->reordered || keep ? (double *)w->mb_a : d_inv_paths;
  double *dd = g->reordered || keep ? (double *)w->mb_b : d_dep;
  if (rows) *rows = dd;
  // rows that a root can reach: vertices up to the last with in-arcs (the
  // relabeling puts the vertices without arcs last) and the roots
  uint64_t live_v = std::min<uint64_t>(V, (uint64_t)in.live_groups * 64);
  for (int i = 0; i < n_roots; i++) live_v = std::max<uint64_t>(live_v, (uint64_t)roots[i] + 1);
  const uint64_t live = live_v * 64;
  fill<double>(sigma, 0.0, live);
  GG_HIP(hipMemcpyAsync(w->ms_roots, roots, 4 * (size_t)n_roots, hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(msbc_seed_sigma_kernel, dim3(1), dim3(1), 0, stream(), w->ms_roots, n_roots,
                     sigma);
  // forward: the multi-root BFS levels, with sigma accumulated per level
  uint32_t levels = 0;
  auto hook = [&](uint32_t level, const uint64_t *fprev, const uint64_t *fnew) -> int {
    GG_TRY(ms_store_level(w, en.fin().bits, level));
    levels = level;
    if (level == 0) return GG_OK;
    MsbcSigma op{sigma, fprev, fnew};
    ProfScope p("msbc_sigma");
    for (int b = 0; b < 3; b++) {
      if (!en.in.n[b]) continue;
      LaneArgs a{in.index, in.nbr, nullptr, en.fin().q[b], en.in.n[b], in.light_max};
      hipLaunchKernelGGL(lane_vertex_kernel<MsbcSigma>, dim3(grid_waves(en.in.n[b])), dim3(kBlock),
                         0, stream(), a, op);
    }
    if (in.n_chunks && en.in.total) {
      HeavyArgs h{in.nbr, nullptr, in.chunk_v, in.chunk_h, in.chunk_beg, in.chunk_end, in.n_chunks,
                  nullptr, nullptr};
      hipLaunchKernelGGL(lane_heavy_kernel<MsbcSigma>, dim3(grid_waves(in.n_chunks)), dim3(kBlock),
                         0, stream(), h, op);
    }
    return GG_OK;
  };
  // (host-driven: sigma runs on the compacted queues of each level; on the
  // device-driven loop, launches over every vertex were measured slower)
  MsLevels lv;
  GG_TRY(ms_bfs_run(g, w, en, roots, n_roots, depth, hook, true, &lv));
  const bool masks = lv.masks;
  hipLaunchKernelGGL(msbc_inv_kernel, dim3(grid1(live)), dim3(kBlock), 0, stream(), sigma, dd,
                     live);
  // backward: level L-1 pulls D from level L along out-edges (the reference
  // pushes along the transposed graph, bc.cpp:121-168)
  for (uint32_t L = levels; L >= 1; L--) {
    const uint64_t *lb = w->ms_levels + (size_t)(L - 1) * w->words;
    GG_HIP(hipMemcpyAsync(en.fin().bits, lb, w->words * 8, hipMemcpyDeviceToDevice, stream()));
    GG_TRY(en.compact_in(out.degree));
    MsbcBack op{dd, depth, lb, w->ms_levels + (size_t)L * w->words,
                masks ? w->ms_lmask + (size_t)(L - 1) * V : nullptr,
                masks ? w->ms_lmask + (size_t)L * V : nullptr, L};
    ProfScope p("msbc_back");
    for (int b = 0; b < 3; b++) {
      if (!en.in.n[b]) continue;
      LaneArgs a{out.index, out.nbr, nullptr, en.fin().q[b], en.in.n[b], out.light_max};
      hipLaunchKernelGGL(lane_vertex_kernel<MsbcBack>, dim3(grid_waves(en.in.n[b])), dim3(kBlock),
                         0, stream(), a, op);
    }
    // the level's vertices are binned by out-degree: with the big bin empty
    // and its bound at most the heavy-list bound, no heavy list is active
    if (out.n_chunks && (en.in.n[2] || en.big_max > out.light_max)) {
      HeavyArgs h{out.nbr, nullptr, out.chunk_v, out.chunk_h, out.chunk_beg, out.chunk_end,
                  out.n_chunks, nullptr, nullptr};
      hipLaunchKernelGGL(lane_heavy_kernel<MsbcBack>, dim3(grid_waves(out.n_chunks)), dim3(kBlock),
                         0, stream(), h, op);
    }
    GG_TRY(en.note_iteration(false));
  }
  if (g->reordered && !keep)
    hipLaunchKernelGGL(msbc_final_unpermute_kernel, dim3(grid1(n)), dim3(kBlock), 0, stream(), dd,
                       sigma, g->iperm, V, live, d_dep, d_inv_paths, lanes);
  else
    hipLaunchKernelGGL(msbc_final_kernel, dim3(grid1(n)), dim3(kBlock), 0, stream(), dd, sigma,
                       live, n, lanes);
  en.st.device_ms = tm.stop();
  GG_HIP(hipGetLastError());
  en.st.reached = (uint64_t)n_roots;
  put_stats(stats, en.st);
  return GG_OK;
}

}  // namespace gg

extern "C" {

int gg_bfs_batch(gg_graph *g, const uint32_t *h_roots, int n_roots, uint32_t *d_depth,
                 gg_run_stats *stats) {
  if (!d_depth) return fail(GG_EINVAL, "gg_bfs_batch: null output");
  return bfs_batch_run(g, h_roots, n_roots, d_depth, stats, nullptr);
}

int gg_sssp_batch(gg_graph *g, const uint32_t *h_roots, int n_roots, float *d_dist,
                  gg_run_stats *stats) {
  if (!d_dist) return fail(GG_EINVAL, "gg_sssp_batch: null output");
  return sssp_batch_run(g, h_roots, n_roots, d_dist, stats, nullptr);
}

int gg_bc_batch(gg_graph *g, const uint32_t *h_roots, int n_roots, double *d_dep,
                double *d_inv_paths, gg_run_stats *stats) {
  if (!d_dep || !d_inv_paths) return fail(GG_EINVAL, "gg_bc_batch: null output");
  return bc_batch_run(g, h_roots, n_roots, d_dep, d_inv_paths, stats, nullptr);
}

}  // extern "C"

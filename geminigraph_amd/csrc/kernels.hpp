// Device building blocks of the edge engine: wave primitives, frontier
// activation and compaction, and the degree-binned push and group/heavy pull
// skeletons that every toolkit op plugs into.
//
// Mapping of the reference's OpenMP engine (core/graph.hpp:1406-1867) onto
// CDNA4 waves:
//   * a 64-vertex bitmap word (WORD_OFFSET) is one wavefront: lane i owns
//     vertex 64*word + i;
//   * edge kernels only set bits of the output frontier (atomic OR,
//     core/bitmap.hpp:56-58); a compaction pass then turns the bitmap into
//     three queues binned by out-degree, with one atomic per workgroup, and
//     produces the popcount and out-degree sum that drive the push/pull
//     switch (core/graph.hpp:1503-1509);
//   * sparse (push) mode consumes those queues: small lists are expanded by
//     an in-wave load-balanced search (all 64 lanes walk consecutive edges of
//     up to 64 vertices), mid-size lists by one wave per vertex, huge lists by
//     every wave of the grid;
//   * dense (pull) mode sweeps the CSC one 64-vertex group per wave with the
//     same in-wave load balancing and a segmented wave reduction, while
//     vertices above the group limit are cut into fixed chunks at build time
//     (the GPU form of tune_chunks, core/graph.hpp:1341-1404) that combine
//     through one atomic per chunk -- the split-partials contract of
//     dense_signal/dense_slot (core/graph.hpp:1767-1770);
//   * emit + slot of the reference is a direct update of the destination
//     state with native HBM atomics (atomicCAS, atomicMin, f64 atomicAdd).
#pragma once

#include "gg_internal.hpp"

namespace gg {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ULL << __lane_id()) - 1ULL; }
__device__ __forceinline__ int wave_in_block() { return (int)(threadIdx.x >> 6); }
__device__ __forceinline__ uint32_t global_wave() {
  return blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
}
__device__ __forceinline__ bool test_bit(const uint64_t *bits, uint32_t v) {
  return (bits[v >> 6] >> (v & 63)) & 1ULL;
}
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Adjacency arrays are streamed once per sweep: load them non-temporally so
// they do not evict the gathered vertex state (the hot sources) from L2.
template <class T>
__device__ __forceinline__ T stream_load(const T *p) {
  return __builtin_nontemporal_load(p);
}
// Orders LDS traffic between lanes of one wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src) {
  uint32_t lo = __shfl((uint32_t)x, src), hi = __shfl((uint32_t)(x >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double shfl_f64(double x, int src) {
  return __longlong_as_double((long long)shfl_u64((uint64_t)__double_as_longlong(x), src));
}
__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t x, int d) {
  uint32_t lo = __shfl_up((uint32_t)x, d), hi = __shfl_up((uint32_t)(x >> 32), d);
  return ((uint64_t)hi << 32) | lo;
}

template <class T> struct Shfl;
template <> struct Shfl<uint32_t> {
  static __device__ __forceinline__ uint32_t at(uint32_t x, int s) { return __shfl(x, s); }
  static __device__ __forceinline__ uint32_t up(uint32_t x, int d) { return __shfl_up(x, d); }
  static __device__ __forceinline__ uint32_t xr(uint32_t x, int m) { return __shfl_xor(x, m); }
};
template <> struct Shfl<float> {
  static __device__ __forceinline__ float at(float x, int s) { return __shfl(x, s); }
  static __device__ __forceinline__ float up(float x, int d) { return __shfl_up(x, d); }
  static __device__ __forceinline__ float xr(float x, int m) { return __shfl_xor(x, m); }
};
template <> struct Shfl<uint64_t> {
  static __device__ __forceinline__ uint64_t at(uint64_t x, int s) { return shfl_u64(x, s); }
  static __device__ __forceinline__ uint64_t up(uint64_t x, int d) { return shfl_up_u64(x, d); }
  static __device__ __forceinline__ uint64_t xr(uint64_t x, int m) {
    return shfl_u64(x, lane_id() ^ m);
  }
};
template <> struct Shfl<double> {
  static __device__ __forceinline__ double at(double x, int s) { return shfl_f64(x, s); }
  static __device__ __forceinline__ double up(double x, int d) {
    return __longlong_as_double((long long)shfl_up_u64((uint64_t)__double_as_longlong(x), d));
  }
  static __device__ __forceinline__ double xr(double x, int m) {
    return shfl_f64(x, lane_id() ^ m);
  }
};

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
  for (int m = 32; m >= 1; m >>= 1) x += shfl_u64(x, lane_id() ^ m);
  return x;
}
__device__ __forceinline__ double wave_sum_f64(double x) {
  for (int m = 32; m >= 1; m >>= 1) x += shfl_f64(x, lane_id() ^ m);
  return x;
}
// Exclusive prefix sum of a u32 across the wave; *total = sum of all lanes.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t *total) {
  uint32_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(inc, d);
    if (lane_id() >= d) inc += y;
  }
  *total = __shfl(inc, 63);
  return inc - x;
}
// Largest lane k with excl[k] <= j (excl non-decreasing, excl[0] = 0).
__device__ __forceinline__ int wave_owner(uint32_t excl, uint32_t j) {
  int k = 0;
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) {
    uint32_t ec = __shfl(excl, k + s);
    if (ec <= j) k += s;
  }
  return k;
}

// Adjacency entries examined, spread over 64 cache lines (optional).
__device__ __forceinline__ void count_examined(unsigned long long *exam, uint64_t n) {
  if (exam && n) atomicAdd(&exam[(blockIdx.x & 63) * 16], (unsigned long long)n);
}

// Sets v in the bitmap; returns whether this call flipped the bit.
__device__ __forceinline__ bool activate(uint64_t *bits, uint32_t v) {
  uint64_t bit = 1ULL << (v & 63);
  uint64_t old = atomicOr((unsigned long long *)&bits[v >> 6], (unsigned long long)bit);
  return !(old & bit);
}

// ------------------------------------------------------------ launch sizes
inline unsigned grid1(uint64_t n, unsigned cap = 1u << 16) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}
inline unsigned grid_waves(uint64_t waves) {
  uint64_t g = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  return (unsigned)(g < 1 ? 1 : g);
}
// ----------------------------------------------------------- event timing
struct DeviceTimer {
  hipEvent_t a = nullptr, b = nullptr;
  DeviceTimer() {
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, stream());
  }
  double stop() {
    hipEventRecord(b, stream());
    hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    return ms;
  }
  ~DeviceTimer() {
    hipEventDestroy(a);
    hipEventDestroy(b);
  }
};

inline int check_graph(const gg_graph *g) {
  if (!g) return fail(GG_EINVAL, "null graph");
  GG_TRY(ensure_init());
  return GG_OK;
}

inline void put_stats(gg_run_stats *dst, const gg_run_stats &src) {
  if (dst) *dst = src;
}

// ------------------------------------------------------------- compaction
// Bitmap -> three queues binned by bin_deg (<= small_max, <= big_max, above),
// plus popcount and degree sum.  Optionally ORs the words into `visited`
// (process_vertices(visited->set_bit) over active_out, bfs.cpp:79-85).
__device__ __forceinline__ void compact_words(const uint64_t *bits, size_t words,
                                                   const uint32_t *bin_deg, uint32_t small_max,
                                                   uint32_t big_max, uint32_t *q0, uint32_t *q1,
                                                   uint32_t *q2, Counters *cnt, uint64_t *visited,
                                                   uint64_t *clear) {
  __shared__ uint32_t s_tot[3][kWavesPerBlock];
  __shared__ unsigned long long s_edges[kWavesPerBlock];
  __shared__ uint32_t s_base[3];
  const int lane = lane_id(), wv = wave_in_block();
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t base = (size_t)blockIdx.x * kBlock; base < words; base += stride) {
    size_t wi = base + threadIdx.x;
    uint64_t word = wi < words ? bits[wi] : 0;
    if (visited && word) visited[wi] |= word;
    if (clear && wi < words) clear[wi] = 0;
    uint32_t c0 = 0, c1 = 0, c2 = 0;
    uint64_t es = 0;
    for (uint64_t x = word; x; x &= x - 1) {
      uint32_t v = (uint32_t)(wi * 64 + (size_t)__builtin_ctzll(x));
      uint32_t d = bin_deg[v];
      es += d;
      if (d <= small_max)
        c0++;
      else if (d <= big_max)
        c1++;
      else
        c2++;
    }
    uint32_t t0, t1, t2;
    uint32_t e0 = wave_excl_scan(c0, &t0);
    uint32_t e1 = wave_excl_scan(c1, &t1);
    uint32_t e2 = wave_excl_scan(c2, &t2);
    uint64_t ws = wave_sum_u64(es);
    if (lane == 0) {
      s_tot[0][wv] = t0;
      s_tot[1][wv] = t1;
      s_tot[2][wv] = t2;
      s_edges[wv] = ws;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t b0 = 0, b1 = 0, b2 = 0;
      unsigned long long be = 0;
      for (int i = 0; i < kWavesPerBlock; i++) {
        b0 += s_tot[0][i];
        b1 += s_tot[1][i];
        b2 += s_tot[2][i];
        be += s_edges[i];
      }
      s_base[0] = b0 ? atomicAdd(&cnt->n[0], b0) : 0;
      s_base[1] = b1 ? atomicAdd(&cnt->n[1], b1) : 0;
      s_base[2] = b2 ? atomicAdd(&cnt->n[2], b2) : 0;
      if (b0 + b1 + b2) {
        atomicAdd(&cnt->total, b0 + b1 + b2);
        atomicAdd((unsigned long long *)&cnt->edges, be);
      }
    }
    __syncthreads();
    uint32_t o0 = s_base[0] + e0, o1 = s_base[1] + e1, o2 = s_base[2] + e2;
    for (int i = 0; i < wv; i++) {
      o0 += s_tot[0][i];
      o1 += s_tot[1][i];
      o2 += s_tot[2][i];
    }
    for (uint64_t x = word; x; x &= x - 1) {
      uint32_t v = (uint32_t)(wi * 64 + (size_t)__builtin_ctzll(x));
      uint32_t d = bin_deg[v];
      if (d <= small_max)
        q0[o0++] = v;
      else if (d <= big_max)
        q1[o1++] = v;
      else
        q2[o2++] = v;
    }
    __syncthreads();
  }
}

// Position of the r-th set bit of w (r < popcount(w)).
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t w, uint32_t r) {
  uint32_t pos = 0;
#pragma unroll
  for (int sh = 32; sh >= 1; sh >>= 1) {
    const uint32_t c = (uint32_t)__popcll(w & ((1ULL << sh) - 1));
    if (r >= c) {
      r -= c;
      pos += sh;
      w >>= sh;
    }
  }
  return pos;
}

// The same compaction for sparse frontiers, launched with kCompactBlock
// threads: a block takes one contiguous range of words and appends all of it
// with one atomic per counter (the per-slice atomics above, from thousands of
// slices, serialise in L2 when a frontier touches most of them), and a wave
// spreads the set bits of its 64 words over its lanes (a prefix of
// popcounts, then the owner word and the bit's position per lane), so a
// dense word costs no lane more than its share.  Slower than the version
// above on dense frontiers, where every word is full.
constexpr int kCompactBlock = 1024;
constexpr int kCompactWaves = kCompactBlock / 64;
inline unsigned compact_grid(size_t words) {
  size_t g = (words + kCompactBlock - 1) / kCompactBlock;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(g, 1024));
}
__device__ __forceinline__ void compact_words_spread(const uint64_t *bits, size_t words,
                                                     const uint32_t *bin_deg, uint32_t small_max,
                                                     uint32_t big_max, uint32_t *q0, uint32_t *q1,
                                                     uint32_t *q2, Counters *cnt,
                                                     uint64_t *visited, uint64_t *clear,
                                                     uint32_t nb) {
  __shared__ uint32_t s_tot[3][kCompactWaves];
  __shared__ unsigned long long s_edges[kCompactWaves];
  __shared__ uint32_t s_base[3];
  const int lane = lane_id(), wv = (int)(threadIdx.x >> 6);
  const uint64_t below = lanemask_lt();
  // this block's words (one of nb blocks): [beg, end), whole multiples of
  // the block size
  const size_t per = (words + (size_t)nb * kCompactBlock - 1) /
                     ((size_t)nb * kCompactBlock) * kCompactBlock;
  const size_t beg = (size_t)blockIdx.x * per;
  const size_t end = beg + per < words ? beg + per : words;
  // the j-th set bit of a wave's 64 words (vertex and bin degree)
  auto bit_at = [&](uint64_t word, uint32_t excl, uint32_t total, size_t wbase, uint32_t j,
                    uint32_t &v, uint32_t &d) {
    const bool ok = j < total;
    const uint32_t jj = ok ? j : total - 1;
    const int k = wave_owner(excl, jj);
    const uint64_t wk = shfl_u64(word, k);
    v = (uint32_t)((wbase + (size_t)k) * 64 + nth_set_bit(wk, jj - (uint32_t)__shfl(excl, k)));
    d = ok ? bin_deg[v] : 0;
  };
  // pass 1: this wave's totals over the whole range
  uint32_t c0 = 0, c1 = 0, c2 = 0;
  uint64_t es = 0;
  for (size_t base = beg; base < end; base += kCompactBlock) {
    const size_t wi = base + threadIdx.x;
    const uint64_t word = wi < end ? bits[wi] : 0;
    if (visited && word) visited[wi] |= word;
    if (clear && wi < end) clear[wi] = 0;
    uint32_t total;
    const uint32_t excl = wave_excl_scan((uint32_t)__popcll(word), &total);
    for (uint32_t j0 = 0; j0 < total; j0 += 64) {
      uint32_t v, d;
      bit_at(word, excl, total, base + (size_t)wv * 64, j0 + lane, v, d);
      const bool ok = j0 + lane < total;
      es += d;
      c0 += (uint32_t)__popcll(__ballot(ok && d <= small_max));
      c1 += (uint32_t)__popcll(__ballot(ok && d > small_max && d <= big_max));
      c2 += (uint32_t)__popcll(__ballot(ok && d > big_max));
    }
  }
  const uint64_t ws = wave_sum_u64(es);
  if (lane == 0) {
    s_tot[0][wv] = c0;
    s_tot[1][wv] = c1;
    s_tot[2][wv] = c2;
    s_edges[wv] = ws;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t b0 = 0, b1 = 0, b2 = 0;
    unsigned long long be = 0;
    for (int i = 0; i < kCompactWaves; i++) {
      b0 += s_tot[0][i];
      b1 += s_tot[1][i];
      b2 += s_tot[2][i];
      be += s_edges[i];
    }
    s_base[0] = b0 ? atomicAdd(&cnt->n[0], b0) : 0;
    s_base[1] = b1 ? atomicAdd(&cnt->n[1], b1) : 0;
    s_base[2] = b2 ? atomicAdd(&cnt->n[2], b2) : 0;
    if (b0 + b1 + b2) {
      atomicAdd(&cnt->total, b0 + b1 + b2);
      atomicAdd((unsigned long long *)&cnt->edges, be);
    }
  }
  __syncthreads();
  // pass 2: each wave writes its vertices from its offsets on
  uint32_t o0 = s_base[0], o1 = s_base[1], o2 = s_base[2];
  for (int i = 0; i < wv; i++) {
    o0 += s_tot[0][i];
    o1 += s_tot[1][i];
    o2 += s_tot[2][i];
  }
  if (!(c0 | c1 | c2)) return;
  for (size_t base = beg; base < end; base += kCompactBlock) {
    const size_t wi = base + threadIdx.x;
    const uint64_t word = wi < end ? bits[wi] : 0;
    uint32_t total;
    const uint32_t excl = wave_excl_scan((uint32_t)__popcll(word), &total);
    for (uint32_t j0 = 0; j0 < total; j0 += 64) {
      uint32_t v, d;
      bit_at(word, excl, total, base + (size_t)wv * 64, j0 + lane, v, d);
      const bool ok = j0 + lane < total;
      const bool b0 = ok && d <= small_max, b1 = ok && d > small_max && d <= big_max,
                 b2 = ok && d > big_max;
      const uint64_t m0 = __ballot(b0), m1 = __ballot(b1), m2 = __ballot(b2);
      if (b0) q0[o0 + __popcll(m0 & below)] = v;
      if (b1) q1[o1 + __popcll(m1 & below)] = v;
      if (b2) q2[o2 + __popcll(m2 & below)] = v;
      o0 += (uint32_t)__popcll(m0);
      o1 += (uint32_t)__popcll(m1);
      o2 += (uint32_t)__popcll(m2);
    }
  }
}
__global__ __launch_bounds__(kBlock) void compact_kernel(const uint64_t *bits, size_t words,
                                                         const uint32_t *bin_deg,
                                                         uint32_t small_max, uint32_t big_max,
                                                         uint32_t *q0, uint32_t *q1, uint32_t *q2,
                                                         Counters *cnt, uint64_t *visited,
                                                         uint64_t *clear);

// ------------------------------------------------------------------- push
struct PushArgs {
  const uint64_t *index;
  const uint32_t *nbr;
  const float *w;
  const uint32_t *q0, *q1, *q2;
  const uint32_t *s0, *s1, *s2;  // message snapshots (raw bits) or null
  uint32_t n0, n1, n2;
  uint32_t waves0, waves1, waves2;
  uint64_t *out;
  unsigned long long *exam;
};

// Op: using Msg; Msg msg(v, snap_bits); bool edge(src, msg, dst, w) -> candidate;
//     void after(bool fresh, dst).
template <class Op>
__global__ __launch_bounds__(kBlock) void push_kernel(PushArgs a, Op op) {
  const uint32_t gw = global_wave();
  const int lane = lane_id();
  using Msg = typename Op::Msg;
  if (gw < a.waves0) {
    // small bin: in-wave load-balanced expansion of 64 frontier vertices
    uint32_t i = gw * 64 + lane;
    bool valid = i < a.n0;
    uint32_t v = valid ? a.q0[i] : 0;
    uint64_t b = valid ? a.index[v] : 0;
    uint32_t deg = valid ? (uint32_t)(a.index[v + 1] - b) : 0;
    Msg m = op.msg(v, (valid && a.s0) ? a.s0[i] : 0u);
    uint32_t total;
    uint32_t excl = wave_excl_scan(deg, &total);
    for (uint32_t j0 = 0; j0 < total; j0 += 64) {
      uint32_t j = j0 + lane;
      int k = wave_owner(excl, j);
      uint32_t src = __shfl(v, k);
      uint64_t bb = shfl_u64(b, k);
      uint32_t ex = __shfl(excl, k);
      Msg mm = Shfl<Msg>::at(m, k);
      if (j < total) {
        uint64_t e = bb + (j - ex);
        uint32_t dst = a.nbr[e];
        if (op.edge(src, mm, dst, a.w ? a.w[e] : 0.f)) op.after(activate(a.out, dst), dst);
      }
    }
    if (lane == 0) count_examined(a.exam, total);
  } else if (gw < a.waves0 + a.waves1) {
    // mid bin: one wave per vertex
    uint32_t i = gw - a.waves0;
    uint32_t v = a.q1[i];
    uint64_t b = a.index[v], e1 = a.index[v + 1];
    Msg m = op.msg(v, a.s1 ? a.s1[i] : 0u);
    for (uint64_t e = b + lane; e < e1; e += 64) {
      uint32_t dst = a.nbr[e];
      if (op.edge(v, m, dst, a.w ? a.w[e] : 0.f)) op.after(activate(a.out, dst), dst);
    }
    if (lane == 0) count_examined(a.exam, e1 - b);
  } else {
    // huge bin: every remaining wave takes chunks of every huge vertex (the
    // grid is rounded up to whole blocks: waves past waves2 have no share)
    uint32_t hw = gw - a.waves0 - a.waves1;
    if (hw >= a.waves2) return;
    uint64_t done = 0;
    for (uint32_t h = 0; h < a.n2; h++) {
      uint32_t v = a.q2[h];
      uint64_t b = a.index[v], e1 = a.index[v + 1];
      Msg m = op.msg(v, a.s2 ? a.s2[h] : 0u);
      uint64_t nch = (e1 - b + kChunkEdges - 1) / kChunkEdges;
      // chunk c goes to huge-bin wave (c + rot) % waves2, rotated per vertex
      // so vertices with few chunks spread over all the waves
      const uint32_t rot = (uint32_t)(((uint64_t)h * 2654435761u) % a.waves2);
      for (uint64_t c = (hw + a.waves2 - rot) % a.waves2; c < nch; c += a.waves2) {
        uint64_t cb = b + c * kChunkEdges;
        uint64_t ce = cb + kChunkEdges < e1 ? cb + kChunkEdges : e1;
        for (uint64_t e = cb + lane; e < ce; e += 64) {
          uint32_t dst = a.nbr[e];
          if (op.edge(v, m, dst, a.w ? a.w[e] : 0.f)) op.after(activate(a.out, dst), dst);
        }
        done += ce - cb;
      }
    }
    if (lane == 0) count_examined(a.exam, done);
  }
}

struct NoAfter {
  __device__ void after(bool, uint32_t) const {}
};

// ================================================================ PageRank
// toolkits/pagerank.cpp:24-129
struct PrPull : NoAfter {
  using T = double;
  const double *curr;
  double *next;
  const uint32_t *outdeg;
  int last;
  __device__ double identity() const { return 0.0; }
  __device__ bool dst_active(uint32_t) const { return true; }
  __device__ double value(uint32_t src, float) const { return curr[src]; }
  __device__ double combine(double a, double b) const { return a + b; }
  __device__ bool finish(uint32_t v, double sum, uint32_t, double &dsum) const {
    // next = 1 - d + d * next, divided by the out-degree except in the last
    // iteration (pagerank.cpp:77-97)
    const double d = 0.85;
    double nx = 1 - d + d * sum;
    if (!last) {
      uint32_t od = outdeg[v];
      if (od > 0) {
        nx /= od;
        dsum += fabs(nx - curr[v]) * od;
      } else {
        dsum += fabs(nx - curr[v]);
      }
    }
    next[v] = nx;
    return false;
  }
};
// ------------------------------------------------------------------- pull
struct PullArgs {
  const uint64_t *index;
  const uint32_t *nbr;
  const float *w;
  uint32_t V;
  uint32_t light_max;
  uint32_t ngroups;
  uint64_t *out;
  unsigned long long *exam;
  // scalar accumulator of the op (PageRank delta), may be null: kScalarSlots
  // partial sums 16 doubles (one 128-byte line) apart, so that the waves'
  // atomics spread over many L2 channels instead of serialising on one line
  double *dacc;
  // per-vertex prefix length of the in-list to read (null: the whole list)
  const uint32_t *count;
};

// Reduction pull over one 64-vertex group per wave.
// Op: using T; T identity(); bool dst_active(v); T value(src, w); T combine(a, b);
//     bool finish(v, acc, deg, double &dsum) -> activation; void after(bool, v).
template <class Op>
__global__ __launch_bounds__(kBlock) void pull_group_kernel(PullArgs a, Op op) {
  using T = typename Op::T;
  __shared__ T lacc[kWavesPerBlock][64];
  const uint32_t g = global_wave();
  const int lane = lane_id();
  const int wv = wave_in_block();
  if (g >= a.ngroups) return;  // whole wave leaves together
  uint32_t v = g * 64 + lane;
  bool inr = v < a.V;
  uint64_t b = inr ? a.index[v] : 0;
  uint64_t e1 = inr ? (a.count ? b + a.count[v] : a.index[v + 1]) : 0;
  uint64_t deg64 = e1 - b;
  bool light = inr && deg64 <= a.light_max && op.dst_active(v);
  uint32_t deg = light ? (uint32_t)deg64 : 0u;
  lacc[wv][lane] = op.identity();
  wave_sync();
  uint32_t total;
  uint32_t excl = wave_excl_scan(deg, &total);
  // Four 64-entry steps per round: their adjacency loads, then their
  // gathers, are issued back to back; the reductions then run in step order
  // (the same combination order as one step at a time).
  constexpr int kSteps = 4;
  for (uint32_t j0 = 0; j0 < total; j0 += 64 * kSteps) {
    int ks[kSteps];
    uint32_t us[kSteps];
    float ws[kSteps];
#pragma unroll
    for (int t = 0; t < kSteps; t++) {
      const uint32_t j = j0 + t * 64 + lane;
      const bool act = j < total;
      const int k = wave_owner(excl, j);
      const uint64_t bb = shfl_u64(b, k);
      const uint32_t ex = __shfl(excl, k);
      const uint64_t e = bb + (j - ex);
      ks[t] = act ? k : 64;
      us[t] = act ? stream_load(&a.nbr[e]) : 0u;
      ws[t] = act && a.w ? stream_load(&a.w[e]) : 0.f;
    }
    T xs[kSteps];
#pragma unroll
    for (int t = 0; t < kSteps; t++) xs[t] = ks[t] < 64 ? op.value(us[t], ws[t]) : op.identity();
#pragma unroll
    for (int t = 0; t < kSteps; t++) {
      if (j0 + t * 64 >= total) break;  // wave-uniform
      T x = xs[t];
      const int kk = ks[t];
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        T y = Shfl<T>::up(x, d);
        int ko = __shfl_up(kk, d);
        if (lane >= d && ko == kk) x = op.combine(x, y);
      }
      const int nk = __shfl_down(kk, 1);
      const bool last = (lane == 63) || nk != kk;
      if (kk < 64 && last) lacc[wv][kk] = op.combine(lacc[wv][kk], x);
      wave_sync();
    }
  }
  T acc = lacc[wv][lane];
  double dsum = 0;
  bool cand = light ? op.finish(v, acc, (uint32_t)deg64, dsum) : false;
  // this wave owns word g; heavy chunks may set other bits of it
  uint64_t m = __ballot(cand);
  if (lane == 0 && m) atomicOr((unsigned long long *)&a.out[g], (unsigned long long)m);
  op.after(cand, v);
  if (a.dacc) {
    double s = wave_sum_f64(dsum);
    if (lane == 0 && s != 0.0) unsafeAtomicAdd(&a.dacc[(g % kScalarSlots) * 16], s);
  }
  if (lane == 0) count_examined(a.exam, total);
}

struct HeavyArgs {
  const uint32_t *nbr;
  const float *w;
  const uint32_t *chunk_v, *chunk_h;
  const uint64_t *chunk_beg, *chunk_end;
  uint32_t n_chunks;
  uint64_t *out;
  unsigned long long *exam;
  // per-chunk partials (8 bytes each): when set, chunks store their partial
  // and heavy_reduce_kernel commits once per vertex
  void *part;
};

// One wave per heavy chunk; the chunk's partial goes through op.commit.
// Op: T identity(); bool dst_active(v); T value(src, w); T combine(a,b);
//     bool commit(v, h, partial) -> activation; void after(fresh, v).
template <class Op>
__global__ __launch_bounds__(kBlock) void pull_heavy_kernel(HeavyArgs a, Op op) {
  using T = typename Op::T;
  const int lane = lane_id();
  const uint32_t W = gridDim.x * kWavesPerBlock;
  // persistent grid: each wave walks chunks c, c + W, ...
  for (uint32_t c = global_wave(); c < a.n_chunks; c += W) {
    const uint32_t v = a.chunk_v[c];
    if (!op.dst_active(v)) {
      if (a.part && lane == 0)
        reinterpret_cast<T *>(static_cast<uint64_t *>(a.part) + c)[0] = op.identity();
      continue;
    }
    const uint64_t cb = a.chunk_beg[c], ce = a.chunk_end[c];
    T x = op.identity();
    // chunks hold at most kChunkEdges = 16 x 64 entries: load all neighbour
    // ids first, then gather, so each lane keeps 16 independent loads in flight
    constexpr int kPer = kChunkEdges / 64;
    uint32_t u[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint64_t e = cb + lane + 64 * k;
      u[k] = e < ce ? stream_load(&a.nbr[e]) : 0xFFFFFFFFu;
    }
    T y[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint64_t e = cb + lane + 64 * k;
      y[k] = u[k] != 0xFFFFFFFFu ? op.value(u[k], a.w ? stream_load(&a.w[e]) : 0.f)
                                 : op.identity();
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) x = op.combine(x, y[k]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x = op.combine(x, Shfl<T>::xr(x, m));
    if (lane == 0) {
      if (a.part)
        reinterpret_cast<T *>(static_cast<uint64_t *>(a.part) + c)[0] = x;
      else if (op.commit(v, a.chunk_h[c], x))
        op.after(activate(a.out, v), v);
      count_examined(a.exam, ce - cb);
    }
  }
}

// One wave per heavy vertex: combines its chunks' partials (lane-strided,
// then a fixed xor tree: the same order every run) and commits once, so
// chunks of one hub never contend on one address.
template <class Op>
__global__ __launch_bounds__(kBlock) void heavy_reduce_kernel(const uint32_t *heavy_v,
                                                             const uint32_t *heavy_first,
                                                             uint32_t n_heavy, const void *part,
                                                             uint64_t *out, Op op) {
  using T = typename Op::T;
  const uint32_t h = global_wave();
  const int lane = lane_id();
  if (h >= n_heavy) return;
  const uint32_t v = heavy_v[h];
  if (!op.dst_active(v)) return;
  const T *p = reinterpret_cast<const T *>(static_cast<const uint64_t *>(part));
  T x = op.identity();
  for (uint32_t c = heavy_first[h] + lane; c < heavy_first[h + 1]; c += 64)
    x = op.combine(x, *reinterpret_cast<const T *>(reinterpret_cast<const uint64_t *>(p) + c));
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x = op.combine(x, Shfl<T>::xr(x, m));
  if (lane == 0 && op.commit(v, h, x)) op.after(activate(out, v), v);
}

}  // namespace gg

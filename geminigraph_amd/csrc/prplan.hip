// PageRank over destination bins (toolkits/pagerank.cpp:24-129 computes the
// same sums).
//
// A PageRank iteration sums curr[u] over the in-arcs u -> v of every vertex.
// Gathered per arc in destination order (the plain pull), a random FP64
// gather costs an L2 request even when it hits (about 200 G/s on the chip),
// and on RMAT-24 tens of millions of them miss to the Infinity Cache (about
// 55 G/s).  On a relabelled graph (descending out-degree) the plan splits the
// destinations into bins of 2^14 vertices (by default), whose sums live in LDS, and
// the arcs by source:
//   hot (source < H <= 2^18, 2 MiB of values): each bin's hot arcs, sorted by
//     source and packed as (source << 14 | destination - bin start) in one
//     u32.  Consecutive lanes gather neighbouring sources, so one L2 request
//     serves several lanes, and every add goes to LDS;
//   cold (source >= H): propagation blocking.  The cold arcs are stored by
//     source range (2^14 sources by default), then by bin, then by source.  Phase 1
//     (emit) loads one range's values into LDS and writes every arc's value
//     in storage order -- a pure stream.  Phase 2 reads, for each bin, the
//     run that every range holds for it: (value, local destination) pairs
//     added into LDS.
// One bin kernel does the hot arcs, the cold runs and the PageRank update of
// its vertices.  A bin with more arcs than kSegArcs is split into segments
// that add their partial sums into a global array; a small kernel finishes
// those bins.
// Streamed bytes per arc: hot 4 (+ an L2 hit shared by several lanes), cold
// 2 + 8 (emit) + 8 + 2 (bin) (+ 8 per cold source).
#include <atomic>
#include <mutex>
#include <string.h>

#include <algorithm>
#include <rocprim/rocprim.hpp>

#include "kernels.hpp"

namespace gg {

// Bins hold 2^bb destinations and ranges 2^rb sources (bb, rb <= 14: 128
// KiB of LDS doubles); a packed hot arc keeps 32 - bb bits of source.
constexpr int kPrBlock = 1024;             // bin and emit workgroups
constexpr uint32_t kSegArcs = 1u << 18;    // arcs per bin segment
constexpr uint32_t kEmitArcs = 1u << 18;   // arcs per emit part

constexpr int kPrWaves = kPrBlock / 64;

// One bin segment: hot arcs [hot_lo, hot_hi) of the bin and a share of its
// cold runs, cut into equal item counts for the kPrWaves waves: wave w takes
// the items from position cut[w] to cut[w + 1], a position being (run << 32 |
// offset in the run).  slot 0: the segment is the whole bin and applies the
// update; slot s > 0: a piece of a split bin, adding into gsum row s - 1.
struct PrSeg {
  uint32_t bin, slot, hot_lo, hot_hi;
  uint64_t cut[kPrWaves + 1];
};

// ------------------------------------------------------------------ build
// hot_cnt[v] = in-neighbours of v below H (lower bound in the sorted list).
__global__ void pr_hot_count_kernel(const uint64_t *index, const uint32_t *nbr, uint32_t V,
                                    uint32_t H, uint32_t *hot_cnt) {
  const uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s) {
    uint64_t lo = index[v], hi = index[v + 1];
    const uint64_t b = lo;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (nbr[mid] < H)
        lo = mid + 1;
      else
        hi = mid;
    }
    hot_cnt[v] = (uint32_t)(lo - b);
  }
}

// One wave per destination v: key (bin, source, local destination) of each
// hot in-arc, at hot_off[v] onwards.
// (With weights, wbits[] receives each arc's weight bits beside its key.)
__global__ void pr_hot_keys_kernel(const uint64_t *index, const uint32_t *nbr, const float *w,
                                   uint32_t V, const uint32_t *hot_cnt, const uint64_t *hot_off,
                                   int bb, uint64_t *keys, uint32_t *wbits) {
  const uint32_t W = gridDim.x * (blockDim.x / 64);
  const int lane = lane_id();
  for (uint32_t v = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); v < V; v += W) {
    const uint64_t b = index[v], o = hot_off[v];
    const uint32_t n = hot_cnt[v];
    const uint64_t hi = (uint64_t)(v >> bb) << 32;
    const uint32_t dl = v & ((1u << bb) - 1);
    for (uint32_t i = lane; i < n; i += 64) {
      keys[o + i] = hi | (uint64_t)nbr[b + i] << bb | dl;
      if (w) wbits[o + i] = __float_as_uint(w[b + i]);
    }
  }
}

__global__ void pr_hot_pack_kernel(const uint64_t *keys, uint64_t n, uint32_t *arcs) {
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s)
    arcs[i] = (uint32_t)keys[i];
}

// One wave per cold source u: key (range, bin, local source, local
// destination) of each out-arc, at its CSR position - cold_base.
__global__ void pr_cold_keys_kernel(const uint64_t *index, const uint32_t *nbr, const float *w,
                                    uint32_t H, uint32_t V, uint64_t cold_base, int bb, int rb,
                                    int bbits, uint64_t *keys, uint32_t *wbits) {
  const uint32_t W = gridDim.x * (blockDim.x / 64);
  const int lane = lane_id();
  for (uint32_t u = H + blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); u < V; u += W) {
    const uint64_t b = index[u], e = index[u + 1];
    const uint64_t r = (u - H) >> rb, su = (u - H) & ((1u << rb) - 1);
    const uint64_t hi = (r << (bbits + rb + bb)) | (su << bb);
    const uint32_t dmask = (1u << bb) - 1;
    for (uint64_t a = b + lane; a < e; a += 64) {
      const uint32_t d = nbr[a];
      keys[a - cold_base] = hi | ((uint64_t)(d >> bb) << (rb + bb)) | (d & dmask);
      if (w) wbits[a - cold_base] = __float_as_uint(w[a]);
    }
  }
}

// Local source and destination of each sorted cold arc; run (range r, bin b)
// = [run_beg, run_end) at index b * nranges + r.
__global__ void pr_cold_split_kernel(const uint64_t *keys, uint64_t n, int bb, int rb, int bbits,
                                     uint32_t nranges, uint16_t *src16, uint16_t *dst16,
                                     uint32_t *run_beg, uint32_t *run_end) {
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t bmask = (1ULL << bbits) - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += s) {
    const uint64_t k = keys[i];
    src16[i] = (uint16_t)((k >> bb) & ((1u << rb) - 1));
    dst16[i] = (uint16_t)(k & ((1u << bb) - 1));
    const int sh = rb + bb;
    const uint64_t pair = k >> sh;
    const uint32_t r = (uint32_t)(pair >> bbits), b = (uint32_t)(pair & bmask);
    const uint64_t slot = (uint64_t)b * nranges + r;
    if (i == 0 || (keys[i - 1] >> sh) != pair) run_beg[slot] = (uint32_t)i;
    if (i + 1 == n || (keys[i + 1] >> sh) != pair) run_end[slot] = (uint32_t)(i + 1);
  }
}

// -------------------------------------------------------------- iteration
// The per-iteration kernels combine 8-byte values of the in-neighbours of
// every vertex under a BinOp: PageRank sums FP64 contributions; the batched
// BFS ORs 64-root frontier masks.  A BinOp gives the value type T, the
// source value, the LDS and global combining atomics, and the per-vertex
// finish (which reports whether the vertex joins the output frontier).  With
// VT is how a source value is stored (the per-vertex input and the emitted
// cold values), widen(VT) gives the combining type T.  With
// kSkipDone, done_mask(v0, valid) gives the wave's 64 destinations v0 ..
// v0+63 (v0 a multiple of 64; lanes with valid false are not read) whose
// result is already settled: the bin kernel then reads no source value for
// their arcs.
struct PrSumOp {
  using T = double;
  using VT = double;
  __device__ static T widen(VT v) { return v; }
  static constexpr bool kSkipDone = false;
  static constexpr bool kWeighted = false;  // true: relax(value, weight) per arc
  PrPull op;
  __device__ uint64_t done_mask(uint32_t, bool) const { return 0; }
  __device__ T value(uint32_t u) const { return op.curr[u]; }
  __device__ static void lds_add(T *acc, uint32_t i, T x) { atomicAdd(&acc[i], x); }
  __device__ static void global_add(T *p, T x) {
    if (x != 0.0) unsafeAtomicAdd(p, x);
  }
  __device__ bool finish(uint32_t v, T sum, double &dsum) const {
    op.finish(v, sum, 0, dsum);
    return false;
  }
};
struct MsOrOp {
  using T = uint64_t;
  using VT = uint64_t;
  __device__ static T widen(VT v) { return v; }
  static constexpr bool kSkipDone = true;  // every root has reached the vertex
  static constexpr bool kWeighted = false;
  const uint64_t *fcur, *vis;
  uint64_t *fnext;
  uint64_t full;
  const uint64_t *settled;  // bit v set once vis[v] covers full, or null (test vis)
  __device__ uint64_t done_mask(uint32_t v0, bool valid) const {
    if (settled) return settled[v0 >> 6];
    return __ballot(valid && (vis[v0 + lane_id()] & full) == full);
  }
  __device__ T value(uint32_t u) const { return fcur[u]; }
  __device__ static void lds_add(T *acc, uint32_t i, T x) {
    if (x) atomicOr((unsigned long long *)&acc[i], (unsigned long long)x);
  }
  __device__ static void global_add(T *p, T x) {
    if (x) atomicOr((unsigned long long *)p, (unsigned long long)x);
  }
  __device__ bool finish(uint32_t v, T acc, double &) const {
    const uint64_t nw = acc & full & ~vis[v];
    if (nw) fnext[v] = nw;
    return nw != 0;
  }
};

// CC (cc.cpp:57-79): the minimum label over each vertex's in-neighbours,
// taken from the labels at the start of the step (BSP: the finish writes the
// label array, the gathers read the copy).  Values are stored complemented,
// ~label (32 bits, never 0 as labels are below 2^32 - 1), and combined by
// max, so 0 (the plan's initial accumulator and split-bin row) is the
// identity; the finish lowers a label and reports the vertex for the output
// frontier.
struct CcMinOp {
  using T = uint64_t;
  using VT = uint32_t;
  __device__ static T widen(VT v) { return v; }
  static constexpr bool kSkipDone = false;
  static constexpr bool kWeighted = false;
  const uint32_t *nlab;  // ~label at the start of the step, per vertex
  uint32_t *label;
  __device__ uint64_t done_mask(uint32_t, bool) const { return 0; }
  __device__ T value(uint32_t u) const { return nlab[u]; }
  __device__ static void lds_add(T *acc, uint32_t i, T x) {
    atomicMax((unsigned long long *)&acc[i], (unsigned long long)x);
  }
  __device__ static void global_add(T *p, T x) {
    if (x) atomicMax((unsigned long long *)p, (unsigned long long)x);
  }
  __device__ bool finish(uint32_t v, T acc, double &) const {
    if (!acc) return false;  // no in-arcs
    const uint32_t m = ~(uint32_t)acc;
    if (m >= label[v]) return false;
    label[v] = m;
    return true;
  }
};

// SSSP (sssp.cpp:58-79): per vertex the minimum over its in-arcs of
// fl(dist[u] + w) with dist at the start of the step (the snapshot holds the
// f32 bits); like CC, the arc values are complemented 64-bit keys combined
// by max (distances are >= 0, so the bit order is the numeric order).  The
// finish applies the slot: lowered below 1e9 and below dist[v].
struct SsspMinOp {
  using T = uint64_t;
  using VT = uint32_t;  // the f32 bits
  __device__ static T widen(VT v) { return v; }
  static constexpr bool kSkipDone = false;
  static constexpr bool kWeighted = true;
  const uint32_t *snap;
  float *dist;
  __device__ uint64_t done_mask(uint32_t, bool) const { return 0; }
  __device__ T value(uint32_t u) const { return snap[u]; }
  __device__ T relax(T x, float w) const {
    const float r = __uint_as_float((uint32_t)x) + w;  // one f32 add, round to nearest
    return ~(uint64_t)__float_as_uint(r);
  }
  __device__ static void lds_add(T *acc, uint32_t i, T x) {
    atomicMax((unsigned long long *)&acc[i], (unsigned long long)x);
  }
  __device__ static void global_add(T *p, T x) {
    if (x) atomicMax((unsigned long long *)p, (unsigned long long)x);
  }
  __device__ bool finish(uint32_t v, T acc, double &) const {
    if (!acc) return false;  // no in-arcs
    const float m = __uint_as_float((uint32_t)~acc);
    if (!(m < 1e9f) || !(m < dist[v])) return false;
    dist[v] = m;
    return true;
  }
};

// Emit: part p = arcs [lo, hi) of range r.  The range's values go to LDS
// (one coalesced read), then every arc's value is written in storage order.
// Values are copied as VT words (8 or 4 bytes).
template <class VT>
__global__ __launch_bounds__(kPrBlock) void pr_emit_kernel(const uint32_t *gate,
                                                           const uint32_t *parts, const VT *curr,
                                                           const uint16_t *src16, uint32_t H,
                                                           uint32_t V, int rb, VT *vals) {
  if (gate && *gate != kGateDense) return;
  extern __shared__ __align__(8) unsigned char val_raw[];
  VT *val = reinterpret_cast<VT *>(val_raw);  // 2^rb values
  const uint32_t r = parts[3 * blockIdx.x], lo = parts[3 * blockIdx.x + 1],
                 hi = parts[3 * blockIdx.x + 2];
  const uint32_t u0 = H + (r << rb);
  const uint32_t n = V - u0 < (1u << rb) ? V - u0 : (1u << rb);
  for (uint32_t i = threadIdx.x; i < n; i += kPrBlock) val[i] = curr[u0 + i];
  __syncthreads();
  constexpr int kU = 8;
  for (uint32_t a = lo + threadIdx.x; a < hi; a += kPrBlock * kU) {
    uint16_t s[kU];
#pragma unroll
    for (int t = 0; t < kU; t++) {
      const uint32_t q = a + t * kPrBlock;
      s[t] = q < hi ? stream_load(&src16[q]) : (uint16_t)0;
    }
#pragma unroll
    for (int t = 0; t < kU; t++) {
      const uint32_t q = a + t * kPrBlock;
      if (q < hi) __builtin_nontemporal_store(val[s[t]], &vals[q]);
    }
  }
}

// Bin segment: hot arcs, then cold runs, combined in LDS; then the finish of
// the bin's vertices (whole bin) or a combine into the split bin's row.
// `out`, when given, receives the bitmap words of the vertices whose finish
// reports true (a whole bin owns its words).
template <class Op>
__global__ __launch_bounds__(kPrBlock) void pr_bin_kernel(
    const uint32_t *gate, const float *hot_w, const float *cold_w, const PrSeg *segs,
    const uint32_t *hot, const uint32_t *run_beg, const uint32_t *run_end,
    uint32_t nranges, const uint16_t *dst16, const typename Op::VT *vals, uint32_t V, int bb,
    Op op, typename Op::T *gsum, uint64_t *out, double *dacc) {
  using T = typename Op::T;
  if (gate && *gate != kGateDense) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *acc = reinterpret_cast<T *>(lds_raw);  // 2^bb values
  const uint32_t nb = 1u << bb, dmask = nb - 1;
  const PrSeg &sg = segs[blockIdx.x];  // read in place: cut[] is indexed by wave
  const uint32_t base = sg.bin << bb;
  const uint32_t n = V - base < nb ? V - base : nb;
  for (uint32_t i = threadIdx.x; i < nb; i += kPrBlock) acc[i] = T(0);
  // settled destinations of this bin, one bit each (bins hold at most 2^14)
  __shared__ uint32_t sdone[Op::kSkipDone ? (1u << 14) / 32 : 1];
  __shared__ uint32_t s_ndone;
  bool skip = false;  // this segment tests sdone before reading a source value
  if constexpr (Op::kSkipDone) {
    if (threadIdx.x == 0) s_ndone = 0;
    __syncthreads();
    for (uint32_t i0 = threadIdx.x - lane_id(); i0 < nb; i0 += kPrBlock) {
      // lanes past the last vertex count as settled
      const uint64_t past = i0 >= n ? ~0ULL : n - i0 >= 64 ? 0ULL : ~0ULL << (n - i0);
      const uint64_t m = past | (i0 < n ? op.done_mask(base + i0, i0 + lane_id() < n) : 0ULL);
      if (lane_id() == 0) {
        sdone[i0 >> 5] = (uint32_t)m;
        sdone[(i0 >> 5) + 1] = (uint32_t)(m >> 32);
        atomicAdd(&s_ndone, (uint32_t)__popcll(m));
      }
    }
    __syncthreads();
    const uint32_t nd = s_ndone;
    if (nd == nb) {  // nothing left to find in this bin
      if (sg.slot) return;
      for (uint32_t i0 = threadIdx.x - lane_id(); i0 < n; i0 += kPrBlock)
        if (out && lane_id() == 0) out[(base + i0) >> 6] = 0;
      return;
    }
    // testing costs an LDS read per arc: worth it once a quarter is settled
    skip = 4 * nd >= nb;
  } else {
    __syncthreads();
  }
  auto settled = [&](uint32_t d) -> bool {
    if constexpr (Op::kSkipDone) return skip && ((sdone[d >> 5] >> (d & 31)) & 1u);
    return false;
  };
  {  // hot arcs: sorted by source, so lanes share gathered lines
    constexpr int kU = 8;
    for (uint32_t a = sg.hot_lo + threadIdx.x; a < sg.hot_hi; a += kPrBlock * kU) {
      uint32_t k[kU];
#pragma unroll
      for (int t = 0; t < kU; t++) {
        const uint32_t q = a + t * kPrBlock;
        k[t] = q < sg.hot_hi ? stream_load(&hot[q]) : 0u;
      }
      T x[kU];
#pragma unroll
      for (int t = 0; t < kU; t++) x[t] = settled(k[t] & dmask) ? T(0) : op.value(k[t] >> bb);
      if constexpr (Op::kWeighted) {
#pragma unroll
        for (int t = 0; t < kU; t++) {
          const uint32_t q = a + t * kPrBlock;
          x[t] = op.relax(x[t], q < sg.hot_hi ? stream_load(&hot_w[q]) : 0.f);
        }
      }
#pragma unroll
      for (int t = 0; t < kU; t++)
        if (a + t * kPrBlock < sg.hot_hi) Op::lds_add(acc, k[t] & dmask, x[t]);
    }
  }
  {  // cold runs: this wave's share, 64 runs per step, flattened by an in-wave scan
    const int lane = lane_id();
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t *rb = run_beg + (uint64_t)sg.bin * nranges;
    const uint32_t *re = run_end + (uint64_t)sg.bin * nranges;
    const uint32_t end_r = (uint32_t)(sg.cut[wv + 1] >> 32), end_off = (uint32_t)sg.cut[wv + 1];
    uint32_t ri = (uint32_t)(sg.cut[wv] >> 32), off = (uint32_t)sg.cut[wv];
    for (;;) {
      const uint32_t r = ri + lane;
      uint32_t beg = 0, len = 0;
      if (r < end_r || (r == end_r && end_off)) {
        const uint32_t lo = r == ri ? off : 0;
        const uint32_t b0 = rb[r];
        const uint32_t hi = r < end_r ? re[r] - b0 : end_off;
        beg = b0 + lo;
        len = hi > lo ? hi - lo : 0;
      }
      uint32_t total;
      const uint32_t excl = wave_excl_scan(len, &total);
      constexpr int kU = 4;
      for (uint32_t j0 = 0; j0 < total; j0 += 64 * kU) {
        uint32_t pos[kU];
#pragma unroll
        for (int t = 0; t < kU; t++) {
          const uint32_t j = j0 + t * 64 + lane;
          const uint32_t jj = j < total ? j : total - 1;
          const int o = wave_owner(excl, jj);
          pos[t] = __shfl(beg, o) + (jj - __shfl(excl, o));
        }
        uint16_t d[kU];
        T x[kU];
        if (skip) {  // destinations first, then only the values that count
#pragma unroll
          for (int t = 0; t < kU; t++) d[t] = stream_load(&dst16[pos[t]]);
#pragma unroll
          for (int t = 0; t < kU; t++)
            x[t] = settled(d[t]) ? T(0) : Op::widen(stream_load(&vals[pos[t]]));
        } else {
#pragma unroll
          for (int t = 0; t < kU; t++) {
            d[t] = stream_load(&dst16[pos[t]]);
            x[t] = Op::widen(stream_load(&vals[pos[t]]));
          }
        }
        if constexpr (Op::kWeighted) {
#pragma unroll
          for (int t = 0; t < kU; t++) x[t] = op.relax(x[t], stream_load(&cold_w[pos[t]]));
        }
#pragma unroll
        for (int t = 0; t < kU; t++)
          if (j0 + t * 64 + lane < total) Op::lds_add(acc, d[t], x[t]);
      }
      if (end_r < ri + 64 || (end_r == ri + 64 && !end_off)) break;
      ri += 64;
      off = 0;
    }
  }
  __syncthreads();
  if (sg.slot) {
    T *gs = gsum + ((uint64_t)(sg.slot - 1) << bb);
    for (uint32_t i = threadIdx.x; i < n; i += kPrBlock) Op::global_add(&gs[i], acc[i]);
    return;
  }
  double dsum = 0;
  // the block's waves take 64 consecutive vertices per step (one bitmap word)
  for (uint32_t i0 = threadIdx.x - lane_id(); i0 < n; i0 += kPrBlock) {
    const uint32_t i = i0 + lane_id();
    const bool on = i < n && op.finish(base + i, acc[i], dsum);
    const uint64_t m = __ballot(on);
    if (out && lane_id() == 0) out[(base + i0) >> 6] = m;
  }
  if (dacc) {
    const double w = wave_sum_f64(dsum);
    if (lane_id() == 0 && w != 0.0)
      unsafeAtomicAdd(
          &dacc[((blockIdx.x * (kPrBlock / 64) + (threadIdx.x >> 6)) % kScalarSlots) * 16], w);
  }
}

// The finish of the split bins' vertices, clearing their rows for the next
// iteration.
template <class Op>
__global__ __launch_bounds__(kBlock) void pr_split_finish_kernel(const uint32_t *gate,
                                                                 const uint32_t *split_bin,
                                                                 uint32_t nsplit, uint32_t V,
                                                                 int bb, Op op,
                                                                 typename Op::T *gsum,
                                                                 uint64_t *out, double *dacc) {
  using T = typename Op::T;
  if (gate && *gate != kGateDense) return;
  const uint64_t total = (uint64_t)nsplit << bb;
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  double dsum = 0;
  // a wave covers 64 consecutive vertices of one bin (bins are whole words)
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x - lane_id(); i0 < total;
       i0 += s) {
    const uint64_t i = i0 + lane_id();
    const uint32_t v = (split_bin[i0 >> bb] << bb) + (uint32_t)(i & ((1u << bb) - 1));
    bool on = false;
    if (i < total && v < V) {
      const T t = gsum[i];
      gsum[i] = T(0);
      on = op.finish(v, t, dsum);
    }
    const uint64_t m = __ballot(on);
    if (out && lane_id() == 0 && (v - lane_id()) < V) out[(v - lane_id()) >> 6] = m;
  }
  if (dacc) {
    const double w = wave_sum_f64(dsum);
    if (lane_id() == 0 && w != 0.0)
      unsafeAtomicAdd(&dacc[(global_wave() % kScalarSlots) * 16], w);
  }
}

// ------------------------------------------------------------------ plan
static int plan_alloc(PrPlan *p, void **ptr, size_t bytes) {
  GG_TRY(device_alloc(ptr, bytes));
  p->allocations.push_back(*ptr);
  return GG_OK;
}

void pr_plan_free(PrPlan *p) {
  if (!p) return;
  for (void *x : p->allocations) device_free(x);
  delete p;
}

// Temporary device buffers, freed on every exit path.
struct Scratch {
  std::vector<void *> ptrs;
  int get(void **p, size_t bytes) {
    int r = device_alloc(p, bytes);
    if (r == GG_OK) ptrs.push_back(*p);
    return r;
  }
  ~Scratch() {
    for (void *p : ptrs) device_free(p);
  }
};

static int bits_for(uint64_t n) {  // bits that hold 0 .. n-1 (at least 1)
  int b = 1;
  while (b < 64 && (1ULL << b) < n) b++;
  return b;
}

// Sorts keys[0, n) (and, when vals is not null, the u32 values beside them);
// keys and vals then point at the sorted copies.
static int sort_keys(Scratch &t, uint64_t *&keys, uint64_t n, int end_bit,
                     uint32_t **vals = nullptr) {
  uint64_t *alt = nullptr;
  void *tmp = nullptr;
  GG_TRY(t.get((void **)&alt, n * 8));
  rocprim::double_buffer<uint64_t> kb(keys, alt);
  size_t tmp_bytes = 0;
  if (vals && *vals) {
    uint32_t *valt = nullptr;
    GG_TRY(t.get((void **)&valt, n * 4));
    rocprim::double_buffer<uint32_t> vb(*vals, valt);
    GG_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kb, vb, (size_t)n, 0, end_bit,
                                     stream()));
    GG_TRY(t.get(&tmp, tmp_bytes));
    GG_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, (size_t)n, 0, end_bit, stream()));
    *vals = vb.current();
  } else {
    GG_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, kb, (size_t)n, 0, end_bit, stream()));
    GG_TRY(t.get(&tmp, tmp_bytes));
    GG_HIP(rocprim::radix_sort_keys(tmp, tmp_bytes, kb, (size_t)n, 0, end_bit, stream()));
  }
  keys = kb.current();
  return GG_OK;
}

// Hot arcs per bin, sorted by source; the bins' first hot arcs on the host.
static int plan_hot(const gg_graph *g, PrPlan *p, std::vector<uint64_t> &hot_bin_off) {
  const Dir &in = in_dir(g);
  const uint32_t V = g->vertices;
  Scratch t;
  uint32_t *cnt = nullptr;
  uint64_t *off = nullptr, *keys = nullptr;
  void *tmp = nullptr;
  GG_TRY(t.get((void **)&cnt, 4 * (size_t)V));
  GG_TRY(t.get((void **)&off, 8 * (size_t)V));
  hipLaunchKernelGGL(pr_hot_count_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), in.index,
                     in.nbr, V, p->H, cnt);
  size_t tmp_bytes = 0;
  GG_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, cnt, off, (uint64_t)0, (size_t)V,
                                 rocprim::plus<uint64_t>(), stream()));
  GG_TRY(t.get(&tmp, tmp_bytes));
  GG_HIP(rocprim::exclusive_scan(tmp, tmp_bytes, cnt, off, (uint64_t)0, (size_t)V,
                                 rocprim::plus<uint64_t>(), stream()));
  // bin b starts at the first hot arc of vertex b << bin_bits (< V)
  hot_bin_off.assign((size_t)p->nbins + 1, 0);
  GG_HIP(hipMemcpy2DAsync(hot_bin_off.data(), 8, off, (size_t)8 << p->bin_bits, 8, p->nbins,
                          hipMemcpyDeviceToHost, stream()));
  uint64_t last_off = 0;
  uint32_t last_cnt = 0;
  GG_HIP(hipMemcpyAsync(&last_off, off + V - 1, 8, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipMemcpyAsync(&last_cnt, cnt + V - 1, 4, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  p->n_hot = last_off + last_cnt;
  hot_bin_off[p->nbins] = p->n_hot;
  GG_TRY(plan_alloc(p, (void **)&p->hot, std::max<uint64_t>(p->n_hot, 1) * 4));
  if (in.w) GG_TRY(plan_alloc(p, (void **)&p->hot_w, std::max<uint64_t>(p->n_hot, 1) * 4));
  if (!p->n_hot) return GG_OK;
  GG_TRY(t.get((void **)&keys, p->n_hot * 8));
  uint32_t *wbits = nullptr;
  if (in.w) GG_TRY(t.get((void **)&wbits, p->n_hot * 4));
  hipLaunchKernelGGL(pr_hot_keys_kernel, dim3(grid1((uint64_t)V * 64)), dim3(kBlock), 0, stream(),
                     in.index, in.nbr, in.w, V, cnt, off, p->bin_bits, keys, wbits);
  GG_TRY(sort_keys(t, keys, p->n_hot, 32 + bits_for(p->nbins), &wbits));
  hipLaunchKernelGGL(pr_hot_pack_kernel, dim3(grid1(p->n_hot)), dim3(kBlock), 0, stream(), keys,
                     p->n_hot, p->hot);
  if (in.w)
    GG_HIP(hipMemcpyAsync(p->hot_w, wbits, p->n_hot * 4, hipMemcpyDeviceToDevice, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

// Cold arcs by (range, bin, source, destination), the run table (also
// returned on the host) and the emit parts.
static int plan_cold(const gg_graph *g, PrPlan *p, std::vector<uint32_t> &hb,
                     std::vector<uint32_t> &he) {
  const Dir &out = out_dir(g);
  const uint32_t V = g->vertices;
  const uint64_t n = p->n_cold;
  const uint64_t range = 1ULL << p->range_bits;
  p->nranges = (uint32_t)(((uint64_t)V - p->H + range - 1) / range);
  const uint64_t nruns = (uint64_t)p->nranges * p->nbins;
  GG_TRY(plan_alloc(p, (void **)&p->src16, std::max<uint64_t>(n, 1) * 2));
  GG_TRY(plan_alloc(p, (void **)&p->dst16, std::max<uint64_t>(n, 1) * 2));
  if (out.w) GG_TRY(plan_alloc(p, (void **)&p->cold_w, std::max<uint64_t>(n, 1) * 4));
  GG_TRY(plan_alloc(p, (void **)&p->run_beg, std::max<uint64_t>(nruns, 1) * 4));
  GG_TRY(plan_alloc(p, (void **)&p->run_end, std::max<uint64_t>(nruns, 1) * 4));
  GG_HIP(hipMemsetAsync(p->run_beg, 0, std::max<uint64_t>(nruns, 1) * 4, stream()));
  GG_HIP(hipMemsetAsync(p->run_end, 0, std::max<uint64_t>(nruns, 1) * 4, stream()));
  hb.assign(nruns, 0);
  he.assign(nruns, 0);
  if (!n) return GG_OK;
  const int bbits = bits_for(p->nbins);
  const int end_bit = bits_for(p->nranges) + bbits + p->range_bits + p->bin_bits;
  if (end_bit > 64) return fail(GG_EINVAL, "PageRank plan: graph too large for 64-bit arc keys");
  Scratch t;
  uint64_t *keys = nullptr;
  GG_TRY(t.get((void **)&keys, n * 8));
  uint32_t *wbits = nullptr;
  if (out.w) GG_TRY(t.get((void **)&wbits, n * 4));
  hipLaunchKernelGGL(pr_cold_keys_kernel, dim3(grid1((uint64_t)(V - p->H) * 64)), dim3(kBlock), 0,
                     stream(), out.index, out.nbr, out.w, p->H, V, p->cold_base, p->bin_bits,
                     p->range_bits, bbits, keys, wbits);
  GG_TRY(sort_keys(t, keys, n, end_bit, &wbits));
  if (out.w)
    GG_HIP(hipMemcpyAsync(p->cold_w, wbits, n * 4, hipMemcpyDeviceToDevice, stream()));
  hipLaunchKernelGGL(pr_cold_split_kernel, dim3(grid1(n)), dim3(kBlock), 0, stream(), keys, n,
                     p->bin_bits, p->range_bits, bbits, p->nranges, p->src16, p->dst16,
                     p->run_beg, p->run_end);
  GG_HIP(hipMemcpyAsync(hb.data(), p->run_beg, nruns * 4, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipMemcpyAsync(he.data(), p->run_end, nruns * 4, hipMemcpyDeviceToHost, stream()));
  // emit parts: range r holds the CSR arcs of its sources, in one block
  std::vector<uint64_t> rb((size_t)p->nranges + 1);
  GG_HIP(hipMemcpy2DAsync(rb.data(), 8, out.index + p->H, (size_t)range * 8, 8, p->nranges,
                          hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  rb[p->nranges] = g->arcs;
  std::vector<uint32_t> parts;
  for (uint32_t r = 0; r < p->nranges; r++) {
    const uint64_t lo = rb[r] - p->cold_base, hi = rb[r + 1] - p->cold_base;
    for (uint64_t a = lo; a < hi; a += kEmitArcs) {
      parts.push_back(r);
      parts.push_back((uint32_t)a);
      parts.push_back((uint32_t)std::min<uint64_t>(a + kEmitArcs, hi));
    }
  }
  p->nparts = (uint32_t)(parts.size() / 3);
  GG_TRY(plan_alloc(p, (void **)&p->parts, std::max<size_t>(parts.size(), 3) * 4));
  if (!parts.empty())
    GG_HIP(hipMemcpyAsync(p->parts, parts.data(), parts.size() * 4, hipMemcpyHostToDevice,
                          stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

// Bin segments: a bin with at most kSegArcs arcs is one segment that applies
// the update itself; a larger bin is split into hot pieces and cold pieces
// (whole runs) that add into a row of gsum, finished afterwards.  Each
// segment's cold items are cut into kPrWaves equal shares.
static void cut_runs(PrSeg &sg, const uint32_t *len, uint32_t r0, uint32_t r1) {
  uint64_t total = 0;
  for (uint32_t r = r0; r < r1; r++) total += len[r];
  uint32_t r = r0;
  uint64_t before = 0;  // items in runs r0 .. r-1
  for (int w = 0; w <= kPrWaves; w++) {
    const uint64_t t = total * w / kPrWaves;
    while (r < r1 && before + len[r] <= t) before += len[r++];
    sg.cut[w] = (uint64_t)r << 32 | (uint32_t)(t - before);
  }
  sg.cut[kPrWaves] = (uint64_t)r1 << 32;
}

static int plan_segments(PrPlan *p, const std::vector<uint64_t> &hot_bin_off,
                         const std::vector<uint32_t> &hb, const std::vector<uint32_t> &he) {
  std::vector<PrSeg> segs;
  std::vector<uint32_t> split;
  const uint32_t R = p->nranges;
  std::vector<uint32_t> len(R);
  auto seg = [&](uint32_t b, uint32_t slot, uint64_t h0, uint64_t h1, uint32_t r0, uint32_t r1) {
    PrSeg sg{b, slot, (uint32_t)h0, (uint32_t)h1, {}};
    cut_runs(sg, len.data(), r0, r1);
    segs.push_back(sg);
  };
  for (uint32_t b = 0; b < p->nbins; b++) {
    const uint64_t h0 = hot_bin_off[b], h1 = hot_bin_off[b + 1];
    uint64_t cold = 0;
    for (uint32_t r = 0; r < R; r++) {
      len[r] = he[(uint64_t)b * R + r] - hb[(uint64_t)b * R + r];
      cold += len[r];
    }
    if ((h1 - h0) + cold <= kSegArcs) {
      seg(b, 0, h0, h1, 0, R);
      continue;
    }
    const uint32_t slot = (uint32_t)split.size() + 1;
    split.push_back(b);
    for (uint64_t a = h0; a < h1; a += kSegArcs)
      seg(b, slot, a, std::min<uint64_t>(a + kSegArcs, h1), 0, 0);
    uint32_t r0 = 0;
    uint64_t acc = 0;
    for (uint32_t r = 0; r < R; r++) {
      acc += len[r];
      if (acc >= kSegArcs || r + 1 == R) {
        if (acc) seg(b, slot, 0, 0, r0, r + 1);
        r0 = r + 1;
        acc = 0;
      }
    }
  }
  p->nseg = (uint32_t)segs.size();
  p->nsplit = (uint32_t)split.size();
  GG_TRY(plan_alloc(p, (void **)&p->segs, std::max<size_t>(segs.size(), 1) * sizeof(PrSeg)));
  GG_TRY(plan_alloc(p, (void **)&p->split_bin, std::max<size_t>(split.size(), 1) * 4));
  if (!segs.empty())
    GG_HIP(hipMemcpyAsync(p->segs, segs.data(), segs.size() * sizeof(PrSeg), hipMemcpyHostToDevice,
                          stream()));
  if (!split.empty())
    GG_HIP(hipMemcpyAsync(p->split_bin, split.data(), split.size() * 4, hipMemcpyHostToDevice,
                          stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

static int plan_build(const gg_graph *g, uint32_t H, int bin_bits, int range_bits, PrPlan *p) {
  const Dir &out = out_dir(g);
  const uint32_t V = g->vertices;
  p->transposed = g->transposed;
  p->H = H;
  p->bin_bits = bin_bits;
  p->range_bits = range_bits;
  p->nbins = (uint32_t)(((uint64_t)V + (1ULL << bin_bits) - 1) >> bin_bits);
  p->cold_base = g->arcs;
  if (H < V) {
    GG_HIP(hipMemcpyAsync(&p->cold_base, out.index + H, 8, hipMemcpyDeviceToHost, stream()));
    GG_HIP(hipStreamSynchronize(stream()));
  }
  p->n_cold = g->arcs - p->cold_base;
  std::vector<uint64_t> hot_bin_off;
  std::vector<uint32_t> hb, he;
  GG_TRY(plan_hot(g, p, hot_bin_off));
  if (H < V) GG_TRY(plan_cold(g, p, hb, he));
  return plan_segments(p, hot_bin_off, hb, he);
}

// Threads that traverse one graph at once (serving workers) share its plan:
// the first to need it builds it while the others wait.  A plan is replaced
// when the plan options change; each caller holds a reference to the plan it
// got for the length of its call, so the replaced plan is freed (its device
// memory with hipFree, which waits for the device) when the last traversal
// that uses it returns.
static std::mutex g_plan_mu;

int pr_plan_get(gg_graph *g, PlanRef *out) {
  out->reset();
  const int64_t hs = options().pr_hot_sources;
  if (!g->reordered || hs <= 0 || !g->vertices) return GG_OK;
  if (g->arcs >= ((uint64_t)1 << 32)) return GG_OK;  // positions are 32-bit
  const int bb = (int)options().pr_bin_bits, rb = (int)options().pr_range_bits;
  const uint32_t H =
      (uint32_t)std::min<int64_t>({hs, (int64_t)g->vertices, (int64_t)1 << (32 - bb)});
  std::lock_guard<std::mutex> lk(g_plan_mu);
  std::shared_ptr<PrPlan> p = g->pr_plan;
  if (p && (p->H != H || p->transposed != g->transposed || p->bin_bits != bb ||
            p->range_bits != rb)) {
    g->pr_plan.reset();  // freed by its last holder
    p.reset();
  }
  if (!p) {
    PrPlan *np = new PrPlan;
    static std::atomic<uint64_t> serials{0};
    np->serial = ++serials;
    int r = plan_build(g, H, bb, rb, np);
    if (r != GG_OK) {
      std::string msg = gg_last_error();
      pr_plan_free(np);
      return fail(r, msg);
    }
    p = std::shared_ptr<PrPlan>(np, pr_plan_free);
    g->pr_plan = p;
  }
  *out = p;
  return GG_OK;
}

// Buffers of the per-iteration kernels, kept in the traversal context.
static int plan_buffers(const PrPlan *p, Work *w) {
  const uint64_t gs_n = (uint64_t)p->nsplit << p->bin_bits;
  if (w->pr_vals_cap < p->n_cold || w->pr_gsum_cap < gs_n || !w->pr_vals) {
    // an allocation and a memset must not land in a captured graph: the
    // captured paths size these buffers before capture starts
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GG_HIP(hipStreamIsCapturing(stream(), &cs));
    if (cs != hipStreamCaptureStatusNone)
      return fail(GG_EINVAL, "plan buffers would be allocated inside a stream capture");
    device_free(w->pr_vals);
    device_free(w->pr_gsum);
    w->pr_vals = w->pr_gsum = nullptr;
    w->pr_vals_cap = w->pr_gsum_cap = 0;
    GG_TRY(device_alloc((void **)&w->pr_vals, std::max<uint64_t>(p->n_cold, 1) * 8));
    GG_TRY(device_alloc((void **)&w->pr_gsum, std::max<uint64_t>(gs_n, 1) * 8));
    GG_HIP(hipMemsetAsync(w->pr_gsum, 0, std::max<uint64_t>(gs_n, 1) * 8, stream()));
    w->pr_vals_cap = p->n_cold;
    w->pr_gsum_cap = gs_n;
  }
  return GG_OK;
}

// One pass of the plan: emit the cold values of `src` (one VT per vertex),
// then the bins, then the split bins.
template <class Op>
static int plan_pass(const gg_graph *g, const PrPlan *p, Work *w, const typename Op::VT *src,
                     const Op &op, uint64_t *out, double *dacc, const uint32_t *gate,
                     const char *names[3]) {
  using T = typename Op::T;
  using VT = typename Op::VT;
  static_assert(sizeof(T) == 8, "plan accumulators are 8 bytes");
  static_assert(sizeof(VT) <= 8, "plan values fit the 8-byte cold buffer");
  const uint32_t V = g->vertices;
  GG_TRY(plan_buffers(p, w));
  VT *vals = reinterpret_cast<VT *>(w->pr_vals);
  T *gsum = reinterpret_cast<T *>(w->pr_gsum);
  if (p->nparts) {
    ProfScope ps(names[0]);
    hipLaunchKernelGGL(pr_emit_kernel<VT>, dim3(p->nparts), dim3(kPrBlock),
                       sizeof(VT) << p->range_bits, stream(), gate, p->parts, src, p->src16, p->H,
                       V, p->range_bits, vals);
  }
  if (p->nseg) {
    ProfScope ps(names[1]);
    hipLaunchKernelGGL(pr_bin_kernel<Op>, dim3(p->nseg), dim3(kPrBlock), (size_t)8 << p->bin_bits,
                       stream(), gate, p->hot_w, p->cold_w, p->segs, p->hot, p->run_beg, p->run_end, p->nranges, p->dst16,
                       vals, V, p->bin_bits, op, gsum, out, dacc);
  }
  if (p->nsplit) {
    ProfScope ps(names[2]);
    const uint64_t gs_n = (uint64_t)p->nsplit << p->bin_bits;
    hipLaunchKernelGGL(pr_split_finish_kernel<Op>, dim3(grid1(gs_n, 4096)), dim3(kBlock), 0,
                       stream(), gate, p->split_bin, p->nsplit, V, p->bin_bits, op, gsum, out, dacc);
  }
  GG_HIP(hipGetLastError());
  return GG_OK;
}

int pr_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const PrPull &op,
                      double *dacc) {
  static const char *names[3] = {"pr_emit", "pr_bins", "pr_split_finish"};
  return plan_pass(g, p, w, op.curr, PrSumOp{op}, nullptr, dacc, nullptr, names);
}

int cc_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const uint32_t *nlab,
                      uint32_t *label, uint64_t *out) {
  static const char *names[3] = {"cc_emit", "cc_bins", "cc_split_finish"};
  return plan_pass(g, p, w, nlab, CcMinOp{nlab, label}, out, nullptr, nullptr, names);
}

int sssp_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const uint32_t *snap,
                        float *dist, uint64_t *out) {
  if (!p->hot_w || (p->n_cold && !p->cold_w))
    return fail(GG_EINVAL, "SSSP plan step: the plan has no arc weights");
  static const char *names[3] = {"sssp_emit", "sssp_bins", "sssp_split_finish"};
  return plan_pass(g, p, w, snap, SsspMinOp{snap, dist}, out, nullptr, nullptr, names);
}

int ms_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const uint64_t *fcur,
                      const uint64_t *vis, uint64_t full, uint64_t *fnext, uint64_t *out,
                      const uint32_t *gate, const uint64_t *settled) {
  static const char *names[3] = {"ms_emit", "ms_bins", "ms_split_finish"};
  return plan_pass(g, p, w, fcur, MsOrOp{fcur, vis, fnext, full, settled}, out, nullptr, gate,
                   names);
}

}  // namespace gg

// PageRank as streams (toolkits/pagerank.cpp:24-129 computes the same sums).
//
// A PageRank iteration sums curr[u] over the in-arcs u -> v of every vertex.
// Gathered per arc, on RMAT-24 each iteration misses L2 about 86M times and
// moves ~12.5 GB for 3.7 GB of algorithmic bytes: the hubs' values stay in
// L2, but millions of low-degree sources each cost a 128-byte line for 8
// useful bytes.  On a relabelled graph (descending out-degree) the hubs are
// the low ids: the first 2^18 sources (2 MiB of FP64) carry about 69% of the
// arcs.  The plan splits the arcs by source:
//   hot (source < H): packed once into one array of (source, destination)
//     in destination order and streamed in chunks of 1024 arcs -- sixteen
//     contiguous arcs per lane, sixteen independent gathers in flight that
//     hit L2, segment sums within the lane and one segmented scan across the
//     wave;
//   cold: the cold arcs are kept as per-bin lists (bins of kPrBinDst
//     destinations), sorted by source.  Each iteration a workgroup loads the
//     values of one range of kPrRange cold sources into LDS (a coalesced
//     read) and fills that range's run of every bin list from LDS (coalesced
//     writes); then workgroups add each bin's (destination, value) list into
//     LDS;
//   finish: one pass over V applies the PageRank update to hot + cold.
// Streamed bytes per arc: hot 6 (+ an L2 or LDS hit), cold 4 + 8 + 8 + 2 = 22
// (+ 8 per cold source).
#include <string.h>

#include <algorithm>
#include <rocprim/rocprim.hpp>

#include "kernels.hpp"

namespace gg {

constexpr uint32_t kPrBinDst = 16384;     // destinations per bin: 128 KiB of LDS doubles
constexpr uint64_t kPrSegArcs = 1 << 17;  // arcs per accumulate workgroup
constexpr int kAccBlock = 1024;           // accumulate workgroup: one per CU (LDS)
constexpr int kHotBlock = 1024;           // hot stream workgroup: one per CU (LDS)
constexpr uint32_t kHub = 16384;          // hub values cached in LDS: 128 KiB
constexpr uint32_t kPrPer = 16;           // hot arcs per lane
constexpr uint32_t kPrHotChunk = 64 * kPrPer;
constexpr uint16_t kNoKey = 0xFFFF;       // padding arcs of the last hot chunk
constexpr uint32_t kPrRange = 16384;      // cold sources per copy workgroup: 128 KiB of LDS
constexpr int kCopyBlock = 1024;
constexpr uint32_t kCopyBins = 1024;      // bins per copy workgroup (LDS run table)
constexpr uint64_t kCopyArcs = 1 << 17;   // arcs per copy workgroup (ranges are split)

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

// One wave per destination: copies its hot prefix into the packed arrays.
__global__ void pr_pack_hot_kernel(const uint64_t *index, const uint32_t *nbr, uint32_t V,
                                   const uint32_t *hot_cnt, const uint64_t *hot_off,
                                   uint32_t *hot_src, uint32_t *dst32) {
  const uint32_t W = gridDim.x * (blockDim.x / 64);
  const int lane = lane_id();
  for (uint32_t v = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); v < V; v += W) {
    const uint64_t b = index[v], o = hot_off[v];
    const uint32_t n = hot_cnt[v];
    for (uint32_t i = lane; i < n; i += 64) {
      hot_src[o + i] = nbr[b + i];
      dst32[o + i] = v;
    }
  }
}

// Chunk-relative destinations and padding to whole chunks; *bad is set if a
// chunk spans kNoKey destinations or more.
__global__ void pr_hot_dst_kernel(const uint32_t *dst32, uint64_t n_hot, uint64_t n_pad,
                                  uint32_t *first, uint32_t *hot_src, uint16_t *hot_dst,
                                  uint32_t *bad) {
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n_pad; a += s) {
    if (a >= n_hot) {
      hot_src[a] = 0;
      hot_dst[a] = kNoKey;
      continue;
    }
    const uint32_t f = dst32[a - a % kPrHotChunk];
    if (a % kPrHotChunk == 0) first[a / kPrHotChunk] = f;
    const uint32_t off = dst32[a] - f;
    if (off >= kNoKey) atomicOr(bad, 1u);
    hot_dst[a] = (uint16_t)(off < kNoKey ? off : 0);
  }
}

// Bin of every cold out-arc (key) and its cold index (value), for a stable
// sort by bin; bin sizes by atomics.
__global__ void pr_bin_keys_kernel(const uint32_t *nbr, uint64_t cold_base, uint64_t n_cold,
                                   uint32_t *keys, uint32_t *vals, uint32_t *bin_cnt) {
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n_cold; a += s) {
    const uint32_t b = nbr[cold_base + a] / kPrBinDst;
    keys[a] = b;
    vals[a] = (uint32_t)a;
    atomicAdd(&bin_cnt[b], 1u);
  }
}

// Source of every cold out-arc (one wave per cold source).
__global__ void pr_arc_src_kernel(const uint64_t *out_index, uint32_t H, uint32_t V,
                                  uint64_t cold_base, uint32_t *arc_src) {
  const uint32_t W = gridDim.x * (blockDim.x / 64);
  const int lane = lane_id();
  for (uint32_t u = H + blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); u < V; u += W) {
    const uint64_t b = out_index[u], e = out_index[u + 1];
    for (uint64_t a = b + lane; a < e; a += 64) arc_src[a - cold_base] = u;
  }
}

// Bin lists in bin order: source and local destination of each position.
__global__ void pr_lists_kernel(const uint32_t *sorted_a, const uint32_t *arc_src,
                                const uint32_t *nbr, uint64_t cold_base, uint64_t n_cold,
                                uint32_t *src_list, uint16_t *dst_local) {
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n_cold; p += s) {
    const uint32_t a = sorted_a[p];
    src_list[p] = arc_src[a];
    dst_local[p] = (uint16_t)(nbr[cold_base + a] % kPrBinDst);
  }
}

// run_pos[r * nbins + b] = first position in bin b's list whose source is
// at least H + r * kPrRange (lists are sorted by source).
__global__ void pr_run_pos_kernel(const uint64_t *bin_off, const uint32_t *src_list, uint32_t H,
                                  uint32_t nbins, uint32_t nranges, uint64_t *run_pos) {
  const uint64_t total = (uint64_t)(nranges + 1) * nbins;
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += s) {
    const uint32_t r = (uint32_t)(i / nbins), b = (uint32_t)(i % nbins);
    const uint64_t key = (uint64_t)H + (uint64_t)r * kPrRange;
    uint64_t lo = bin_off[b], hi = bin_off[b + 1];
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (src_list[mid] < key)
        lo = mid + 1;
      else
        hi = mid;
    }
    run_pos[i] = lo;
  }
}

// -------------------------------------------------------------- iteration
// Cold copy: a workgroup takes one range of kPrRange cold sources and a
// strided subset of the bins (part q of Q: bins q, q + Q, ...), loads the
// range's values into LDS, gathers its runs of those bins' lists into an
// LDS table (start, exclusive-scanned length), and then copies the
// concatenated runs item by item: reads of the static source list and
// writes of the values run along each list.
__global__ __launch_bounds__(kCopyBlock) void pr_cold_copy_kernel(
    const double *curr, const uint32_t *src_list, const uint64_t *run_pos,
    const uint32_t *copy_blk, uint32_t H, uint32_t V, uint32_t nbins, double *vals) {
  __shared__ double val[kPrRange];
  __shared__ uint64_t rstart[kCopyBins];
  __shared__ uint32_t rscan[kCopyBins + 1];
  __shared__ uint32_t wsum[kCopyBlock / 64];
  const uint32_t info = copy_blk[2 * blockIdx.x], parts = copy_blk[2 * blockIdx.x + 1];
  const uint32_t r = info >> 16, q = info & 0xFFFF;
  const uint32_t u0 = H + r * kPrRange;
  const uint32_t n = V - u0 < kPrRange ? V - u0 : kPrRange;
  for (uint32_t i = threadIdx.x; i < n; i += kCopyBlock) val[i] = curr[u0 + i];
  // this workgroup's bins: q, q + parts, ... (at most kCopyBins of them)
  const uint32_t nb = (nbins - q + parts - 1) / parts;
  const uint64_t *pos = run_pos + (uint64_t)r * nbins;
  uint32_t len = 0;
  if (threadIdx.x < nb) {
    const uint32_t b = q + threadIdx.x * parts;
    rstart[threadIdx.x] = pos[b];
    len = (uint32_t)(pos[nbins + b] - pos[b]);
  }
  // block-wide exclusive scan of the run lengths
  uint32_t wtot;
  const uint32_t wex = wave_excl_scan(len, &wtot);
  if (lane_id() == 0) wsum[threadIdx.x >> 6] = wtot;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t i = 0; i < (threadIdx.x >> 6); i++) base += wsum[i];
  if (threadIdx.x < nb) rscan[threadIdx.x] = base + wex;
  uint32_t total = 0;
  for (uint32_t i = 0; i < kCopyBlock / 64; i++) total += wsum[i];
  if (threadIdx.x == 0) rscan[nb] = total;
  __syncthreads();
  constexpr int kU = 4;
  for (uint32_t j0 = threadIdx.x; j0 < total; j0 += kCopyBlock * kU) {
    uint64_t pp[kU];
    uint32_t su[kU];
#pragma unroll
    for (int t = 0; t < kU; t++) {
      const uint32_t j = j0 + t * kCopyBlock;
      pp[t] = ~0ULL;
      if (j < total) {
        uint32_t lo = 0, hi = nb;  // last run with rscan <= j
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (rscan[mid] <= j)
            lo = mid;
          else
            hi = mid;
        }
        pp[t] = rstart[lo] + (j - rscan[lo]);
      }
    }
#pragma unroll
    for (int t = 0; t < kU; t++) su[t] = pp[t] != ~0ULL ? stream_load(&src_list[pp[t]]) : u0;
#pragma unroll
    for (int t = 0; t < kU; t++)
      if (pp[t] != ~0ULL) vals[pp[t]] = val[su[t] - u0];
  }
}

// Cold accumulate: one workgroup per segment (at most kPrSegArcs arcs of one
// bin).  A bin with one segment stores its sums; segments of a larger bin
// add theirs (the sums start at zero every iteration).
__global__ __launch_bounds__(kAccBlock) void pr_cold_accumulate_kernel(
    const uint64_t *seg, const uint64_t *seg_end, const uint8_t *bin_multi,
    const uint16_t *dst_local, const double *vals, uint32_t V, double *cold) {
  __shared__ double sum[kPrBinDst];
  const uint64_t packed = seg[blockIdx.x];
  const uint32_t bin = (uint32_t)(packed >> 40);
  const uint64_t p0 = packed & ((1ULL << 40) - 1), p1 = seg_end[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < kPrBinDst; i += kAccBlock) sum[i] = 0.0;
  __syncthreads();
  constexpr int kU = 8;  // loads in flight per thread
  for (uint64_t p = p0 + threadIdx.x; p < p1; p += kAccBlock * kU) {
    uint16_t d[kU];
    double x[kU];
#pragma unroll
    for (int t = 0; t < kU; t++) {
      const uint64_t q = p + (uint64_t)t * kAccBlock;
      d[t] = q < p1 ? stream_load(&dst_local[q]) : (uint16_t)0;
      x[t] = q < p1 ? stream_load(&vals[q]) : 0.0;
    }
#pragma unroll
    for (int t = 0; t < kU; t++)
      if (p + (uint64_t)t * kAccBlock < p1) atomicAdd(&sum[d[t]], x[t]);
  }
  __syncthreads();
  const uint32_t base = bin * kPrBinDst;
  const uint32_t n = V - base < kPrBinDst ? V - base : kPrBinDst;
  if (!bin_multi[bin]) {
    for (uint32_t i = threadIdx.x; i < n; i += kAccBlock) cold[base + i] = sum[i];
  } else {
    for (uint32_t i = threadIdx.x; i < n; i += kAccBlock)
      if (sum[i] != 0.0) atomicAdd(&cold[base + i], sum[i]);
  }
}

// Hot stream: chunk c holds packed arcs [1024c, 1024c + 1024); lane L owns
// arcs 16L .. 16L+15.  Destinations ascend along the chunk, so each lane's
// arcs form runs of equal destination: runs inside a lane are complete
// segments and are stored; a lane's first and last runs are joined across
// lanes by a segmented scan.  Only a segment containing the chunk's first or
// last arc can continue into a neighbouring chunk: those are added with an
// atomic, every other segment is stored (the sums start at zero).
__global__ __launch_bounds__(kHotBlock) void pr_hot_stream_kernel(const uint32_t *hot_src,
                                                                  const uint16_t *hot_dst,
                                                                  const uint32_t *first,
                                                                  uint32_t nchunks,
                                                                  const double *curr, double *hot,
                                                                  uint32_t n_hub) {
  // the values of the n_hub hottest sources (about 30% of all arcs on RMAT),
  // kept in LDS: their gathers cost no L2 request
  __shared__ double hub[kHub];
  for (uint32_t i = threadIdx.x; i < n_hub; i += kHotBlock) hub[i] = curr[i];
  __syncthreads();
  const int lane = lane_id();
  constexpr uint32_t kWaves = kHotBlock / 64;
  const uint32_t W = gridDim.x * kWaves;
  for (uint32_t c = blockIdx.x * kWaves + (threadIdx.x >> 6); c < nchunks; c += W) {
    const uint64_t base = (uint64_t)c * kPrHotChunk + (uint64_t)lane * kPrPer;
    uint32_t s[kPrPer];
    uint16_t d[kPrPer];
    {
      const u32x4 *ps = reinterpret_cast<const u32x4 *>(hot_src + base);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const u32x4 v = stream_load(ps + q);
        s[4 * q] = v.x;
        s[4 * q + 1] = v.y;
        s[4 * q + 2] = v.z;
        s[4 * q + 3] = v.w;
      }
      const u32x4 *pd = reinterpret_cast<const u32x4 *>(hot_dst + base);
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const u32x4 v = stream_load(pd + q);
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int t = 0; t < 4; t++) {
          d[8 * q + 2 * t] = (uint16_t)(w4[t] & 0xFFFF);
          d[8 * q + 2 * t + 1] = (uint16_t)(w4[t] >> 16);
        }
      }
    }
    double x[kPrPer];
#pragma unroll
    for (int i = 0; i < (int)kPrPer; i++)
      x[i] = d[i] == kNoKey ? 0.0 : s[i] < n_hub ? hub[s[i]] : curr[s[i]];
    const uint32_t fdst = first[c];
    const uint32_t first_key = (uint32_t)__builtin_amdgcn_readlane((int)d[0], 0);
    const uint32_t last_key = (uint32_t)__builtin_amdgcn_readlane((int)d[kPrPer - 1], 63);
    // runs within the lane: running sums first (no stores among the loads'
    // consumers, so nothing waits on a store), then the stores
    double run[kPrPer];
#pragma unroll
    for (int i = 0; i < (int)kPrPer; i++)
      run[i] = i > 0 && d[i] == d[i - 1] ? run[i - 1] + x[i] : x[i];
    const uint32_t hk = d[0], tk = d[kPrPer - 1];
    const bool single = hk == tk;
    double hs = run[kPrPer - 1];
#pragma unroll
    for (int i = (int)kPrPer - 2; i >= 0; i--)
      if (d[i] == hk && d[i + 1] != hk) hs = run[i];
    const double ts = run[kPrPer - 1];
#pragma unroll
    for (int i = 0; i < (int)kPrPer - 1; i++)  // interior runs: whole segments
      if (d[i] != d[i + 1] && d[i] != hk && d[i] != kNoKey) hot[fdst + d[i]] = run[i];
    // R = sum of the segment of tk up to the end of this lane
    const uint32_t prev_tk = __shfl_up(tk, 1);
    const bool start = !single || lane == 0 || prev_tk != tk;
    int head = start ? lane : 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(head, off);
      if (lane >= off && y > head) head = y;
    }
    double R = ts;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double y = shfl_f64(R, lane >= off ? lane - off : lane);
      if (lane >= off && lane - off >= head) R += y;
    }
    const double prevR = shfl_f64(R, lane > 0 ? lane - 1 : 0);
    auto emit = [&](uint32_t key, double val) {
      if (key == first_key || key == last_key)
        atomicAdd(&hot[fdst + key], val);
      else
        hot[fdst + key] = val;
    };
    if (!single && hk != kNoKey) emit(hk, hs + (lane > 0 && prev_tk == hk ? prevR : 0.0));
    const uint32_t next_hk = __shfl_down(hk, 1);
    if (tk != kNoKey && (lane == 63 || next_hk != tk)) emit(tk, R);
  }
}

// finish(v, hot + cold) for every vertex, clearing both sums for the next
// iteration.
__global__ __launch_bounds__(kBlock) void pr_plan_finish_kernel(double *hot, double *cold,
                                                                uint32_t V, PrPull op,
                                                                double *dacc) {
  const uint32_t s = gridDim.x * blockDim.x;
  double dsum = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s) {
    const double t = hot[v] + cold[v];
    hot[v] = 0.0;
    cold[v] = 0.0;
    op.finish(v, t, 0, dsum);
  }
  const double w = wave_sum_f64(dsum);
  if (lane_id() == 0 && w != 0.0)
    unsafeAtomicAdd(&dacc[(global_wave() % kScalarSlots) * 16], w);
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

static int plan_hot(const gg_graph *g, PrPlan *p, bool *ok) {
  const Dir &in = in_dir(g);
  const uint32_t V = g->vertices;
  Scratch t;
  uint32_t *cnt = nullptr, *dst32 = nullptr, *bad = nullptr;
  uint64_t *off = nullptr;
  void *tmp = nullptr;
  GG_TRY(t.get((void **)&cnt, 4 * (size_t)V));
  GG_TRY(t.get((void **)&off, 8 * (size_t)V));
  GG_TRY(t.get((void **)&bad, 4));
  hipLaunchKernelGGL(pr_hot_count_kernel, dim3(grid1(V)), dim3(kBlock), 0, stream(), in.index,
                     in.nbr, V, p->H, cnt);
  size_t tmp_bytes = 0;
  GG_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, cnt, off, (uint64_t)0, (size_t)V,
                                 rocprim::plus<uint64_t>(), stream()));
  GG_TRY(t.get(&tmp, tmp_bytes));
  GG_HIP(rocprim::exclusive_scan(tmp, tmp_bytes, cnt, off, (uint64_t)0, (size_t)V,
                                 rocprim::plus<uint64_t>(), stream()));
  uint64_t last_off = 0;
  uint32_t last_cnt = 0;
  GG_HIP(hipMemcpyAsync(&last_off, off + V - 1, 8, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipMemcpyAsync(&last_cnt, cnt + V - 1, 4, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  p->n_hot = last_off + last_cnt;
  p->n_hchunks = (uint32_t)((p->n_hot + kPrHotChunk - 1) / kPrHotChunk);
  const uint64_t n_pad = std::max<uint64_t>((uint64_t)p->n_hchunks * kPrHotChunk, kPrHotChunk);
  GG_TRY(plan_alloc(p, (void **)&p->hot_src, n_pad * 4));
  GG_TRY(plan_alloc(p, (void **)&p->hot_dst, n_pad * 2));
  GG_TRY(plan_alloc(p, (void **)&p->hchunk_first, std::max<size_t>(p->n_hchunks, 1) * 4));
  GG_TRY(t.get((void **)&dst32, std::max<uint64_t>(p->n_hot, 1) * 4));
  GG_HIP(hipMemsetAsync(bad, 0, 4, stream()));
  hipLaunchKernelGGL(pr_pack_hot_kernel, dim3(grid1((uint64_t)V * 64)), dim3(kBlock), 0, stream(),
                     in.index, in.nbr, V, cnt, off, p->hot_src, dst32);
  hipLaunchKernelGGL(pr_hot_dst_kernel, dim3(grid1(n_pad)), dim3(kBlock), 0, stream(), dst32,
                     p->n_hot, n_pad, p->hchunk_first, p->hot_src, p->hot_dst, bad);
  uint32_t h_bad = 0;
  GG_HIP(hipMemcpyAsync(&h_bad, bad, 4, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  *ok = h_bad == 0;
  return GG_OK;
}

static int plan_cold(const gg_graph *g, PrPlan *p) {
  const Dir &out = out_dir(g);
  const uint32_t V = g->vertices;
  const uint64_t n = p->n_cold;
  p->nbins = (V + kPrBinDst - 1) / kPrBinDst;
  GG_TRY(plan_alloc(p, (void **)&p->src_list, std::max<uint64_t>(n, 1) * 4));
  GG_TRY(plan_alloc(p, (void **)&p->dst_local, std::max<uint64_t>(n, 1) * 2));
  Scratch t;
  uint32_t *keys = nullptr, *keys2 = nullptr, *vals = nullptr, *vals2 = nullptr, *cnt = nullptr;
  void *tmp = nullptr;
  GG_TRY(t.get((void **)&keys, std::max<uint64_t>(n, 1) * 4));
  GG_TRY(t.get((void **)&keys2, std::max<uint64_t>(n, 1) * 4));
  GG_TRY(t.get((void **)&vals, std::max<uint64_t>(n, 1) * 4));
  GG_TRY(t.get((void **)&vals2, std::max<uint64_t>(n, 1) * 4));
  GG_TRY(t.get((void **)&cnt, (size_t)p->nbins * 4));
  GG_HIP(hipMemsetAsync(cnt, 0, (size_t)p->nbins * 4, stream()));
  if (n) {
    hipLaunchKernelGGL(pr_bin_keys_kernel, dim3(grid1(n)), dim3(kBlock), 0, stream(), out.nbr,
                       p->cold_base, n, keys, vals, cnt);
    int bits = 1;
    while ((1u << bits) < p->nbins) bits++;
    rocprim::double_buffer<uint32_t> kb(keys, keys2);
    rocprim::double_buffer<uint32_t> vb(vals, vals2);
    size_t tmp_bytes = 0;
    GG_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kb, vb, (size_t)n, 0, bits, stream()));
    GG_TRY(t.get(&tmp, tmp_bytes));
    GG_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, (size_t)n, 0, bits, stream()));
    // keys2 is free again: the source of every cold arc
    uint32_t *arc_src = kb.current() == keys ? keys2 : keys;
    hipLaunchKernelGGL(pr_arc_src_kernel, dim3(grid1((uint64_t)(V - p->H) * 64)), dim3(kBlock), 0,
                       stream(), out.index, p->H, V, p->cold_base, arc_src);
    hipLaunchKernelGGL(pr_lists_kernel, dim3(grid1(n)), dim3(kBlock), 0, stream(), vb.current(),
                       arc_src, out.nbr, p->cold_base, n, p->src_list, p->dst_local);
  }
  std::vector<uint32_t> hc(p->nbins);
  GG_HIP(hipMemcpyAsync(hc.data(), cnt, (size_t)p->nbins * 4, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  // segments of at most kPrSegArcs arcs, none spanning two bins
  std::vector<uint64_t> seg, seg_end;
  std::vector<uint8_t> multi(p->nbins, 0);
  uint64_t pos = 0;
  for (uint32_t b = 0; b < p->nbins; b++) {
    const uint64_t e = pos + hc[b];
    multi[b] = hc[b] > kPrSegArcs;
    for (uint64_t q = pos; q < e; q += kPrSegArcs) {
      seg.push_back(((uint64_t)b << 40) | q);
      seg_end.push_back(std::min(q + kPrSegArcs, e));
    }
    pos = e;
  }
  p->nseg = (uint32_t)seg.size();
  GG_TRY(plan_alloc(p, (void **)&p->seg, std::max<size_t>(seg.size(), 1) * 8));
  GG_TRY(plan_alloc(p, (void **)&p->seg_end, std::max<size_t>(seg.size(), 1) * 8));
  GG_TRY(plan_alloc(p, (void **)&p->bin_multi, p->nbins));
  if (!seg.empty()) {
    GG_HIP(hipMemcpyAsync(p->seg, seg.data(), seg.size() * 8, hipMemcpyHostToDevice, stream()));
    GG_HIP(hipMemcpyAsync(p->seg_end, seg_end.data(), seg.size() * 8, hipMemcpyHostToDevice,
                          stream()));
  }
  GG_HIP(hipMemcpyAsync(p->bin_multi, multi.data(), p->nbins, hipMemcpyHostToDevice, stream()));
  // copy ranges over the cold sources that have out-arcs
  uint64_t *boff = nullptr;
  GG_TRY(t.get((void **)&boff, ((size_t)p->nbins + 1) * 8));
  std::vector<uint64_t> off((size_t)p->nbins + 1, 0);
  for (uint32_t b = 0; b < p->nbins; b++) off[b + 1] = off[b] + hc[b];
  GG_HIP(hipMemcpyAsync(boff, off.data(), off.size() * 8, hipMemcpyHostToDevice, stream()));
  p->nranges = (uint32_t)(((uint64_t)V - p->H + kPrRange - 1) / kPrRange);
  const uint64_t npos = (uint64_t)(p->nranges + 1) * p->nbins;
  GG_TRY(plan_alloc(p, (void **)&p->run_pos, npos * 8));
  hipLaunchKernelGGL(pr_run_pos_kernel, dim3(grid1(npos)), dim3(kBlock), 0, stream(), boff,
                     p->src_list, p->H, p->nbins, p->nranges, p->run_pos);
  // copy workgroups: a range is split into parts of about kCopyArcs arcs and
  // of at most kCopyBins bins
  std::vector<uint64_t> rb((size_t)p->nranges + 1);
  GG_HIP(hipMemcpy2DAsync(rb.data(), 8, out.index + p->H, (size_t)kPrRange * 8, 8, p->nranges,
                          hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  rb[p->nranges] = g->arcs;
  std::vector<uint32_t> blk;
  for (uint32_t r = 0; r < p->nranges; r++) {
    const uint64_t arcs = rb[r + 1] - rb[r];
    uint32_t parts = (uint32_t)std::max<uint64_t>(1, (arcs + kCopyArcs - 1) / kCopyArcs);
    parts = std::max(parts, (p->nbins + kCopyBins - 1) / kCopyBins);
    parts = std::min(parts, std::min(p->nbins, 0xFFFFu));
    if (!arcs) continue;
    for (uint32_t q = 0; q < parts; q++) {
      blk.push_back(r << 16 | q);
      blk.push_back(parts);
    }
  }
  if (p->nranges >= (1u << 16)) return fail(GG_EINVAL, "PageRank plan: too many source ranges");
  p->ncopy = (uint32_t)(blk.size() / 2);
  GG_TRY(plan_alloc(p, (void **)&p->copy_blk, std::max<size_t>(blk.size(), 2) * 4));
  if (!blk.empty())
    GG_HIP(hipMemcpyAsync(p->copy_blk, blk.data(), blk.size() * 4, hipMemcpyHostToDevice, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

static int plan_build(const gg_graph *g, uint32_t H, PrPlan *p, bool *ok) {
  const Dir &out = out_dir(g);
  p->transposed = g->transposed;
  p->H = H;
  GG_HIP(hipMemcpyAsync(&p->cold_base, out.index + H, 8, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  p->n_cold = g->arcs - p->cold_base;
  GG_TRY(plan_hot(g, p, ok));
  if (!*ok) return GG_OK;
  return plan_cold(g, p);
}

int pr_plan_get(gg_graph *g, const PrPlan **out) {
  *out = nullptr;
  const int64_t hs = options().pr_hot_sources;
  if (!g->reordered || hs <= 0 || (uint64_t)hs >= g->vertices) return GG_OK;
  if (g->arcs >= ((uint64_t)1 << 32)) return GG_OK;  // slots and sort values are 32-bit
  const uint32_t H = (uint32_t)hs;
  PrPlan *p = g->pr_plan;
  if (p && (p->H != H || p->transposed != g->transposed)) {
    pr_plan_free(p);
    g->pr_plan = p = nullptr;
  }
  if (!p) {
    p = new PrPlan;
    bool ok = true;
    int r = plan_build(g, H, p, &ok);
    if (r != GG_OK) {
      std::string msg = gg_last_error();
      pr_plan_free(p);
      return fail(r, msg);
    }
    if (!ok) {  // a hot chunk spans too many destinations: keep the plain pull
      pr_plan_free(p);
      p = new PrPlan;
      p->transposed = g->transposed;
      p->H = H;
    }
    g->pr_plan = p;
  }
  if (p->n_hchunks && p->n_cold) *out = p;
  return GG_OK;
}

int pr_plan_iteration(const gg_graph *g, const PrPlan *p, Work *w, const PrPull &op,
                      double *dacc) {
  const uint32_t V = g->vertices;
  if (w->pr_vals_cap < p->n_cold || !w->pr_cold || !w->pr_hot) {
    device_free(w->pr_vals);
    device_free(w->pr_cold);
    device_free(w->pr_hot);
    w->pr_vals = w->pr_cold = w->pr_hot = nullptr;
    w->pr_vals_cap = 0;
    GG_TRY(device_alloc((void **)&w->pr_vals, p->n_cold * 8));
    GG_TRY(device_alloc((void **)&w->pr_cold, (size_t)V * 8));
    GG_TRY(device_alloc((void **)&w->pr_hot, (size_t)V * 8));
    GG_HIP(hipMemsetAsync(w->pr_cold, 0, (size_t)V * 8, stream()));
    GG_HIP(hipMemsetAsync(w->pr_hot, 0, (size_t)V * 8, stream()));
    w->pr_vals_cap = p->n_cold;
  }
  {
    ProfScope ps("pr_cold_scatter");
    if (p->ncopy)
      hipLaunchKernelGGL(pr_cold_copy_kernel, dim3(p->ncopy), dim3(kCopyBlock), 0, stream(),
                         op.curr, p->src_list, p->run_pos, p->copy_blk, p->H, V, p->nbins,
                         w->pr_vals);
  }
  if (p->nseg) {
    ProfScope ps("pr_cold_accumulate");
    hipLaunchKernelGGL(pr_cold_accumulate_kernel, dim3(p->nseg), dim3(kAccBlock), 0, stream(),
                       p->seg, p->seg_end, p->bin_multi, p->dst_local, w->pr_vals, V, w->pr_cold);
  }
  {
    ProfScope ps("pr_hot_stream");
    int cus = 0;
    GG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    const uint32_t blocks = std::min<uint32_t>((uint32_t)std::max(cus, 1),
                                               (p->n_hchunks + kHotBlock / 64 - 1) / (kHotBlock / 64));
    hipLaunchKernelGGL(pr_hot_stream_kernel, dim3(std::max(blocks, 1u)), dim3(kHotBlock), 0,
                       stream(), p->hot_src, p->hot_dst, p->hchunk_first, p->n_hchunks, op.curr,
                       w->pr_hot, std::min(p->H, kHub));
  }
  {
    ProfScope ps("pr_finish");
    hipLaunchKernelGGL(pr_plan_finish_kernel, dim3(grid1(V, 4096)), dim3(kBlock), 0, stream(),
                       w->pr_hot, w->pr_cold, V, op, dacc);
  }
  GG_HIP(hipGetLastError());
  return GG_OK;
}

}  // namespace gg

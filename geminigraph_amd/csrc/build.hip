// Synthetic RMAT input and the on-device CSR/CSC build.
//
// Replaces Graph<E>::load_directed (core/graph.hpp:772-1339) and
// Graph<E>::load_undirected_from_directed (core/graph.hpp:350-757): the
// reference reads the file five times and builds each direction with a serial
// count pass, a prefix sum and an atomic scatter whose neighbour order is
// nondeterministic.  Here the edge list is uploaded once; each direction is
// one radix sort of packed (key << vbits | neighbour) 64-bit words, so the
// adjacency lists come out sorted by neighbour, ties in input order (stable
// sort), which makes every run reproducible.  Degrees are the run lengths
// of the sorted endpoint ids, offsets an exclusive scan.
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <rocprim/rocprim.hpp>

#include "gg_internal.hpp"

namespace gg {

// ------------------------------------------------------------------- RMAT
// Definition shared with the CPU oracle (oracle/rmat.c): splitmix64 in
// counter mode, 32 random bits per level, integer quadrant thresholds.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t mix64h(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct PermKeys {
  uint64_t mask, m1, a1, m2, a2, m3, a3;
  int h;
};

static PermKeys perm_keys(int scale, uint64_t seed) {
  const uint64_t G = 0x9E3779B97F4A7C15ULL;
  PermKeys p;
  p.mask = (scale >= 32) ? 0xFFFFFFFFULL : ((1ULL << scale) - 1ULL);
  uint64_t k1 = mix64h(seed ^ 0x632BE59BD9B4E019ULL);
  uint64_t k2 = mix64h(k1 + G);
  uint64_t k3 = mix64h(k2 + G);
  p.m1 = (k1 & p.mask) | 1ULL;
  p.a1 = (k1 >> 32) & p.mask;
  p.m2 = (k2 & p.mask) | 1ULL;
  p.a2 = (k2 >> 32) & p.mask;
  p.m3 = (k3 & p.mask) | 1ULL;
  p.a3 = (k3 >> 32) & p.mask;
  p.h = (scale + 1) / 2;
  return p;
}

__device__ __forceinline__ uint32_t permute(uint32_t x, const PermKeys &p) {
  uint64_t v = x;
  v = (v * p.m1 + p.a1) & p.mask;
  v ^= v >> p.h;
  v = (v * p.m2 + p.a2) & p.mask;
  v ^= v >> p.h;
  v = (v * p.m3 + p.a3) & p.mask;
  return (uint32_t)v;
}

__global__ void rmat_kernel(uint32_t *out, uint64_t ne, int scale, uint64_t key, uint64_t ta,
                            uint64_t tab, uint64_t tabc, int do_perm, PermKeys pk,
                            int weighted) {
  const uint64_t G = 0x9E3779B97F4A7C15ULL;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += stride) {
    uint32_t src = 0, dst = 0;
    uint64_t w = 0;
    for (int l = 0; l < scale; l++) {
      if ((l & 1) == 0) w = mix64(key + (16ULL * i + (uint64_t)(l >> 1) + 1ULL) * G);
      uint64_t r = (l & 1) ? (w >> 32) : (w & 0xFFFFFFFFULL);
      uint32_t bit = 1u << (scale - 1 - l);
      if (r < ta) {
      } else if (r < tab) {
        dst |= bit;
      } else if (r < tabc) {
        src |= bit;
      } else {
        src |= bit;
        dst |= bit;
      }
    }
    if (do_perm) {
      src = permute(src, pk);
      dst = permute(dst, pk);
    }
    if (weighted) {
      uint64_t k = mix64(key + (16ULL * i + 15ULL + 1ULL) * G) >> 40;
      float wt = (float)(k + 1ULL) * (1.0f / 16777216.0f);
      uint32_t *r = out + 3 * i;
      r[0] = src;
      r[1] = dst;
      r[2] = __float_as_uint(wt);
    } else {
      uint32_t *r = out + 2 * i;
      r[0] = src;
      r[1] = dst;
    }
  }
}

// ------------------------------------------------------------------ build
// Degrees by sorting endpoints (the reference's degree pass is a counting
// loop, core/graph.hpp:799-822): an atomic histogram of 2|E| increments on
// random addresses runs at the chip's random-atomic rate (~25 G/s, 21.6 ms
// for RMAT-24); sorting the endpoint ids and measuring their runs streams.
// which = 0: sources, 1: destinations, 2: both (2|E| ids, symmetric graphs).
__global__ void endpoint_kernel(const uint32_t *edges, uint64_t ne, int rec_words, int which,
                                uint32_t *ids) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += stride) {
    if (which == 2) {
      ids[2 * i] = edges[i * rec_words];
      ids[2 * i + 1] = edges[i * rec_words + 1];
    } else {
      ids[i] = edges[i * rec_words + which];
    }
  }
}
// start[v] = the position of v's run in the sorted ids
__global__ void run_start_kernel(const uint32_t *ids, uint64_t n, uint64_t *start) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (i == 0 || ids[i - 1] != ids[i]) start[ids[i]] = i;
}
// deg[v] = the length of v's run (untouched, so 0, for absent ids)
__global__ void run_length_kernel(const uint32_t *ids, uint64_t n, const uint64_t *start,
                                  uint32_t *deg) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (i + 1 == n || ids[i + 1] != ids[i]) deg[ids[i]] = (uint32_t)(i + 1 - start[ids[i]]);
}

// Endpoints must be vertices (bit 1).  Weights must be >= 0 and not NaN
// (bit 2): the SSSP kernels keep f32 distances as unsigned bit patterns
// (atomicMin, max of complements), an order that holds for non-negative
// floats only, where the reference's CAS-loop write_min
// (core/atomic.hpp:35-41) compares any floats.  -0.0 passes (its sums equal
// +0.0's); +inf passes (it never lowers a distance).
__global__ void check_range_kernel(const uint32_t *edges, uint64_t ne, int rec_words, uint32_t V,
                                   uint32_t *bad) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += stride) {
    uint32_t s = edges[i * rec_words], d = edges[i * rec_words + 1];
    if (s >= V || d >= V) atomicOr(bad, 1u);
    if (rec_words == 3 && !weight_ok(edges[i * rec_words + 2])) atomicOr(bad, 2u);
  }
}

// keys for one direction: key vertex in the high bits, neighbour in the low
// vbits.  mode 0: key = src, 1: key = dst, 2: both (symmetric, 2E keys).
// perm (relabeling) renames both endpoints; rank (hub-first lists without
// relabeling) replaces only the neighbour by its degree rank, so each list
// sorts hubs first and split_keys_kernel maps the ranks back to ids.
__global__ void make_keys_kernel(const uint32_t *edges, uint64_t ne, int rec_words, int vbits,
                                 int mode, const uint32_t *perm, const uint32_t *rank,
                                 uint64_t *keys, float *vals) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += stride) {
    uint64_t s = edges[i * rec_words], d = edges[i * rec_words + 1];
    if (perm) {
      s = perm[s];
      d = perm[d];
    }
    const uint64_t rs = rank ? rank[s] : s, rd = rank ? rank[d] : d;
    float w = rec_words == 3 ? __uint_as_float(edges[i * rec_words + 2]) : 0.f;
    if (mode == 0) {
      keys[i] = (s << vbits) | rd;
      if (vals) vals[i] = w;
    } else if (mode == 1) {
      keys[i] = (d << vbits) | rs;
      if (vals) vals[i] = w;
    } else {
      keys[2 * i] = (s << vbits) | rd;
      keys[2 * i + 1] = (d << vbits) | rs;
      if (vals) {
        vals[2 * i] = w;
        vals[2 * i + 1] = w;
      }
    }
  }
}

// Relabeling helpers: sort key = complemented out-degree (descending order),
// then vertices with in-arcs before those without, then id (the stable
// sort), and perm = inverse of the sorted ids.  Half of an RMAT graph's
// vertices have no in-arcs; grouped at the end of their out-degree class
// (the isolated ones, out-degree 0, form the tail of the ids) the bottom-up
// BFS step skips their groups whole instead of reading records beside the
// vertices it tests.  in_deg is null on a symmetric graph (no in-arcs iff
// no arcs).
__global__ void relabel_keys_kernel(const uint32_t *deg, const uint32_t *in_deg, uint32_t V,
                                    uint64_t *keys, uint32_t *ids) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s) {
    const uint32_t noin = (in_deg ? in_deg[v] : deg[v]) == 0;
    keys[v] = ((uint64_t)~deg[v] << 1) | noin;
    ids[v] = v;
  }
}
// Hybrid order: the vertices of degree rank < hubs keep that rank; the others
// follow in file order, those without in-arcs last.  Consecutive internal
// ids outside the hubs then have ascending file ids, so outputs indexed by
// file id (BFS parents, depths) are written near one another by the waves
// that find neighbouring vertices.
__global__ void hybrid_keys_kernel(const uint32_t *rank, const uint32_t *deg,
                                   const uint32_t *in_deg, uint32_t V, uint32_t hubs,
                                   uint64_t *keys, uint32_t *ids) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s) {
    const uint64_t noin = (in_deg ? in_deg[v] : deg[v]) == 0;
    keys[v] = rank[v] < hubs ? rank[v] : hubs + (noin << 32) + v;
    ids[v] = v;
  }
}
__global__ void invert_perm_kernel(const uint32_t *iperm, uint32_t V, uint32_t *perm) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < V; n += s) perm[iperm[n]] = n;
}
__global__ void gather_u32_kernel(const uint32_t *src, const uint32_t *idx, uint32_t V,
                                  uint32_t *dst) {
  uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < V; n += s) dst[n] = src[idx[n]];
}

__global__ void split_keys_kernel(const uint64_t *keys, uint64_t n, int vbits,
                                  const uint32_t *unrank, uint32_t *nbr) {
  const uint64_t mask = (1ULL << vbits) - 1ULL;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t low = (uint32_t)(keys[i] & mask);
    nbr[i] = unrank ? unrank[low] : low;
  }
}

static unsigned grid_for(uint64_t n, unsigned block = 256, unsigned cap = 1u << 20) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Exclusive scan of u32 degrees into u64 offsets (V+1 entries).
static int scan_degrees(const uint32_t *deg, uint64_t *index, uint32_t V) {
  size_t tmp_bytes = 0;
  auto in = rocprim::make_transform_iterator(
      deg, [] __device__(uint32_t x) -> uint64_t { return (uint64_t)x; });
  GG_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, in, index, (uint64_t)0, (size_t)V + 1,
                                 rocprim::plus<uint64_t>(), stream()));
  void *tmp = nullptr;
  GG_TRY(device_alloc(&tmp, tmp_bytes));
  hipError_t e = rocprim::exclusive_scan(tmp, tmp_bytes, in, index, (uint64_t)0, (size_t)V + 1,
                                         rocprim::plus<uint64_t>(), stream());
  hipStreamSynchronize(stream());
  device_free(tmp);
  if (e != hipSuccess) return fail(GG_EHIP, std::string("exclusive_scan: ") + hipGetErrorString(e));
  return GG_OK;
}

static int sort_keys(uint64_t *&keys, uint64_t *&keys_alt, float *&vals, float *&vals_alt,
                     uint64_t n, int end_bit) {
  size_t tmp_bytes = 0;
  rocprim::double_buffer<uint64_t> kb(keys, keys_alt);
  rocprim::double_buffer<float> vb(vals, vals_alt);
  if (vals)
    GG_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kb, vb, (size_t)n, 0, end_bit, stream()));
  else
    GG_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, kb, (size_t)n, 0, end_bit, stream()));
  void *tmp = nullptr;
  GG_TRY(device_alloc(&tmp, tmp_bytes));
  hipError_t e;
  if (vals)
    e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, (size_t)n, 0, end_bit, stream());
  else
    e = rocprim::radix_sort_keys(tmp, tmp_bytes, kb, (size_t)n, 0, end_bit, stream());
  hipStreamSynchronize(stream());
  device_free(tmp);
  if (e != hipSuccess) return fail(GG_EHIP, std::string("radix_sort: ") + hipGetErrorString(e));
  keys = kb.current();
  keys_alt = kb.alternate();
  if (vals) {
    vals = vb.current();
    vals_alt = vb.alternate();
  }
  return GG_OK;
}

// deg (V+1 entries, zeroed by the caller) of one endpoint of every record;
// start: V+1 scratch u64 (the offsets array before its scan).
static int degrees_by_sort(const uint32_t *edges, uint64_t ne, int rec_words, int which,
                           uint32_t V, uint32_t *deg, uint64_t *start) {
  const uint64_t n = which == 2 ? 2 * ne : ne;
  if (!n) return GG_OK;
  int vbits = 1;
  while (((uint64_t)1 << vbits) < (uint64_t)V) vbits++;
  uint32_t *ids = nullptr, *ids_alt = nullptr;
  void *tmp = nullptr;
  int r = device_alloc((void **)&ids, n * 4);
  if (r == GG_OK) r = device_alloc((void **)&ids_alt, n * 4);
  if (r == GG_OK) {
    hipLaunchKernelGGL(endpoint_kernel, dim3(grid_for(ne)), dim3(256), 0, stream(), edges, ne,
                       rec_words, which, ids);
    rocprim::double_buffer<uint32_t> kb(ids, ids_alt);
    size_t tmp_bytes = 0;
    hipError_t e = rocprim::radix_sort_keys(nullptr, tmp_bytes, kb, (size_t)n, 0, vbits, stream());
    if (e == hipSuccess) r = device_alloc(&tmp, tmp_bytes);
    if (e == hipSuccess && r == GG_OK)
      e = rocprim::radix_sort_keys(tmp, tmp_bytes, kb, (size_t)n, 0, vbits, stream());
    if (e == hipSuccess && r == GG_OK) {
      hipLaunchKernelGGL(run_start_kernel, dim3(grid_for(n)), dim3(256), 0, stream(),
                         kb.current(), n, start);
      hipLaunchKernelGGL(run_length_kernel, dim3(grid_for(n)), dim3(256), 0, stream(),
                         kb.current(), n, start, deg);
      e = hipStreamSynchronize(stream());
    }
    if (e != hipSuccess && r == GG_OK)
      r = fail(GG_EHIP, std::string("degrees: ") + hipGetErrorString(e));
  }
  device_free(tmp);
  device_free(ids);
  device_free(ids_alt);
  return r;
}

int graph_alloc(gg_graph *g, void **p, size_t bytes) {
  GG_TRY(device_alloc(p, bytes));
  g->allocations.push_back(*p);
  g->device_bytes += bytes;
  return GG_OK;
}

__global__ void vinfo_kernel(const uint64_t *index, const uint32_t *nbr, const uint32_t *other_deg,
                             const uint32_t *iperm, uint32_t V, uint4 *vinfo,
                             uint32_t *first_orig) {
  const uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s) {
    const uint64_t b = index[v], e = index[v + 1];
    const uint32_t f = b < e ? nbr[b] : kUnset;
    const uint32_t od = other_deg[v] < kDegEscape ? other_deg[v] : kDegEscape;
    const uint32_t id = e - b < kDegEscape ? (uint32_t)(e - b) : kDegEscape;
    vinfo[v] = make_uint4(iperm ? iperm[v] : v, (od << 16) | id, f, (uint32_t)b);
    first_orig[v] = f == kUnset ? V : (iperm ? iperm[f] : f);
  }
}

// Per-vertex records of one direction (after both directions are built).
static int build_vinfo(gg_graph *g, Dir &d, const Dir &other) {
  const uint32_t V = g->vertices;
  GG_TRY(graph_alloc(g, (void **)&d.vinfo, (size_t)std::max(V, 1u) * sizeof(uint4)));
  GG_TRY(graph_alloc(g, (void **)&d.first_orig, (size_t)std::max(V, 1u) * 4));
  const unsigned grid = (unsigned)std::min<uint64_t>(((uint64_t)V + kBlock - 1) / kBlock, 65536);
  hipLaunchKernelGGL(vinfo_kernel, dim3(std::max(grid, 1u)), dim3(kBlock), 0, stream(), d.index,
                     d.nbr, other.degree, g->reordered ? g->iperm : nullptr, V, d.vinfo,
                     d.first_orig);
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(stream()));  // the graph may be used from other streams
  return GG_OK;
}

// Heavy-vertex chunk list of one direction (degree > light_max), computed on
// the host from the degree array: the dense-mode analogue of tune_chunks
// (core/graph.hpp:1341-1404), which balances threads by in-edges.
static int build_chunks(gg_graph *g, Dir &d) {
  uint32_t V = g->vertices;
  d.light_max = (uint32_t)options().pull_light_max;
  std::vector<uint32_t> deg(V);
  std::vector<uint64_t> idx((size_t)V + 1);
  GG_HIP(hipMemcpyAsync(deg.data(), d.degree, sizeof(uint32_t) * V, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipMemcpyAsync(idx.data(), d.index, sizeof(uint64_t) * ((size_t)V + 1),
                        hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  std::vector<uint32_t> cv, ch, hv, hf;
  std::vector<uint64_t> cb, ce;
  uint32_t last = 0;  // one past the last vertex with arcs in this direction
  for (uint32_t v = V; v > 0; v--)
    if (deg[v - 1]) {
      last = v;
      break;
    }
  d.live_groups = (last + 63) / 64;
  for (uint32_t v = 0; v < V; v++) {
    if (deg[v] <= d.light_max) continue;
    uint32_t h = (uint32_t)hv.size();
    hv.push_back(v);
    hf.push_back((uint32_t)cv.size());
    for (uint64_t b = idx[v]; b < idx[v + 1]; b += kChunkEdges) {
      cv.push_back(v);
      ch.push_back(h);
      cb.push_back(b);
      ce.push_back(std::min<uint64_t>(b + kChunkEdges, idx[v + 1]));
    }
  }
  // vertices without arcs in this direction (a bottom-up BFS never finds them)
  std::vector<uint64_t> eb((size_t)V / 64 + 1, 0);
  for (uint32_t v = 0; v < V; v++)
    if (!deg[v]) eb[v >> 6] |= 1ULL << (v & 63);
  GG_TRY(graph_alloc(g, (void **)&d.empty, eb.size() * 8));
  GG_HIP(hipMemcpyAsync(d.empty, eb.data(), eb.size() * 8, hipMemcpyHostToDevice, stream()));
  d.n_heavy = (uint32_t)hv.size();
  d.n_chunks = (uint32_t)cv.size();
  hf.push_back((uint32_t)cv.size());
  size_t nh = std::max<size_t>(hv.size(), 1), nc = std::max<size_t>(cv.size(), 1);
  GG_TRY(graph_alloc(g, (void **)&d.heavy_v, nh * 4));
  GG_TRY(graph_alloc(g, (void **)&d.heavy_first, hf.size() * 4));
  GG_HIP(hipMemcpyAsync(d.heavy_first, hf.data(), hf.size() * 4, hipMemcpyHostToDevice, stream()));
  GG_TRY(graph_alloc(g, (void **)&d.chunk_v, nc * 4));
  GG_TRY(graph_alloc(g, (void **)&d.chunk_h, nc * 4));
  GG_TRY(graph_alloc(g, (void **)&d.chunk_beg, nc * 8));
  GG_TRY(graph_alloc(g, (void **)&d.chunk_end, nc * 8));
  if (!hv.empty()) {
    GG_HIP(hipMemcpyAsync(d.heavy_v, hv.data(), hv.size() * 4, hipMemcpyHostToDevice, stream()));
    GG_HIP(hipMemcpyAsync(d.chunk_v, cv.data(), cv.size() * 4, hipMemcpyHostToDevice, stream()));
    GG_HIP(hipMemcpyAsync(d.chunk_h, ch.data(), ch.size() * 4, hipMemcpyHostToDevice, stream()));
    GG_HIP(hipMemcpyAsync(d.chunk_beg, cb.data(), cb.size() * 8, hipMemcpyHostToDevice, stream()));
    GG_HIP(hipMemcpyAsync(d.chunk_end, ce.data(), ce.size() * 8, hipMemcpyHostToDevice, stream()));
  }
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

// What every graph derives from its CSR/CSC: heavy chunk lists, the
// empty-list bitmaps and the per-vertex records (symmetric graphs alias the
// in direction to the out direction).
int graph_finish(gg_graph *g) {
  GG_TRY(build_chunks(g, g->out));
  if (g->symmetric) {
    GG_TRY(build_vinfo(g, g->out, g->out));
    g->in = g->out;
    return GG_OK;
  }
  GG_TRY(build_chunks(g, g->in));
  // the first in-neighbour: a bottom-up BFS step tests it before reading the
  // list (lists are sorted, so on a relabelled graph it is the highest-degree
  // in-neighbour, the likeliest to be in the frontier)
  GG_TRY(build_vinfo(g, g->out, g->in));
  return build_vinfo(g, g->in, g->out);
}

// One direction from the device edge list.
static int build_dir(gg_graph *g, Dir &d, const uint32_t *edges, uint64_t ne, int rec_words,
                     int mode, const uint32_t *rank, const uint32_t *unrank) {
  uint32_t V = g->vertices;
  int vbits = 1;
  while (((uint64_t)1 << vbits) < (uint64_t)V) vbits++;
  uint64_t n = mode == 2 ? 2 * ne : ne;
  uint64_t *keys = nullptr, *keys_alt = nullptr;
  float *vals = nullptr, *vals_alt = nullptr;
  int r = device_alloc((void **)&keys, n * 8);
  if (r == GG_OK) r = device_alloc((void **)&keys_alt, n * 8);
  if (r == GG_OK && g->weighted) r = device_alloc((void **)&vals, n * 4);
  if (r == GG_OK && g->weighted) r = device_alloc((void **)&vals_alt, n * 4);
  if (r == GG_OK) {
    hipLaunchKernelGGL(make_keys_kernel, dim3(grid_for(ne)), dim3(256), 0, stream(), edges, ne,
                       rec_words, vbits, mode, g->reordered ? g->perm : nullptr, rank, keys,
                       vals);
    r = sort_keys(keys, keys_alt, vals, vals_alt, n, 2 * vbits);
  }
  device_free(keys_alt);
  device_free(vals_alt);
  if (r == GG_OK) r = graph_alloc(g, (void **)&d.nbr, n * 4);
  if (r == GG_OK) {
    hipLaunchKernelGGL(split_keys_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), keys, n,
                       vbits, unrank, d.nbr);
    hipError_t e = hipStreamSynchronize(stream());
    if (e != hipSuccess) r = fail(GG_EHIP, std::string("split_keys: ") + hipGetErrorString(e));
  }
  device_free(keys);
  if (r == GG_OK && g->weighted) {
    d.w = vals;  // sorted weights become the adjacency weights
    g->allocations.push_back(vals);
    g->device_bytes += n * 4;
  } else {
    device_free(vals);
  }
  return r;
}

// Degree order: order[k] = the k-th vertex by descending out-degree (ties:
// vertices with in-arcs first, then by id), rank its inverse.  With hubs in
// (0, V), the hybrid order of hybrid_keys_kernel instead.
static int degree_order(gg_graph *g, uint32_t *order, uint32_t *rank, uint32_t hubs = 0) {
  uint32_t V = g->vertices;
  uint64_t *keys = nullptr, *keys_alt = nullptr;
  uint32_t *ids = nullptr, *ids_alt = nullptr;
  int r = device_alloc((void **)&keys, (size_t)V * 8);
  if (r == GG_OK) r = device_alloc((void **)&keys_alt, (size_t)V * 8);
  if (r == GG_OK) r = device_alloc((void **)&ids, (size_t)V * 4);
  if (r == GG_OK) r = device_alloc((void **)&ids_alt, (size_t)V * 4);
  void *tmp = nullptr;
  // sorts (keys, ids) over end_bit bits; order = the sorted ids, rank its inverse
  auto sort_order = [&](int end_bit) -> hipError_t {
    rocprim::double_buffer<uint64_t> kb(keys, keys_alt);
    rocprim::double_buffer<uint32_t> vb(ids, ids_alt);
    size_t tmp_bytes = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, kb, vb, (size_t)V, 0, end_bit,
                                             stream());
    if (e == hipSuccess && !tmp && device_alloc(&tmp, tmp_bytes) != GG_OK) return hipErrorOutOfMemory;
    if (e == hipSuccess)
      e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, (size_t)V, 0, end_bit, stream());
    if (e == hipSuccess)
      e = hipMemcpyAsync(order, vb.current(), (size_t)V * 4, hipMemcpyDeviceToDevice, stream());
    if (e == hipSuccess)
      hipLaunchKernelGGL(invert_perm_kernel, dim3(grid_for(V)), dim3(256), 0, stream(), order, V,
                         rank);
    return e;
  };
  if (r == GG_OK) {
    const uint32_t *in_deg = g->symmetric ? nullptr : g->in.degree;
    hipLaunchKernelGGL(relabel_keys_kernel, dim3(grid_for(V)), dim3(256), 0, stream(),
                       g->out.degree, in_deg, V, keys, ids);
    hipError_t e = sort_order(33);
    if (e == hipSuccess && hubs > 0 && hubs < V) {
      // the temporary space of the first sort serves the second (same sizes)
      hipLaunchKernelGGL(hybrid_keys_kernel, dim3(grid_for(V)), dim3(256), 0, stream(), rank,
                         g->out.degree, in_deg, V, hubs, keys, ids);
      e = sort_order(34);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    device_free(tmp);
    if (e != hipSuccess)
      r = fail(e == hipErrorOutOfMemory ? GG_ENOMEM : GG_EHIP,
               std::string("degree order: ") + hipGetErrorString(e));
  }
  device_free(keys);
  device_free(keys_alt);
  device_free(ids);
  device_free(ids_alt);
  return r;
}

// Degree-sorted relabeling: iperm = the degree order, perm its inverse;
// degree arrays are then permuted into internal order.
static int relabel(gg_graph *g) {
  uint32_t V = g->vertices;
  int r = graph_alloc(g, (void **)&g->perm, (size_t)V * 4);
  if (r == GG_OK) r = graph_alloc(g, (void **)&g->iperm, (size_t)V * 4);
  const int64_t hubs = options().relabel_hubs;
  if (r == GG_OK)
    r = degree_order(g, g->iperm, g->perm, hubs > 0 && hubs < V ? (uint32_t)hubs : 0);
  uint32_t *scratch = nullptr;
  if (r == GG_OK) r = device_alloc((void **)&scratch, (size_t)V * 4);
  if (r == GG_OK) {
    hipLaunchKernelGGL(gather_u32_kernel, dim3(grid_for(V)), dim3(256), 0, stream(),
                       g->out.degree, g->iperm, V, scratch);
    hipError_t e =
        hipMemcpyAsync(g->out.degree, scratch, (size_t)V * 4, hipMemcpyDeviceToDevice, stream());
    if (e == hipSuccess && !g->symmetric) {
      hipLaunchKernelGGL(gather_u32_kernel, dim3(grid_for(V)), dim3(256), 0, stream(),
                         g->in.degree, g->iperm, V, scratch);
      e = hipMemcpyAsync(g->in.degree, scratch, (size_t)V * 4, hipMemcpyDeviceToDevice, stream());
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    if (e != hipSuccess) r = fail(GG_EHIP, std::string("relabel: ") + hipGetErrorString(e));
  }
  device_free(scratch);
  if (r == GG_OK) {
    g->h_perm.resize(V);
    hipError_t e = hipMemcpy(g->h_perm.data(), g->perm, (size_t)V * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = fail(GG_EHIP, std::string("relabel copy: ") + hipGetErrorString(e));
  }
  if (r == GG_OK) g->reordered = 1;
  return r;
}

static int graph_build_device(const uint32_t *d_edges, uint64_t ne, uint32_t V, int weighted,
                              int undirected, unsigned flags, gg_graph **out) {
  if (V == 0) return fail(GG_EINVAL, "graph needs at least one vertex");
  if (V >= 0xFFFFFFFFu) return fail(GG_EINVAL, "vertex count must be < 2^32-1 (sentinel V)");
  int rec_words = weighted ? 3 : 2;
  gg_graph *g = new gg_graph;
  g->vertices = V;
  g->edges = ne;
  g->weighted = weighted;
  g->symmetric = undirected;
  g->arcs = undirected ? 2 * ne : ne;
  int r = GG_OK;
  uint32_t *bad = nullptr;
  if (ne) {
    r = device_alloc((void **)&bad, 4);
    if (r == GG_OK) {
      hipMemsetAsync(bad, 0, 4, stream());
      hipLaunchKernelGGL(check_range_kernel, dim3(grid_for(ne)), dim3(256), 0, stream(), d_edges,
                         ne, rec_words, V, bad);
      uint32_t hb = 0;
      hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, stream());
      hipStreamSynchronize(stream());
      device_free(bad);
      if (hb & 1u) r = fail(GG_EINVAL, "edge endpoint >= vertices");
      else if (hb & 2u) r = fail(GG_EINVAL, kBadWeight);
    }
  }
  if (r == GG_OK) r = graph_alloc(g, (void **)&g->out.degree, ((size_t)V + 1) * 4);
  if (r == GG_OK) r = graph_alloc(g, (void **)&g->out.index, ((size_t)V + 1) * 8);
  if (r == GG_OK && !undirected) r = graph_alloc(g, (void **)&g->in.degree, ((size_t)V + 1) * 4);
  if (r == GG_OK && !undirected) r = graph_alloc(g, (void **)&g->in.index, ((size_t)V + 1) * 8);
  if (r == GG_OK) {
    hipMemsetAsync(g->out.degree, 0, ((size_t)V + 1) * 4, stream());
    if (!undirected) hipMemsetAsync(g->in.degree, 0, ((size_t)V + 1) * 4, stream());
    // (the offsets arrays serve as the run-start scratch before their scan)
    r = degrees_by_sort(d_edges, ne, rec_words, undirected ? 2 : 0, V, g->out.degree,
                        g->out.index);
    if (r == GG_OK && !undirected)
      r = degrees_by_sort(d_edges, ne, rec_words, 1, V, g->in.degree, g->in.index);
    if (r == GG_OK && (flags & GG_BUILD_REORDER)) r = relabel(g);
    if (r == GG_OK) r = scan_degrees(g->out.degree, g->out.index, V);
    if (r == GG_OK && !undirected) r = scan_degrees(g->in.degree, g->in.index, V);
  }
  // hub-first lists on file ids: neighbours sorted by their degree rank (a
  // relabelled graph's lists are in that order already)
  uint32_t *order = nullptr, *rank = nullptr;
  if (r == GG_OK && ne && (flags & GG_BUILD_HUB_FIRST) && !(flags & GG_BUILD_REORDER)) {
    r = device_alloc((void **)&order, (size_t)V * 4);
    if (r == GG_OK) r = device_alloc((void **)&rank, (size_t)V * 4);
    if (r == GG_OK) r = degree_order(g, order, rank);
  }
  if (r == GG_OK) {
    if (ne == 0) {
      r = graph_alloc(g, (void **)&g->out.nbr, 4);
      if (r == GG_OK && !undirected) r = graph_alloc(g, (void **)&g->in.nbr, 4);
      if (r == GG_OK && weighted) r = graph_alloc(g, (void **)&g->out.w, 4);
      if (r == GG_OK && weighted && !undirected) r = graph_alloc(g, (void **)&g->in.w, 4);
    } else if (undirected) {
      r = build_dir(g, g->out, d_edges, ne, rec_words, 2, rank, order);
    } else {
      r = build_dir(g, g->out, d_edges, ne, rec_words, 0, rank, order);
      if (r == GG_OK) r = build_dir(g, g->in, d_edges, ne, rec_words, 1, rank, order);
    }
  }
  device_free(order);
  device_free(rank);
  if (r == GG_OK) r = graph_finish(g);
  if (r != GG_OK) {
    std::string msg = gg_last_error();
    gg_graph_destroy(g);
    return fail(r, msg);
  }
  *out = g;
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" {

int gg_rmat_generate(void *d_edges, uint64_t num_edges, int scale, double a, double b, double c,
                     uint64_t seed, int permute_ids, int weighted) {
  GG_TRY(ensure_init());
  if (!d_edges && num_edges) return fail(GG_EINVAL, "gg_rmat_generate: null output");
  if (scale < 1 || scale > 31) return fail(GG_EINVAL, "gg_rmat_generate: scale must be in [1,31]");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return fail(GG_EINVAL, "gg_rmat_generate: bad quadrant probabilities");
  if (!num_edges) return GG_OK;
  uint64_t key = mix64h(seed ^ 0xD1B54A32D192ED03ULL);
  uint64_t ta = (uint64_t)(a * 4294967296.0);
  uint64_t tab = (uint64_t)((a + b) * 4294967296.0);
  uint64_t tabc = (uint64_t)((a + b + c) * 4294967296.0);
  PermKeys pk = perm_keys(scale, seed);
  hipLaunchKernelGGL(rmat_kernel, dim3(grid_for(num_edges, 256, 1u << 16)), dim3(256), 0,
                     stream(), (uint32_t *)d_edges, num_edges, scale, key, ta, tab, tabc,
                     permute_ids, pk, weighted);
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

int gg_graph_build_ex(const void *edges, int edges_on_device, uint64_t num_edges,
                      uint32_t vertices, int weighted, int undirected, unsigned flags,
                      gg_graph **out) {
  if (!out) return fail(GG_EINVAL, "gg_graph_build: null out");
  if (!edges && num_edges) return fail(GG_EINVAL, "gg_graph_build: null edges");
  if (flags & ~(GG_BUILD_REORDER | GG_BUILD_HUB_FIRST))
    return fail(GG_EINVAL, "gg_graph_build: unknown flags");
  // host input is checked before any device work (the device check covers
  // edge lists already on the device)
  if (!edges_on_device && weighted && !host_weights_ok((const uint32_t *)edges, num_edges))
    return fail(GG_EINVAL, kBadWeight);
  GG_TRY(ensure_init());
  size_t rec = weighted ? 12 : 8;
  if (edges_on_device || num_edges == 0)
    return graph_build_device((const uint32_t *)edges, num_edges, vertices, weighted, undirected,
                              flags, out);
  void *d = nullptr;
  GG_TRY(device_alloc(&d, num_edges * rec));
  hipError_t e = hipMemcpyAsync(d, edges, num_edges * rec, hipMemcpyHostToDevice, stream());
  if (e == hipSuccess) e = hipStreamSynchronize(stream());
  int r = e == hipSuccess ? graph_build_device((const uint32_t *)d, num_edges, vertices, weighted,
                                               undirected, flags, out)
                          : fail(GG_EHIP, std::string("upload: ") + hipGetErrorString(e));
  device_free(d);
  return r;
}

int gg_graph_build(const void *edges, int edges_on_device, uint64_t num_edges, uint32_t vertices,
                   int weighted, int undirected, gg_graph **out) {
  return gg_graph_build_ex(edges, edges_on_device, num_edges, vertices, weighted, undirected, 0u,
                           out);
}

int gg_graph_load(const char *path, uint32_t vertices, int weighted, int undirected,
                  gg_graph **out) {
  return gg_graph_load_ex(path, vertices, weighted, undirected, 0u, out);
}

int gg_graph_load_ex(const char *path, uint32_t vertices, int weighted, int undirected,
                     unsigned flags, gg_graph **out) {
  if (!path || !out) return fail(GG_EINVAL, "gg_graph_load: null argument");
  if (flags & ~(GG_BUILD_REORDER | GG_BUILD_HUB_FIRST))
    return fail(GG_EINVAL, "gg_graph_load: unknown flags");
  GG_TRY(ensure_init());
  struct stat st;
  if (stat(path, &st) != 0) return fail(GG_EIO, std::string("cannot stat ") + path);
  size_t rec = weighted ? 12 : 8;
  uint64_t ne = (uint64_t)st.st_size / rec;  // core/graph.hpp:781-782
  FILE *f = fopen(path, "rb");
  if (!f) return fail(GG_EIO, std::string("cannot open ") + path);
  void *d = nullptr;
  int r = device_alloc(&d, std::max<uint64_t>(ne * rec, 8));
  void *stage = nullptr;
  const size_t chunk = (size_t)48 << 20;  // whole 8- and 12-byte records
  if (r == GG_OK && hipHostMalloc(&stage, chunk, hipHostMallocDefault) != hipSuccess)
    r = fail(GG_ENOMEM, "pinned staging buffer");
  uint64_t done = 0, total = ne * rec;
  while (r == GG_OK && done < total) {
    size_t want = (size_t)std::min<uint64_t>(chunk, total - done);
    size_t got = fread(stage, 1, want, f);
    if (got != want) {
      r = fail(GG_EIO, std::string("short read from ") + path);
      break;
    }
    if (weighted && !host_weights_ok((const uint32_t *)stage, got / rec)) {
      r = fail(GG_EINVAL, kBadWeight);
      break;
    }
    if (hipMemcpyAsync((char *)d + done, stage, got, hipMemcpyHostToDevice, stream()) !=
            hipSuccess ||
        hipStreamSynchronize(stream()) != hipSuccess)
      r = fail(GG_EHIP, "upload of edge list failed");
    done += got;
  }
  fclose(f);
  if (stage) hipHostFree(stage);
  if (r == GG_OK)
    r = graph_build_device((const uint32_t *)d, ne, vertices, weighted, undirected, flags, out);
  device_free(d);
  return r;
}

int gg_graph_destroy(gg_graph *g) {
  if (!g) return GG_OK;
  for (Work *w : g->works)
    if (w) work_free(w);
  g->pr_plan.reset();
  for (void *p : g->allocations) device_free(p);
  delete g;
  return GG_OK;
}

int gg_graph_transpose(gg_graph *g) {
  if (!g) return fail(GG_EINVAL, "null graph");
  if (!g->symmetric) g->transposed ^= 1;
  return GG_OK;
}

int gg_graph_get_info(const gg_graph *g, gg_graph_info *info) {
  if (!g || !info) return fail(GG_EINVAL, "null argument");
  const Dir &o = out_dir(g), &i = in_dir(g);
  info->vertices = g->vertices;
  info->edges = g->edges;
  info->arcs = g->arcs;
  info->weighted = g->weighted;
  info->symmetric = g->symmetric;
  info->transposed = g->transposed;
  info->device_bytes = g->device_bytes;
  info->d_out_degree = o.degree;
  info->d_in_degree = i.degree;
  info->d_out_index = o.index;
  info->d_out_nbr = o.nbr;
  info->d_out_w = o.w;
  info->d_in_index = i.index;
  info->d_in_nbr = i.nbr;
  info->d_in_w = i.w;
  info->reordered = g->reordered;
  info->d_perm = g->perm;
  info->d_iperm = g->iperm;
  return GG_OK;
}

}  // extern "C"

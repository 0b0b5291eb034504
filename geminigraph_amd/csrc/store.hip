// Vertex arrays to and from files, and a cached binary CSR/CSC.
//
// dump_vertex_array / restore_vertex_array (core/graph.hpp:253-307) write or
// read each partition's slice of a vertex array at byte offset
// sizeof(T) * partition_offset[p] of a file of sizeof(T) * vertices bytes;
// gather_vertex_array (:310-324) collects the slices on one rank.  With one
// partition the slice is the whole array: dump writes all of it, restore
// reads all of it (a file of another size is an error, where the reference
// asserts), gather is a device-to-host copy.  Arrays are indexed by
// original vertex ids, as every entry point returns them.
//
// The reference rebuilds its CSR/CSC from the edge list on every launch.
// gg_graph_save writes a built graph (both directions, the relabeling) and
// gg_graph_open loads it without the sort; the derived structures (heavy
// chunks, per-vertex records) are rebuilt on load.
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>

#include "kernels.hpp"

namespace gg {
namespace {

constexpr size_t kStage = (size_t)64 << 20;  // pinned staging buffer

struct Staging {
  void *p = nullptr;
  int get() {
    if (!p && hipHostMalloc(&p, kStage, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return fail(GG_ENOMEM, "pinned staging buffer");
    }
    return GG_OK;
  }
  ~Staging() {
    if (p) hipHostFree(p);
  }
};

int write_device(FILE *f, const void *d, size_t bytes, Staging &s, const char *path) {
  GG_TRY(s.get());
  for (size_t done = 0; done < bytes;) {
    const size_t n = std::min(kStage, bytes - done);
    GG_HIP(hipMemcpyAsync(s.p, (const char *)d + done, n, hipMemcpyDeviceToHost, stream()));
    GG_HIP(hipStreamSynchronize(stream()));
    if (fwrite(s.p, 1, n, f) != n) return fail(GG_EIO, std::string("short write to ") + path);
    done += n;
  }
  return GG_OK;
}

int read_device(FILE *f, void *d, size_t bytes, Staging &s, const char *path) {
  GG_TRY(s.get());
  for (size_t done = 0; done < bytes;) {
    const size_t n = std::min(kStage, bytes - done);
    if (fread(s.p, 1, n, f) != n) return fail(GG_EIO, std::string("short read from ") + path);
    GG_HIP(hipMemcpyAsync((char *)d + done, s.p, n, hipMemcpyHostToDevice, stream()));
    GG_HIP(hipStreamSynchronize(stream()));
    done += n;
  }
  return GG_OK;
}

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

// Graph file: this header, then per direction (out; in unless symmetric)
// degree[V] u32, index[V+1] u64, nbr[arcs] u32, w[arcs] f32 (weighted only),
// then perm[V] and iperm[V] u32 (relabelled only).
struct GraphHeader {
  char magic[8];
  uint32_t version;
  uint32_t vertices;
  uint64_t edges, arcs;
  uint32_t weighted, symmetric, reordered, transposed;
};
constexpr char kMagic[8] = {'G', 'G', 'C', 'S', 'R', 'C', 'S', 'C'};
constexpr uint32_t kVersion = 1;

uint64_t graph_file_bytes(const GraphHeader &h) {
  const uint64_t V = h.vertices;
  const uint64_t dir = V * 4 + (V + 1) * 8 + h.arcs * 4 + (h.weighted ? h.arcs * 4 : 0);
  return sizeof(GraphHeader) + dir * (h.symmetric ? 1 : 2) + (h.reordered ? V * 8 : 0);
}

int write_dir(FILE *f, const Dir &d, const GraphHeader &h, Staging &s, const char *path) {
  GG_TRY(write_device(f, d.degree, (size_t)h.vertices * 4, s, path));
  GG_TRY(write_device(f, d.index, ((size_t)h.vertices + 1) * 8, s, path));
  GG_TRY(write_device(f, d.nbr, (size_t)h.arcs * 4, s, path));
  if (h.weighted) GG_TRY(write_device(f, d.w, (size_t)h.arcs * 4, s, path));
  return GG_OK;
}

int read_dir(FILE *f, gg_graph *g, Dir &d, const GraphHeader &h, Staging &s, const char *path) {
  const size_t V = h.vertices;
  // degree keeps the slot past the last vertex that the build allocates
  GG_TRY(graph_alloc(g, (void **)&d.degree, (V + 1) * 4));
  GG_HIP(hipMemsetAsync(d.degree + V, 0, 4, stream()));
  GG_TRY(graph_alloc(g, (void **)&d.index, (V + 1) * 8));
  GG_TRY(graph_alloc(g, (void **)&d.nbr, std::max<size_t>(h.arcs, 1) * 4));
  GG_TRY(read_device(f, d.degree, V * 4, s, path));
  GG_TRY(read_device(f, d.index, (V + 1) * 8, s, path));
  GG_TRY(read_device(f, d.nbr, (size_t)h.arcs * 4, s, path));
  if (h.weighted) {
    GG_TRY(graph_alloc(g, (void **)&d.w, std::max<size_t>(h.arcs, 1) * 4));
    GG_TRY(read_device(f, d.w, (size_t)h.arcs * 4, s, path));
  }
  return GG_OK;
}

// A saved graph is checked before any kernel walks it: offsets start at 0,
// never decrease and end at the arc count, degrees match them, neighbours
// are vertices, and the relabeling is a permutation with its inverse.
__global__ void check_dir_kernel(const uint64_t *index, const uint32_t *nbr, const uint32_t *deg,
                                 uint32_t V, uint64_t arcs, uint32_t *bad) {
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t v = i0; v < V; v += s) {
    const uint64_t b = index[v], e = index[v + 1];
    if (e < b || e > arcs || deg[v] != e - b || (v == 0 && b != 0)) atomicOr(bad, 1u);
  }
  if (i0 == 0 && index[V] != arcs) atomicOr(bad, 1u);
  for (uint64_t a = i0; a < arcs; a += s)
    if (nbr[a] >= V) atomicOr(bad, 2u);
}
__global__ void check_weights_kernel(const float *w, uint64_t arcs, uint32_t *bad) {
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; a < arcs; a += s)
    if (!weight_ok(__float_as_uint(w[a]))) atomicOr(bad, 8u);
}
__global__ void check_perm_kernel(const uint32_t *perm, const uint32_t *iperm, uint32_t V,
                                  uint32_t *bad) {
  const uint32_t s = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += s) {
    const uint32_t p = perm[i];
    if (p >= V || iperm[p] != i) atomicOr(bad, 4u);
  }
}

int check_graph_arrays(const gg_graph *g) {
  uint32_t *bad = nullptr;
  GG_TRY(device_alloc((void **)&bad, 4));
  hipMemsetAsync(bad, 0, 4, stream());
  const unsigned grid = grid1((uint64_t)g->vertices + g->arcs, 8192);
  hipLaunchKernelGGL(check_dir_kernel, dim3(grid), dim3(kBlock), 0, stream(), g->out.index,
                     g->out.nbr, g->out.degree, g->vertices, g->arcs, bad);
  if (!g->symmetric)
    hipLaunchKernelGGL(check_dir_kernel, dim3(grid), dim3(kBlock), 0, stream(), g->in.index,
                       g->in.nbr, g->in.degree, g->vertices, g->arcs, bad);
  if (g->weighted) {
    hipLaunchKernelGGL(check_weights_kernel, dim3(grid), dim3(kBlock), 0, stream(), g->out.w,
                       g->arcs, bad);
    if (!g->symmetric)
      hipLaunchKernelGGL(check_weights_kernel, dim3(grid), dim3(kBlock), 0, stream(), g->in.w,
                         g->arcs, bad);
  }
  if (g->reordered)
    hipLaunchKernelGGL(check_perm_kernel, dim3(grid1(g->vertices, 8192)), dim3(kBlock), 0,
                       stream(), g->perm, g->iperm, g->vertices, bad);
  uint32_t h = 0;
  hipError_t e = hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, stream());
  if (e == hipSuccess) e = hipStreamSynchronize(stream());
  device_free(bad);
  if (e != hipSuccess) return fail(GG_EHIP, hipGetErrorString(e));
  if (h & 8u) return fail(GG_EIO, std::string("saved graph: ") + kBadWeight);
  if (h) return fail(GG_EIO, "saved graph has inconsistent arrays");
  return GG_OK;
}

int open_graph(const char *path, gg_graph *g) {
  File fh;
  fh.f = fopen(path, "rb");
  if (!fh.f) return fail(GG_EIO, std::string("cannot open ") + path);
  GraphHeader h;
  if (fread(&h, sizeof(h), 1, fh.f) != 1 || memcmp(h.magic, kMagic, 8) != 0)
    return fail(GG_EIO, std::string(path) + " is not a saved graph");
  if (h.version != kVersion) return fail(GG_EIO, std::string(path) + ": unknown graph file version");
  if (h.vertices == 0 || h.vertices >= 0xFFFFFFFFu || h.weighted > 1 || h.symmetric > 1 ||
      h.reordered > 1 || h.transposed > 1 || h.arcs != (h.symmetric ? 2 * h.edges : h.edges))
    return fail(GG_EIO, std::string(path) + ": bad graph header");
  struct stat st;
  if (stat(path, &st) != 0 || (uint64_t)st.st_size != graph_file_bytes(h))
    return fail(GG_EIO, std::string(path) + ": file size does not match its header");
  g->vertices = h.vertices;
  g->edges = h.edges;
  g->arcs = h.arcs;
  g->weighted = (int)h.weighted;
  g->symmetric = (int)h.symmetric;
  g->transposed = (int)h.transposed;
  Staging s;
  GG_TRY(read_dir(fh.f, g, g->out, h, s, path));
  if (!h.symmetric) GG_TRY(read_dir(fh.f, g, g->in, h, s, path));
  if (h.reordered) {
    const size_t V = h.vertices;
    GG_TRY(graph_alloc(g, (void **)&g->perm, V * 4));
    GG_TRY(graph_alloc(g, (void **)&g->iperm, V * 4));
    GG_TRY(read_device(fh.f, g->perm, V * 4, s, path));
    GG_TRY(read_device(fh.f, g->iperm, V * 4, s, path));
    g->reordered = 1;
  }
  GG_TRY(check_graph_arrays(g));
  if (g->reordered) {
    g->h_perm.resize(h.vertices);
    GG_HIP(hipMemcpy(g->h_perm.data(), g->perm, (size_t)h.vertices * 4, hipMemcpyDeviceToHost));
  }
  return graph_finish(g);
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" {

int gg_dump_vertex_array(const gg_graph *g, const void *d_array, size_t elem_bytes,
                         const char *path) {
  GG_TRY(check_graph(g));
  if (!d_array || !path || !elem_bytes) return fail(GG_EINVAL, "gg_dump_vertex_array: null argument");
  File fh;
  fh.f = fopen(path, "wb");
  if (!fh.f) return fail(GG_EIO, std::string("cannot create ") + path);
  Staging s;
  GG_TRY(write_device(fh.f, d_array, elem_bytes * g->vertices, s, path));
  if (fclose(fh.f) != 0) {
    fh.f = nullptr;
    return fail(GG_EIO, std::string("cannot close ") + path);
  }
  fh.f = nullptr;
  return GG_OK;
}

int gg_restore_vertex_array(const gg_graph *g, void *d_array, size_t elem_bytes,
                            const char *path) {
  GG_TRY(check_graph(g));
  if (!d_array || !path || !elem_bytes)
    return fail(GG_EINVAL, "gg_restore_vertex_array: null argument");
  struct stat st;
  if (stat(path, &st) != 0) return fail(GG_EIO, std::string("cannot stat ") + path);
  if ((uint64_t)st.st_size != (uint64_t)elem_bytes * g->vertices)
    return fail(GG_EIO, std::string(path) + ": size is not elem_bytes x vertices");
  File fh;
  fh.f = fopen(path, "rb");
  if (!fh.f) return fail(GG_EIO, std::string("cannot open ") + path);
  Staging s;
  return read_device(fh.f, d_array, elem_bytes * g->vertices, s, path);
}

int gg_gather_vertex_array(const gg_graph *g, const void *d_array, size_t elem_bytes,
                           void *h_out) {
  GG_TRY(check_graph(g));
  if (!d_array || !h_out || !elem_bytes)
    return fail(GG_EINVAL, "gg_gather_vertex_array: null argument");
  GG_HIP(hipMemcpyAsync(h_out, d_array, elem_bytes * g->vertices, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

int gg_graph_save(const gg_graph *g, const char *path) {
  GG_TRY(check_graph(g));
  if (!path) return fail(GG_EINVAL, "gg_graph_save: null path");
  GraphHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.vertices = g->vertices;
  h.edges = g->edges;
  h.arcs = g->arcs;
  h.weighted = (uint32_t)g->weighted;
  h.symmetric = (uint32_t)g->symmetric;
  h.reordered = (uint32_t)g->reordered;
  h.transposed = (uint32_t)g->transposed;
  File fh;
  fh.f = fopen(path, "wb");
  if (!fh.f) return fail(GG_EIO, std::string("cannot create ") + path);
  if (fwrite(&h, sizeof(h), 1, fh.f) != 1) return fail(GG_EIO, std::string("short write to ") + path);
  Staging s;
  GG_TRY(write_dir(fh.f, g->out, h, s, path));
  if (!g->symmetric) GG_TRY(write_dir(fh.f, g->in, h, s, path));
  if (g->reordered) {
    GG_TRY(write_device(fh.f, g->perm, (size_t)g->vertices * 4, s, path));
    GG_TRY(write_device(fh.f, g->iperm, (size_t)g->vertices * 4, s, path));
  }
  const int c = fclose(fh.f);
  fh.f = nullptr;
  if (c != 0) return fail(GG_EIO, std::string("cannot close ") + path);
  return GG_OK;
}

int gg_graph_open(const char *path, gg_graph **out) {
  if (!path || !out) return fail(GG_EINVAL, "gg_graph_open: null argument");
  *out = nullptr;
  GG_TRY(ensure_init());
  gg_graph *g = new gg_graph;
  const int r = open_graph(path, g);
  if (r != GG_OK) {
    std::string msg = gg_last_error();
    gg_graph_destroy(g);
    return fail(r, msg);
  }
  *out = g;
  return GG_OK;
}

}  // extern "C"

// Runtime pieces of libgg: errors, options, device selection, memory,
// profiling and vertex subsets (device bitmaps).
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "gg_internal.hpp"

namespace gg {

static thread_local std::string t_error;
static thread_local int t_code = 0;

void set_error(int code, const std::string &msg) {
  t_code = code;
  t_error = msg;
}
int fail(int code, const std::string &msg) {
  set_error(code, msg);
  return code;
}

Options &options() {
  static Options o;
  return o;
}

static int g_device = -1;
static hipStream_t g_stream = nullptr;
static std::mutex g_init_mu;

int init_device(int device) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return fail(GG_ENODEV, "no HIP device available");
  if (device < 0 || device >= n) return fail(GG_EINVAL, "device index out of range");
  if (g_device == device && g_stream) return GG_OK;
  GG_HIP(hipSetDevice(device));
  if (g_stream) {
    hipStreamDestroy(g_stream);
    g_stream = nullptr;
  }
  GG_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  // the allocation-free entry point's scratch array, on this device
  return bfs_noalloc_setup(device);
}

int ensure_init() {
  if (g_device >= 0 && g_stream) {
    GG_HIP(hipSetDevice(g_device));
    return GG_OK;
  }
  return init_device(0);
}

static thread_local hipStream_t t_stream = nullptr;
static thread_local int t_slot = 0;

hipStream_t stream() { return t_stream ? t_stream : g_stream; }
int context_slot() { return t_slot; }
void bind_thread(hipStream_t s, int slot) {
  t_stream = s;
  t_slot = slot;
}

int device_alloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(GG_ENOMEM, "hipMalloc(" + std::to_string(bytes) + ") failed: " +
                               hipGetErrorString(e));
  }
  return GG_OK;
}
void device_free(void *p) {
  if (p) (void)hipFree(p);
}

// ---------------------------------------------------------------- profiling
struct ProfEntry {
  double total_ms = 0;
  uint64_t launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};
static bool g_prof_on = false;
static std::map<std::string, ProfEntry> g_prof;
static std::mutex g_prof_mu;
bool prof_on() { return g_prof_on; }

ProfScope::ProfScope(const char *n) : name(n) {
  if (!g_prof_on || !n) return;  // a null name: never timed (no events created)
  if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
    start = stop = nullptr;
    return;
  }
  hipEventRecord(start, stream());
}
ProfScope::~ProfScope() {
  if (!start) return;
  hipEventRecord(stop, stream());
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof[name].pending.emplace_back(start, stop);
}

static void prof_collect() {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (auto &kv : g_prof) {
    for (auto &pr : kv.second.pending) {
      hipEventSynchronize(pr.second);
      float ms = 0;
      hipEventElapsedTime(&ms, pr.first, pr.second);
      kv.second.total_ms += ms;
      kv.second.launches++;
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
    kv.second.pending.clear();
  }
}

}  // namespace gg

using namespace gg;

// ------------------------------------------------------------------ C ABI
extern "C" {

const char *gg_last_error(void) { return t_error.c_str(); }

int gg_init(int device) { return init_device(device); }

int gg_synchronize(void) {
  GG_TRY(ensure_init());
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

// Buffers of gg_malloc / gg_free (the callers' vertex arrays and subsets)
// are cached: a freed buffer waits, per device and exact size, for the next
// gg_malloc of that size, up to alloc_cache_bytes in all; a toolkit that
// allocates its outputs in every call (as compute() does) then pays no
// hipMalloc / hipFree after the first.  gg_free still waits for the work
// queued on the device, as hipFree does, so a buffer handed out again is
// idle.  When hipMalloc fails, the cache is released and the call retried.
namespace {
struct AllocCache {
  std::mutex m;
  std::multimap<std::pair<int, size_t>, void *> idle;  // (device, bytes) -> buffer
  std::map<void *, std::pair<int, size_t>> owned;      // every live or idle buffer
  size_t idle_bytes = 0;
};
AllocCache &alloc_cache() {
  static AllocCache c;
  return c;
}
void release_idle(AllocCache &c) {  // with c.m held
  for (auto &kv : c.idle) {
    c.owned.erase(kv.second);
    (void)hipFree(kv.second);
  }
  c.idle.clear();
  c.idle_bytes = 0;
}
}  // namespace

int gg_malloc(void **d_ptr, size_t bytes) {
  if (!d_ptr) return fail(GG_EINVAL, "gg_malloc: null out pointer");
  GG_TRY(ensure_init());
  if (bytes == 0) bytes = 256;
  int dev = 0;
  GG_HIP(hipGetDevice(&dev));
  AllocCache &c = alloc_cache();
  std::lock_guard<std::mutex> lk(c.m);
  auto it = c.idle.find({dev, bytes});
  if (it != c.idle.end()) {
    *d_ptr = it->second;
    c.idle.erase(it);
    c.idle_bytes -= bytes;
    return GG_OK;
  }
  int r = device_alloc(d_ptr, bytes);
  if (r != GG_OK && c.idle_bytes) {
    release_idle(c);
    r = device_alloc(d_ptr, bytes);
  }
  if (r == GG_OK) c.owned[*d_ptr] = {dev, bytes};
  return r;
}

int gg_free(void *d_ptr) {
  GG_TRY(ensure_init());
  if (!d_ptr) return GG_OK;
  AllocCache &c = alloc_cache();
  std::lock_guard<std::mutex> lk(c.m);
  auto it = c.owned.find(d_ptr);
  if (it == c.owned.end()) {  // not from gg_malloc
    GG_HIP(hipFree(d_ptr));
    return GG_OK;
  }
  const std::pair<int, size_t> key = it->second;
  const size_t cap = (size_t)std::max<int64_t>(0, options().alloc_cache_bytes);
  if (c.idle_bytes + key.second > cap) {
    c.owned.erase(it);
    GG_HIP(hipFree(d_ptr));
    return GG_OK;
  }
  // the queued work that may use it has finished, on the buffer's device
  int cur = key.first;
  GG_HIP(hipGetDevice(&cur));
  if (cur != key.first) GG_HIP(hipSetDevice(key.first));
  const hipError_t e = hipDeviceSynchronize();
  if (cur != key.first) GG_HIP(hipSetDevice(cur));
  GG_HIP(e);
  c.idle.insert({key, d_ptr});
  c.idle_bytes += key.second;
  return GG_OK;
}

// A lower alloc_cache_bytes takes effect at once: kept buffers are released.
void trim_alloc_cache() {
  AllocCache &c = alloc_cache();
  std::lock_guard<std::mutex> lk(c.m);
  if (c.idle_bytes > (size_t)std::max<int64_t>(0, options().alloc_cache_bytes)) release_idle(c);
}

int gg_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes) {
  GG_TRY(ensure_init());
  if (!bytes) return GG_OK;
  if (!h_dst || !d_src) return fail(GG_EINVAL, "gg_memcpy_d2h: null pointer");
  GG_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

int gg_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes) {
  GG_TRY(ensure_init());
  if (!bytes) return GG_OK;
  if (!h_src || !d_dst) return fail(GG_EINVAL, "gg_memcpy_h2d: null pointer");
  GG_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

int gg_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes) {
  GG_TRY(ensure_init());
  if (!bytes) return GG_OK;
  if (!d_src || !d_dst) return fail(GG_EINVAL, "gg_memcpy_d2d: null pointer");
  GG_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, stream()));
  return GG_OK;
}

int gg_mem_info(size_t *free_bytes, size_t *total_bytes) {
  GG_TRY(ensure_init());
  size_t f = 0, t = 0;
  GG_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return GG_OK;
}

static int64_t *option_slot(const char *name) {
  Options &o = options();
  if (!name) return nullptr;
  if (!strcmp(name, "switch_divisor")) return &o.switch_divisor;
  if (!strcmp(name, "plan_switch_divisor")) return &o.plan_switch_divisor;
  if (!strcmp(name, "force_mode")) return &o.force_mode;
  if (!strcmp(name, "push_small_max")) return &o.push_small_max;
  if (!strcmp(name, "push_big_max")) return &o.push_big_max;
  if (!strcmp(name, "pull_light_max")) return &o.pull_light_max;
  if (!strcmp(name, "pull_serial_max")) return &o.pull_serial_max;
  if (!strcmp(name, "huge_grid_waves")) return &o.huge_grid_waves;
  if (!strcmp(name, "count_examined")) return &o.count_examined;
  if (!strcmp(name, "level_masks")) return &o.level_masks;
  if (!strcmp(name, "level_mask_bytes")) return &o.level_mask_bytes;
  if (!strcmp(name, "pr_hot_sources")) return &o.pr_hot_sources;
  if (!strcmp(name, "pr_bin_bits")) return &o.pr_bin_bits;
  if (!strcmp(name, "bfs_alpha")) return &o.bfs_alpha;
  if (!strcmp(name, "pr_range_bits")) return &o.pr_range_bits;
  if (!strcmp(name, "heavy_waves")) return &o.heavy_waves;
  if (!strcmp(name, "bfs_pull_waves")) return &o.bfs_pull_waves;
  if (!strcmp(name, "batch_graphs")) return &o.batch_graphs;
  if (!strcmp(name, "batch_device_loop")) return &o.batch_device_loop;
  if (!strcmp(name, "bfs_ahead")) return &o.bfs_ahead;
  if (!strcmp(name, "bfs_record_starts")) return &o.bfs_record_starts;
  if (!strcmp(name, "alloc_cache_bytes")) return &o.alloc_cache_bytes;
  if (!strcmp(name, "relabel_hubs")) return &o.relabel_hubs;
  if (!strcmp(name, "serve_bfs_workers")) return &o.serve_bfs_workers;
  return nullptr;
}

int gg_set_option(const char *name, int64_t value) {
  int64_t *s = option_slot(name);
  if (!s) return fail(GG_EINVAL, std::string("unknown option: ") + (name ? name : "(null)"));
  if (!strcmp(name, "force_mode") && (value < 0 || value > 2))
    return fail(GG_EINVAL, "force_mode must be 0, 1 or 2");
  const bool flag = !strcmp(name, "count_examined") || !strcmp(name, "level_masks") ||
                    !strcmp(name, "batch_graphs") || !strcmp(name, "batch_device_loop") ||
                    !strcmp(name, "bfs_record_starts");
  if (flag && (value < 0 || value > 1)) return fail(GG_EINVAL, std::string(name) + " must be 0 or 1");
  if (!strcmp(name, "bfs_pull_waves") && (value < 0 || value > (1 << 24)))
    return fail(GG_EINVAL, "bfs_pull_waves must be in [0, 2^24]");
  if (!strcmp(name, "bfs_alpha") && value > (1 << 20))
    return fail(GG_EINVAL, "bfs_alpha must be in [0, 2^20]");
  const bool zero_ok = !strcmp(name, "bfs_pull_waves") || !strcmp(name, "pr_hot_sources") ||
                       !strcmp(name, "bfs_alpha") || !strcmp(name, "relabel_hubs") ||
                       !strcmp(name, "serve_bfs_workers") || !strcmp(name, "alloc_cache_bytes") ||
                       !strcmp(name, "heavy_waves");
  if (zero_ok && value < 0) return fail(GG_EINVAL, std::string(name) + " must be >= 0");
  if (strcmp(name, "force_mode") && !zero_ok && !flag && value <= 0)
    return fail(GG_EINVAL, std::string(name) + " must be positive");
  if (!strcmp(name, "push_small_max") && value > 64 * 64)
    return fail(GG_EINVAL, "push_small_max must be <= 4096");
  if ((!strcmp(name, "pr_bin_bits") || !strcmp(name, "pr_range_bits")) &&
      (value < 8 || value > 14))
    return fail(GG_EINVAL, std::string(name) + " must be in [8, 14]");
  if (!strcmp(name, "bfs_ahead") && value > 8) return fail(GG_EINVAL, "bfs_ahead must be in [1, 8]");
  if (!strcmp(name, "serve_bfs_workers") && value > 31)
    return fail(GG_EINVAL, "serve_bfs_workers must be in [0, 31]");
  if (!strcmp(name, "relabel_hubs") && value > 0xFFFFFFFFll)
    return fail(GG_EINVAL, "relabel_hubs must be < 2^32");
  if (!strcmp(name, "pull_light_max") && value > (1 << 20))
    return fail(GG_EINVAL, "pull_light_max must be <= 2^20");
  *s = value;
  if (!strcmp(name, "alloc_cache_bytes")) trim_alloc_cache();
  return GG_OK;
}

int gg_get_option(const char *name, int64_t *value) {
  int64_t *s = option_slot(name);
  if (!s || !value) return fail(GG_EINVAL, "unknown option or null value");
  *value = *s;
  return GG_OK;
}

int gg_prof_enable(int on) {
  GG_TRY(ensure_init());
  if (!on) prof_collect();
  g_prof_on = on != 0;
  return GG_OK;
}

int gg_prof_reset(void) {
  prof_collect();
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.clear();
  return GG_OK;
}

int gg_prof_read(const char *kernel, double *total_ms, uint64_t *launches) {
  if (!kernel) return fail(GG_EINVAL, "gg_prof_read: null kernel name");
  prof_collect();
  std::lock_guard<std::mutex> lk(g_prof_mu);
  auto it = g_prof.find(kernel);
  if (total_ms) *total_ms = it == g_prof.end() ? 0.0 : it->second.total_ms;
  if (launches) *launches = it == g_prof.end() ? 0 : it->second.launches;
  return GG_OK;
}

// ---------------------------------------------------------- vertex subsets

int gg_subset_create(const gg_graph *g, gg_subset **out) {
  if (!g || !out) return fail(GG_EINVAL, "gg_subset_create: null argument");
  GG_TRY(ensure_init());
  gg_subset *s = new gg_subset;
  s->g = g;
  s->size = g->vertices;
  s->words = ((size_t)g->vertices >> 6) + 1;
  int r = device_alloc((void **)&s->d, s->words * 8);
  if (r != GG_OK) {
    delete s;
    return r;
  }
  GG_HIP(hipMemsetAsync(s->d, 0, s->words * 8, stream()));
  *out = s;
  return GG_OK;
}

int gg_subset_destroy(gg_subset *s) {
  if (!s) return GG_OK;
  device_free(s->d);
  delete s;
  return GG_OK;
}

int gg_subset_clear(gg_subset *s) {
  if (!s) return fail(GG_EINVAL, "null subset");
  GG_HIP(hipMemsetAsync(s->d, 0, s->words * 8, stream()));
  return GG_OK;
}

int gg_subset_fill(gg_subset *s) {
  // core/bitmap.hpp:42-52: full words, then the exact tail of the last word.
  if (!s) return fail(GG_EINVAL, "null subset");
  size_t full = (size_t)s->size >> 6;
  if (full) GG_HIP(hipMemsetAsync(s->d, 0xFF, full * 8, stream()));
  uint64_t tail = 0;
  for (size_t i = full << 6; i < s->size; i++) tail |= 1ULL << (i & 63);
  GG_HIP(hipMemcpyAsync(s->d + full, &tail, 8, hipMemcpyHostToDevice, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

int gg_subset_set_bit(gg_subset *s, uint32_t v) {
  if (!s || v >= s->size) return fail(GG_EINVAL, "gg_subset_set_bit: vertex out of range");
  v = s->g->to_internal(v);
  uint64_t w = 0;
  GG_HIP(hipMemcpyAsync(&w, s->d + (v >> 6), 8, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  w |= 1ULL << (v & 63);
  GG_HIP(hipMemcpyAsync(s->d + (v >> 6), &w, 8, hipMemcpyHostToDevice, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  return GG_OK;
}

int gg_subset_get_bit(const gg_subset *s, uint32_t v, int *bit) {
  if (!s || !bit || v >= s->size) return fail(GG_EINVAL, "gg_subset_get_bit: bad argument");
  v = s->g->to_internal(v);
  uint64_t w = 0;
  GG_HIP(hipMemcpyAsync(&w, s->d + (v >> 6), 8, hipMemcpyDeviceToHost, stream()));
  GG_HIP(hipStreamSynchronize(stream()));
  *bit = (int)((w >> (v & 63)) & 1ULL);
  return GG_OK;
}

uint64_t *gg_subset_device_words(gg_subset *s) { return s ? s->d : nullptr; }

}  // extern "C"

// Threshold autotuner: times candidate values of the engine's options on a
// sample of the graph's own work (BFS from sample roots, or PageRank
// iterations) and keeps the fastest, one option at a time (coordinate
// descent).  The reference has one fixed switch (core/graph.hpp:1509,
// sparse iff active edges < |E| / 20); here the switch, the degree bins, the
// grid sizes and the PageRank plan sizes are all options, and no result
// depends on them, so tuning only moves time.
#include <string.h>

#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace gg {
namespace {

struct Knob {
  const char *name;
  std::vector<int64_t> values;
};

// Sample roots: original ids spread over [0, V) by a fixed multiplicative
// walk, keeping those with out-arcs (a root without out-arcs times nothing).
std::vector<uint32_t> sample_roots(const gg_graph *g, int n) {
  const uint32_t V = g->vertices;
  std::vector<uint32_t> deg(V);
  if (hipMemcpyAsync(deg.data(), out_dir(g).degree, (size_t)V * 4, hipMemcpyDeviceToHost,
                     stream()) != hipSuccess ||
      hipStreamSynchronize(stream()) != hipSuccess)
    return {};
  std::vector<uint32_t> roots;
  uint64_t x = 0x9E3779B97F4A7C15ULL;
  for (uint64_t tries = 0; (int)roots.size() < n && tries < 64ULL * V + 1024; tries++) {
    x = x * 6364136223846793005ULL + 1442695040888963407ULL;
    const uint32_t o = (uint32_t)((x >> 33) % V);
    if (deg[g->to_internal(o)] > 0) roots.push_back(o);
  }
  return roots;
}

struct Sample {
  gg_graph *g;
  int kind;
  std::vector<uint32_t> roots;
  int iterations = 0;
  void *a = nullptr;  // parents (BFS) or ranks (PageRank)
  int run(double *ms) {
    DeviceTimer tm;
    gg_run_stats st;
    if (kind == GG_TUNE_BFS) {
      for (uint32_t r : roots)
        GG_TRY(gg_bfs(g, r, (uint32_t *)a, nullptr, nullptr, 0, &st));
    } else {
      GG_TRY(gg_pagerank(g, iterations, (double *)a, nullptr, &st));
    }
    *ms = tm.stop();
    return GG_OK;
  }
  // best of two runs; a first untimed run builds what the settings need
  // (the PageRank plan) and warms the caches
  int time(double *ms) {
    double t0 = 0, t1 = 0;
    if (kind == GG_TUNE_PAGERANK) {
      gg_run_stats st;
      GG_TRY(gg_pagerank(g, 1, (double *)a, nullptr, &st));
    }
    GG_TRY(run(&t0));
    GG_TRY(run(&t1));
    *ms = std::min(t0, t1);
    return GG_OK;
  }
};

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_tune(gg_graph *g, int kind, int samples, int apply, gg_tune_result *out) {
  GG_TRY(check_graph(g));
  if (!out) return fail(GG_EINVAL, "gg_tune: null result");
  if (kind != GG_TUNE_BFS && kind != GG_TUNE_PAGERANK) return fail(GG_EINVAL, "gg_tune: unknown kind");
  if (samples < 1 || samples > 64) return fail(GG_EINVAL, "gg_tune: samples must be in [1, 64]");
  if (!g->vertices) return fail(GG_EINVAL, "gg_tune: empty graph");
  std::vector<Knob> knobs;
  if (kind == GG_TUNE_BFS) {
    knobs = {{"switch_divisor", {20, 50, 100, 200, 400, 1000}},
             {"bfs_alpha", {0, 8, 16, 24, 32, 64}},
             {"push_big_max", {256, 1024, 4096}},
             {"push_small_max", {16, 32, 64}},
             {"bfs_pull_waves", {4096, 8192, 16384}}};
  } else {
    knobs = {{"pr_hot_sources", {0, 1 << 17, 1 << 18, 1 << 19}},
             {"pr_bin_bits", {12, 13, 14}},
             {"pr_range_bits", {12, 13, 14}}};
  }
  if (knobs.size() > (size_t)GG_TUNE_MAX_OPTIONS) return fail(GG_EINVAL, "gg_tune: too many knobs");
  Sample s;
  s.g = g;
  s.kind = kind;
  const uint32_t V = g->vertices;
  if (kind == GG_TUNE_BFS) {
    s.roots = sample_roots(g, samples);
    if (s.roots.empty()) return fail(GG_EINVAL, "gg_tune: no vertex has out-arcs");
    GG_TRY(device_alloc(&s.a, (size_t)V * 4));  // parents only, the reference's output
  } else {
    s.iterations = samples;
    GG_TRY(device_alloc(&s.a, (size_t)V * 8));
  }
  std::vector<int64_t> before(knobs.size());
  for (size_t k = 0; k < knobs.size(); k++) gg_get_option(knobs[k].name, &before[k]);
  memset(out, 0, sizeof(*out));
  out->kind = kind;
  int rc = GG_OK;
  double best = 0;
  rc = s.time(&best);
  out->baseline_ms = best;
  out->trials = 1;
  std::vector<int64_t> chosen = before;
  for (size_t k = 0; k < knobs.size() && rc == GG_OK; k++) {
    for (int64_t v : knobs[k].values) {
      if (v == chosen[k]) continue;
      if ((rc = gg_set_option(knobs[k].name, v)) != GG_OK) break;
      double t = 0;
      if ((rc = s.time(&t)) != GG_OK) break;
      out->trials++;
      if (t < best) {
        best = t;
        chosen[k] = v;
      }
    }
    if (rc == GG_OK) rc = gg_set_option(knobs[k].name, chosen[k]);
  }
  out->tuned_ms = best;
  out->n_options = (int)knobs.size();
  for (size_t k = 0; k < knobs.size(); k++) {
    strncpy(out->names[k], knobs[k].name, sizeof(out->names[k]) - 1);
    out->values[k] = chosen[k];
    // keep the winners, or put back what was set before
    gg_set_option(knobs[k].name, apply && rc == GG_OK ? chosen[k] : before[k]);
  }
  device_free(s.a);
  return rc;
}

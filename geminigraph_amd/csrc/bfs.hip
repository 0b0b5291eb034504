// BFS (toolkits/bfs.cpp) with its iteration loop driven from the device.
//
// Every iteration is the same fixed sequence of two launches, issued ahead
// of time on one stream:
//   step       -- every block first makes the iteration's decision from the
//                 counts of the frontier just produced (its size, out-arcs
//                 and in-arcs, summed over spread slots by the kernels that
//                 produced it): the push/pull switch
//                 (core/graph.hpp:1503-1509), or the end of the traversal.
//                 All blocks reach the same decision from the same counts;
//                 block 0 also records it (statistics, control block, the
//                 flag the host polls).  Then, in a pull: bottom-up over
//                 64-vertex groups (bfs.cpp:60-76); in a push: frontier
//                 bitmap -> three degree-binned queues, the frontier joins
//                 `visited`, the output bitmap is cleared;
//   push_heavy -- push: a persistent grid over the queues (bfs.cpp:49-58);
//                 pull: the chunks of the heavy in-lists.
// The host never waits for an iteration before issuing the next one: it
// stays one iteration ahead and polls the decision the device writes to
// coherent host memory.  Per-iteration buffers that one launch reads while
// the next fills them (count slots, examined counters, queue counters,
// Beamer's bound) alternate between two copies by iteration parity.
//
// Claims: a push edge, and each chunk of a heavy in-list, claims its vertex
// with a compare-and-swap on the vertex's parent entry, as the reference
// does (bfs.cpp:53, 71); the entry sits at its original id (relabelling
// scatters those, so the atomics do not pile onto a few cache lines).  The
// depth array is optional: without it the BFS writes what the reference's
// does (parents), 4 bytes per vertex found instead of 8.  Results are
// written in original vertex ids as vertices are found (parent[iperm[v]] =
// iperm[u]), so a relabelled graph needs no pass over V at the end.  A
// frontier joins `visited` one iteration later, in the step that reads it
// next (bin for push, the group pull for pull); until then the readers test
// visited | frontier.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <type_traits>

#include "kernels.hpp"

namespace gg {

enum : uint32_t { kBfsPush = 0, kBfsPull = 1, kBfsDone = 2 };
constexpr uint32_t kHubBatch = 512;  // huge frontier vertices staged per push round
constexpr uint32_t kPushChunk = 128;  // arcs of a huge frontier vertex per wave step (one two-step round)
static_assert(kHubBatch % kBlock == 0, "whole rows of the staged batch per thread");
constexpr int kCountStride = 16;  // u64 per count slot: one 128-byte line
// Count slots of one frontier: the step kernels' waves add their counts to
// slot (wave % kBfsSlots); every block of the next decision reads all of
// them, one per lane of its first wave.
constexpr uint32_t kBfsSlots = 64;
constexpr size_t kBfsSetWords = (size_t)kBfsSlots * kCountStride;
// examined counters of one iteration: 64 slots, one 128-byte line each
constexpr size_t kExamWords = 64 * 16;

// Device control block of one traversal (done first: the allocation-free
// variant reads it back after every round of iterations; gg_bfs polls a
// flag in pinned host memory instead and copies the whole block once, at
// the end).
struct BfsCtl {
  uint32_t done, iter;  // iter = iterations started
  uint32_t mode, level, push_iters, V;
  int32_t force;
  uint32_t lv_cap;
  // levels done before the host's first iteration: 1 when the seed kernel
  // pushed the root's out-arcs itself (host iteration n is level n + n0)
  uint32_t n0;
  uint32_t pad;
  uint64_t reached, m_reach, examined, threshold;
  // in-arcs of the vertices not visited yet (the pull's scan bound) before
  // iteration n, in m_unv[n & 1]: every block of iteration n reads it while
  // block 0 writes the next copy
  uint64_t m_unv[2];
  uint64_t alpha;        // 0, or: pull only if frontier out-arcs * alpha > m_unvisited
  uint32_t modes[GG_STATS_RECORDS];
  uint64_t active[GG_STATS_RECORDS], fresh[GG_STATS_RECORDS], exam[GG_STATS_RECORDS],
      unvisited[GG_STATS_RECORDS];
};

// Where found vertices go.  parent is also the claim word (none = V until
// claimed, as bfs.cpp:35); depth is optional (null: parents only, the
// reference's output).
struct BfsOut {
  uint32_t *parent, *depth;  // indexed by original ids
  // by internal id: {original id, out-degree << 16 | in-degree (16 bits each,
  // kDegEscape: read odeg / ideg), first in-neighbour, in-list start (low 32
  // bits)}
  const uint4 *vinfo;
  uint64_t *visited, *fout;
  uint32_t none;             // V
  const uint32_t *odeg, *ideg;  // by internal id
};

// A record's degrees.  A capped field (a degree of 2^16 - 1 or more: a few
// hubs) sends the read to the degree array, on a branch the wave takes only
// when one of its lanes holds one.
__device__ __forceinline__ uint32_t rec_odeg(uint32_t degs, const uint32_t *odeg, uint32_t v) {
  uint32_t d = degs >> 16;
  if (__builtin_expect(__ballot(d == kDegEscape) != 0, 0))
    if (d == kDegEscape) d = odeg[v];
  return d;
}
__device__ __forceinline__ uint32_t rec_ideg(uint32_t degs, const uint32_t *ideg, uint32_t v) {
  uint32_t d = degs & kDegEscape;
  if (__builtin_expect(__ballot(d == kDegEscape) != 0, 0))
    if (d == kDegEscape) d = ideg[v];
  return d;
}

// Lane 0 of a wave adds its counts (vertices found, their out-arcs and their
// in-arcs) to one of kBfsSlots spread slots.
__device__ __forceinline__ void bfs_count(uint64_t *set, uint32_t slot, uint64_t n, uint64_t e,
                                          uint64_t ie) {
  if (!n) return;
  uint64_t *p = set + (size_t)(slot % kBfsSlots) * kCountStride;
  atomicAdd((unsigned long long *)p, (unsigned long long)n);
  atomicAdd((unsigned long long *)(p + 1), (unsigned long long)e);
  atomicAdd((unsigned long long *)(p + 2), (unsigned long long)ie);
}

__device__ __forceinline__ void bfs_wave_flush(uint64_t *set, uint32_t gw, uint64_t cnt,
                                               uint64_t edg, uint64_t ind,
                                               unsigned long long *exam, uint64_t ex) {
  cnt = wave_sum_u64(cnt);
  edg = wave_sum_u64(edg);
  ind = wave_sum_u64(ind);
  ex = wave_sum_u64(ex);
  if (lane_id() == 0) {
    bfs_count(set, gw, cnt, edg, ind);
    count_examined(exam, ex);
  }
}

// Push claim: `visited` holds every vertex of the earlier levels and of the
// frontier (the bin kernel adds the frontier before the push).  An edge whose
// target is neither visited nor already in the output frontier races a
// compare-and-swap on the target's parent entry, at its original id, as
// bfs.cpp:53 does: with relabelled hubs packed into a few bitmap words,
// claiming on those words instead would pile returning atomics onto a few
// cache lines.  The winner writes the depth (if kept) and sets the frontier
// bit (an atomic whose result is not needed).
// Each lane makes two claims at once (two 64-arc steps of a wave in flight):
// the plain bitmap reads, the records and the compare-and-swaps of both are
// issued before either result is used (a dependent chain each, so two chains
// overlap their latencies).
__device__ __forceinline__ void bfs_claim2(const BfsOut &o, uint32_t level, bool ok0,
                                           uint32_t s0, uint32_t d0, bool ok1, uint32_t s1,
                                           uint32_t d1, uint64_t &cnt, uint64_t &edg,
                                           uint64_t &ind) {
  const uint64_t t0 = ok0 ? (o.visited[d0 >> 6] | o.fout[d0 >> 6]) : ~0ULL;
  const uint64_t t1 = ok1 ? (o.visited[d1 >> 6] | o.fout[d1 >> 6]) : ~0ULL;
  const bool c0 = !((t0 >> (d0 & 63)) & 1ULL), c1 = !((t1 >> (d1 & 63)) & 1ULL);
  const uint4 v0 = c0 ? o.vinfo[d0] : make_uint4(0, 0, 0, 0);
  const uint4 v1 = c1 ? o.vinfo[d1] : make_uint4(0, 0, 0, 0);
  // the degrees before any write: a capped one's load then waits with the
  // compare-and-swaps, not for this lane's stores
  const uint32_t od0 = c0 ? rec_odeg(v0.y, o.odeg, d0) : 0, id0 = c0 ? rec_ideg(v0.y, o.ideg, d0) : 0;
  const uint32_t od1 = c1 ? rec_odeg(v1.y, o.odeg, d1) : 0, id1 = c1 ? rec_ideg(v1.y, o.ideg, d1) : 0;
  const bool w0 = c0 && atomicCAS(&o.parent[v0.x], o.none, s0) == o.none;
  const bool w1 = c1 && atomicCAS(&o.parent[v1.x], o.none, s1) == o.none;
  if (w0) {
    cnt++;
    edg += od0;
    ind += id0;
    if (o.depth) o.depth[v0.x] = level;
    __hip_atomic_fetch_or((unsigned long long *)&o.fout[d0 >> 6], 1ULL << (d0 & 63),
                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (w1) {
    cnt++;
    edg += od1;
    ind += id1;
    if (o.depth) o.depth[v1.x] = level;
    __hip_atomic_fetch_or((unsigned long long *)&o.fout[d1 >> 6], 1ULL << (d1 & 63),
                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Progress the host polls, in coherent pinned memory: the deciding thread
// publishes each decision with a relaxed system-scope store, which the host
// uses only to pace its launches (no release, an L2 write-back, per level).
// At the end, the deciding wave copies the control block to final_copy and
// then publishes done with a release: gg_bfs reads its statistics there and
// returns without copying or waiting for the launches still queued (they
// return at once), so the next call's launches reach the GPU sooner.
struct BfsCtl;
struct BfsHostFlag {
  uint32_t decided;  // iterations decided so far
  uint32_t done;
  BfsCtl *final_copy;  // pinned, coherent; null: the caller copies the block itself
};

// The push/pull switch (core/graph.hpp:1509): sparse iff the frontier's
// out-arcs are below |E| / divisor; with alpha, also sparse while the pull
// would have to scan more than alpha times the frontier's out-arcs (the
// unvisited vertices' in-arcs, m_left, bound it).
__device__ __forceinline__ uint32_t bfs_mode(int32_t force, uint64_t threshold, uint64_t alpha,
                                             uint64_t edg, uint64_t m_left) {
  const bool sparse = edg < threshold || (alpha && edg * alpha <= m_left);
  return force == 1 ? kBfsPush : force == 2 ? kBfsPull : sparse ? kBfsPush : kBfsPull;
}

// The decision of iteration n, made by every block of the iteration's first
// launch from the counts of the frontier iteration n - 1 produced
// (kBfsSlots spread slots, read by the block's first wave): the push/pull
// switch (core/graph.hpp:1509), or the end of the traversal.  Every block
// reads the same counts and the same control block values, so all reach the
// same mode; block 0 also records the iteration (statistics, level counts,
// the running sums, the next copy of Beamer's bound) and publishes the
// decision to the host.  The slots are cleared by the iteration's second
// launch, once every block here has read them.  Returns the mode.
__device__ uint32_t bfs_decide_block(BfsCtl *c, uint32_t n, const uint64_t *set,
                                     const unsigned long long *exam, uint32_t *lv,
                                     BfsHostFlag *hf, uint32_t *level_out = nullptr) {
  __shared__ uint32_t s_mode, s_level, s_finish;
  if (threadIdx.x < 64) {
    const int lane = lane_id();
    const uint32_t done = c->done;
    const uint32_t n0 = c->n0;
    const int32_t force = c->force;
    const uint64_t threshold = c->threshold, alpha = c->alpha, m_unv = c->m_unv[n & 1];
    uint64_t cnt = 0, edg = 0, ind = 0, x = 0;
    if (!done) {
      const uint64_t *p = set + (size_t)lane * kCountStride;  // kBfsSlots == 64
      cnt = p[0];
      edg = p[1];
      ind = p[2];
      if (exam) x = exam[lane * 16];
    }
    cnt = wave_sum_u64(cnt);
    edg = wave_sum_u64(edg);
    ind = wave_sum_u64(ind);
    x = wave_sum_u64(x);
    if (lane == 0) {
      uint32_t mode = kBfsDone;
      const uint64_t m_left = m_unv - (ind < m_unv ? ind : m_unv);
      if (!done && cnt) mode = bfs_mode(force, threshold, alpha, edg, m_left);
      s_mode = mode;
      s_level = n + n0 + 1;
      s_finish = blockIdx.x == 0 && !done && !cnt;
      if (blockIdx.x == 0 && !done) {
        const uint32_t i = n + n0;  // the level this iteration runs
        if (i > 0 && i - 1 < GG_STATS_RECORDS) {
          c->fresh[i - 1] = cnt;
          c->exam[i - 1] = x;
        }
        c->examined += x;
        if (!cnt) {
          c->done = 1;
          c->mode = kBfsDone;
          // published below, after the control block's copy
        } else {
          const uint64_t reached = c->reached + cnt;
          c->reached = reached;
          c->m_reach += edg;
          c->m_unv[(n + 1) & 1] = m_left;
          if (i < c->lv_cap) lv[i] = (uint32_t)cnt;
          if (i < GG_STATS_RECORDS) {
            c->modes[i] = mode == kBfsPush ? 0u : 1u;
            c->active[i] = cnt;
            c->unvisited[i] = (uint64_t)c->V - reached;
          }
          c->mode = mode;
          c->level = i + 1;
          c->iter = i + 1;
          if (mode == kBfsPush) c->push_iters += 1;
          __hip_atomic_store(&hf->decided, n + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
      }
    }
  }
  // the end of the traversal: wave 0 of block 0 copies the final control
  // block to host memory, then publishes done (release, system scope)
  if (blockIdx.x == 0 && threadIdx.x < 64 && s_finish) {
    __threadfence();  // lane 0's last records are performed before they are read
    BfsCtl *dst = hf->final_copy;
    if (dst) {
      const uint64_t *from = (const uint64_t *)c;
      uint64_t *to = (uint64_t *)dst;
      for (size_t k = threadIdx.x; k < sizeof(BfsCtl) / 8; k += 64) to[k] = from[k];
      __threadfence_system();
    }
    __builtin_amdgcn_wave_barrier();
    if (threadIdx.x == 0)
      __hip_atomic_store(&hf->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __syncthreads();
  if (level_out) *level_out = s_level;
  return s_mode;
}

// Output arrays to "unreached" (bfs.cpp:35-36), frontiers, count slots and
// examined counters (both copies) to 0.  Vertices without in-arcs start in
// `visited`: no step can find them (the root is seeded separately), so the
// bottom-up step skips their groups.
__global__ void bfs_init_kernel(uint32_t *parent, uint32_t *depth, uint32_t V, uint64_t *visited,
                                const uint64_t *no_in, uint64_t *f0, uint64_t *f1, size_t words,
                                uint64_t *sets, size_t set_words, unsigned long long *exam) {
  const size_t s = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  // 16-byte stores where the output arrays allow them (the parent array,
  // 4V bytes, is most of what this kernel writes)
  const size_t nq = ((uintptr_t)parent | (uintptr_t)depth) & 15 ? 0 : (size_t)V / 4;
  for (size_t i = i0; i < nq; i += s) {
    ((uint4 *)parent)[i] = make_uint4(V, V, V, V);
    if (depth) ((uint4 *)depth)[i] = make_uint4(kUnset, kUnset, kUnset, kUnset);
  }
  for (size_t i = nq * 4 + i0; i < V; i += s) {
    parent[i] = V;
    if (depth) depth[i] = kUnset;
  }
  for (size_t i = i0; i < words; i += s) {
    visited[i] = no_in[i];
    f0[i] = 0;
    f1[i] = 0;
  }
  for (size_t i = i0; i < set_words; i += s) sets[i] = 0;
  if (exam)
    for (size_t i = i0; i < 2 * kExamWords; i += s) exam[i] = 0;
}

// One block: the root (bfs.cpp:37-39), the control block and both queue
// counters.  When the first level is a push (the switch on the root's own
// counts) of at most kSeedArcs out-arcs, the block runs it: it claims the
// root's out-neighbours with a compare-and-swap on their parent entries, as
// the push does (bfs.cpp:53), sets their bits in the frontier that the
// host's iteration 0 reads, leaves their counts in iteration 0's slots, and
// records level 0 as the decision would; the host's iteration n then runs
// level n + 1 (n0 = 1), one iteration of two launches fewer.  Otherwise
// the frontier is the root, with its counts, and n0 = 0.
constexpr uint32_t kSeedArcs = 1024;
__global__ __launch_bounds__(kBlock) void bfs_seed_kernel(
    BfsCtl *c, uint64_t *set0, Counters *qc, unsigned long long *exam0, uint32_t *lv,
    uint64_t *visited, uint64_t *fin, const uint64_t *oindex, const uint32_t *onbr, uint32_t root,
    uint32_t root_orig, uint32_t *parent, uint32_t *depth, const uint4 *vinfo,
    const uint32_t *odeg, const uint32_t *ideg, uint32_t V, int32_t force, uint64_t threshold,
    uint64_t arcs, uint64_t alpha, uint32_t lv_cap) {
  __shared__ uint64_t s_sum[3][kWavesPerBlock];
  const uint4 rv = vinfo[root];
  const uint64_t rb = oindex[root], re = oindex[root + 1];
  const uint32_t r_out = rec_odeg(rv.y, odeg, root), r_in = rec_ideg(rv.y, ideg, root);
  const uint64_t m_left = arcs - (r_in < arcs ? r_in : arcs);
  const bool seed_push =
      bfs_mode(force, threshold, alpha, r_out, m_left) == kBfsPush && re - rb <= kSeedArcs;
  if (threadIdx.x == 0) {
    c->done = 0;
    c->iter = 0;
    c->mode = kBfsDone;
    c->level = 0;
    c->push_iters = 0;
    c->V = V;
    c->force = force;
    c->lv_cap = lv_cap;
    c->n0 = seed_push ? 1u : 0u;
    c->reached = 0;
    c->m_reach = 0;
    c->examined = 0;
    c->threshold = threshold;
    c->m_unv[0] = arcs;
    c->m_unv[1] = arcs;
    c->alpha = alpha;
    const uint64_t bit = 1ULL << (root & 63);
    visited[root >> 6] |= bit;
    parent[root_orig] = root_orig;
    if (depth) depth[root_orig] = 0;
    if (!seed_push) {
      fin[root >> 6] |= bit;
      set0[0] = 1;
      set0[1] = r_out;
      set0[2] = r_in;
    }
    for (int k = 0; k < 2; k++) {
      qc[k].n[0] = qc[k].n[1] = qc[k].n[2] = 0;
      qc[k].total = 0;
      qc[k].edges = 0;
    }
  }
  if (!seed_push) return;
  __syncthreads();  // the root's parent entry is set before any claim (self-loops)
  uint64_t cnt = 0, edg = 0, ind = 0;
  for (uint64_t e = rb + threadIdx.x; e < re; e += kBlock) {
    const uint32_t u = onbr[e];
    const uint4 vi = vinfo[u];
    if (atomicCAS(&parent[vi.x], V, root_orig) == V) {
      if (depth) depth[vi.x] = 1;
      atomicOr((unsigned long long *)&fin[u >> 6], 1ULL << (u & 63));
      cnt++;
      edg += rec_odeg(vi.y, odeg, u);
      ind += rec_ideg(vi.y, ideg, u);
    }
  }
  cnt = wave_sum_u64(cnt);
  edg = wave_sum_u64(edg);
  ind = wave_sum_u64(ind);
  if (lane_id() == 0) {
    s_sum[0][threadIdx.x >> 6] = cnt;
    s_sum[1][threadIdx.x >> 6] = edg;
    s_sum[2][threadIdx.x >> 6] = ind;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    cnt = edg = ind = 0;
    for (int k = 0; k < kWavesPerBlock; k++) {
      cnt += s_sum[0][k];
      edg += s_sum[1][k];
      ind += s_sum[2][k];
    }
    set0[0] = cnt;  // the frontier of level 1, for the host's iteration 0
    set0[1] = edg;
    set0[2] = ind;
    if (exam0) exam0[0] = re - rb;
    // level 0's records, as bfs_decide_block makes them for the root
    c->reached = 1;
    c->m_reach = r_out;
    c->m_unv[0] = m_left;
    if (lv_cap) lv[0] = 1;
    c->modes[0] = 0;
    c->active[0] = 1;
    c->unvisited[0] = (uint64_t)V - 1;
    c->mode = kBfsPush;
    c->level = 1;
    c->iter = 1;
    c->push_iters = 1;
  }
}

// Frontier binning of a push iteration: bitmap -> three degree-binned
// queues, the frontier joins visited, the output bitmap is cleared.
struct BfsBinArgs {
  const uint64_t *fin;
  size_t words;
  const uint32_t *deg;
  uint32_t small_max, big_max;
  uint32_t *q0, *q1, *q2;
  Counters *qc;
};

// The binning in the step launch of a push: each block takes one contiguous
// range of bitmap words (whole multiples of kBlock), and adds its bin sizes to
// the queue counters once, so a few hundred blocks, not one per word round,
// meet on those three counters.  Push frontiers are sparse (most words hold
// no set bit, few more than one), so each thread handles kBinWords words of
// the range at a time: it loads them together, issues the degree loads of up
// to two set bits per word together (more rounds only for denser words), and
// keeps each word's bins as two masks, so the second walk, which writes the
// queue entries, loads nothing.  A block whose range holds no set bit returns
// after the first walk.  Two words per thread keep the step kernel within
// the 64 vector registers of the bottom-up step (the launch has a thread
// for every word of an RMAT-25 bitmap, so one pass covers most ranges).
constexpr int kBinWords = 2;
__device__ __forceinline__ void bfs_bin_step(const BfsBinArgs &b, uint64_t *visited,
                                             uint64_t *clear) {
  __shared__ uint32_t s_tot[3][kWavesPerBlock];
  __shared__ uint32_t s_base[3];
  const int lane = lane_id(), wv = wave_in_block();
  const size_t span = (size_t)gridDim.x * kBlock;
  const size_t per = (b.words + span - 1) / span * kBlock;
  const size_t beg = (size_t)blockIdx.x * per;
  const size_t end = beg + per < b.words ? beg + per : b.words;
  for (size_t base = beg; base < end; base += (size_t)kBinWords * kBlock) {  // block-uniform
    uint64_t wd[kBinWords], ms[kBinWords], mm[kBinWords], rem[kBinWords];
#pragma unroll
    for (int t = 0; t < kBinWords; t++) {
      const size_t wi = base + (size_t)t * kBlock + threadIdx.x;
      wd[t] = wi < end ? b.fin[wi] : 0;
    }
#pragma unroll
    for (int t = 0; t < kBinWords; t++) {
      const size_t wi = base + (size_t)t * kBlock + threadIdx.x;
      if (wi < end) {
        clear[wi] = 0;
        if (wd[t]) visited[wi] |= wd[t];
      }
      ms[t] = mm[t] = 0;
      rem[t] = wd[t];
    }
    bool more = true;
    while (more) {
      uint32_t k0[kBinWords], k1[kBinWords], d0[kBinWords], d1[kBinWords];
#pragma unroll
      for (int t = 0; t < kBinWords; t++) {
        const size_t wi = base + (size_t)t * kBlock + threadIdx.x;
        k0[t] = rem[t] ? (uint32_t)__builtin_ctzll(rem[t]) : 64u;
        rem[t] &= rem[t] - 1;
        k1[t] = rem[t] ? (uint32_t)__builtin_ctzll(rem[t]) : 64u;
        rem[t] &= rem[t] - 1;
        d0[t] = k0[t] < 64 ? b.deg[wi * 64 + k0[t]] : 0;
        d1[t] = k1[t] < 64 ? b.deg[wi * 64 + k1[t]] : 0;
      }
      more = false;
#pragma unroll
      for (int t = 0; t < kBinWords; t++) {
        if (k0[t] < 64) {
          if (d0[t] <= b.small_max) ms[t] |= 1ULL << k0[t];
          else if (d0[t] <= b.big_max) mm[t] |= 1ULL << k0[t];
        }
        if (k1[t] < 64) {
          if (d1[t] <= b.small_max) ms[t] |= 1ULL << k1[t];
          else if (d1[t] <= b.big_max) mm[t] |= 1ULL << k1[t];
        }
        more = more || rem[t] != 0;
      }
    }
    uint32_t c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
    for (int t = 0; t < kBinWords; t++) {
      c0 += (uint32_t)__popcll(ms[t]);
      c1 += (uint32_t)__popcll(mm[t]);
      c2 += (uint32_t)__popcll(wd[t] & ~ms[t] & ~mm[t]);
    }
    if (!__syncthreads_or((c0 | c1 | c2) != 0)) continue;
    uint32_t t0, t1, t2;
    uint32_t o0 = wave_excl_scan(c0, &t0);
    uint32_t o1 = wave_excl_scan(c1, &t1);
    uint32_t o2 = wave_excl_scan(c2, &t2);
    if (lane == 0) {
      s_tot[0][wv] = t0;
      s_tot[1][wv] = t1;
      s_tot[2][wv] = t2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t b0 = 0, b1 = 0, b2 = 0;
      for (int i = 0; i < kWavesPerBlock; i++) {
        b0 += s_tot[0][i];
        b1 += s_tot[1][i];
        b2 += s_tot[2][i];
      }
      s_base[0] = b0 ? atomicAdd(&b.qc->n[0], b0) : 0;
      s_base[1] = b1 ? atomicAdd(&b.qc->n[1], b1) : 0;
      s_base[2] = b2 ? atomicAdd(&b.qc->n[2], b2) : 0;
    }
    __syncthreads();
    o0 += s_base[0];
    o1 += s_base[1];
    o2 += s_base[2];
    for (int i = 0; i < wv; i++) {
      o0 += s_tot[0][i];
      o1 += s_tot[1][i];
      o2 += s_tot[2][i];
    }
#pragma unroll
    for (int t = 0; t < kBinWords; t++) {
      const size_t wi = base + (size_t)t * kBlock + threadIdx.x;
      for (uint64_t w = wd[t]; w; w &= w - 1) {
        const int k = __builtin_ctzll(w);
        const uint32_t v = (uint32_t)(wi * 64 + k);
        if ((ms[t] >> k) & 1ULL)
          b.q0[o0++] = v;
        else if ((mm[t] >> k) & 1ULL)
          b.q1[o1++] = v;
        else
          b.q2[o2++] = v;
      }
    }
    __syncthreads();  // s_tot / s_base are reused by the next chunk
  }
}

struct BfsPushArgs {
  const uint64_t *index;
  const uint32_t *nbr;
  const uint32_t *q0, *q1, *q2;
  const Counters *qc;
};

// Persistent grid over the three bins: 64 small vertices per wave step
// (in-wave load balancing), one mid vertex per wave step, and chunks of every
// huge vertex shared by all waves.
// W waves take part (kBlock threads per block); gw is this wave's index.
__device__ __forceinline__ void bfs_push_step(const BfsCtl *c, const BfsOut &o,
                                              const BfsPushArgs &a, uint64_t *set_out,
                                              unsigned long long *exam, uint32_t W,
                                              uint32_t gw) {
  const uint32_t level = c->level;
  const int lane = lane_id();
  const uint32_t n0 = a.qc->n[0], n1 = a.qc->n[1], n2 = a.qc->n[2];
  const uint32_t waves0 = (n0 + 63) / 64;
  uint64_t cnt = 0, edg = 0, ind = 0, ex = 0;
  for (uint32_t vw = gw; vw < waves0 + n1; vw += W) {
    if (vw < waves0) {
      const uint32_t i = vw * 64 + lane;
      const bool valid = i < n0;
      const uint32_t v = valid ? a.q0[i] : 0;
      const uint32_t vo = valid ? o.vinfo[v].x : 0;
      const uint64_t b = valid ? a.index[v] : 0;
      const uint32_t d = valid ? (uint32_t)(a.index[v + 1] - b) : 0;
      uint32_t total;
      const uint32_t excl = wave_excl_scan(d, &total);
      for (uint32_t j0 = 0; j0 < total; j0 += 128) {
        const uint32_t ja = j0 + lane, jb = j0 + 64 + lane;
        const int ka = wave_owner(excl, ja), kb = wave_owner(excl, jb);
        const uint32_t sa = __shfl(vo, ka), sb = __shfl(vo, kb);
        const uint64_t ba = shfl_u64(b, ka), bbb = shfl_u64(b, kb);
        const uint32_t ea = __shfl(excl, ka), eb = __shfl(excl, kb);
        const bool oka = ja < total, okb = jb < total;
        const uint32_t da = oka ? stream_load(&a.nbr[ba + (ja - ea)]) : 0;
        const uint32_t db = okb ? stream_load(&a.nbr[bbb + (jb - eb)]) : 0;
        bfs_claim2(o, level, oka, sa, da, okb, sb, db, cnt, edg, ind);
      }
      if (lane == 0) ex += total;
    } else {
      const uint32_t v = a.q1[vw - waves0];
      const uint32_t vo = o.vinfo[v].x;
      const uint64_t b = a.index[v], e1 = a.index[v + 1];
      for (uint64_t e = b + lane; e < e1; e += 128) {
        const bool okb = e + 64 < e1;
        const uint32_t da = stream_load(&a.nbr[e]);
        const uint32_t db = okb ? stream_load(&a.nbr[e + 64]) : 0;
        bfs_claim2(o, level, true, vo, da, okb, vo, db, cnt, edg, ind);
      }
      if (lane == 0) ex += e1 - b;
    }
  }
  // Huge vertices: their chunks form one global sequence, chunk g going to
  // wave g % W.  Each block stages kHubBatch of them at a time in LDS (list
  // bounds and the prefix of chunk counts), so a wave finds the vertex of a
  // chunk by a binary search in LDS instead of walking the list.
  __shared__ uint32_t s_v[kHubBatch];  // original ids (the parent written for claims)
  __shared__ uint64_t s_b[kHubBatch], s_e[kHubBatch];
  __shared__ uint64_t s_pre[kHubBatch + 1];
  __shared__ uint64_t s_wsum[kWavesPerBlock];
  uint64_t coff = 0;  // chunks of the earlier batches
  for (uint32_t base = 0; base < n2; base += kHubBatch) {
    const uint32_t nb = n2 - base < kHubBatch ? n2 - base : kHubBatch;
    __syncthreads();  // the previous batch is done with the LDS arrays
    uint64_t mine[kHubBatch / kBlock];
    uint64_t tsum = 0;
#pragma unroll
    for (int r = 0; r < (int)(kHubBatch / kBlock); r++) {
      const uint32_t i = threadIdx.x * (kHubBatch / kBlock) + r;
      mine[r] = 0;
      if (i < nb) {
        const uint32_t v = a.q2[base + i];
        const uint64_t b = a.index[v], e1 = a.index[v + 1];
        s_v[i] = o.vinfo[v].x;
        s_b[i] = b;
        s_e[i] = e1;
        mine[r] = (e1 - b + kPushChunk - 1) / kPushChunk;
      }
      tsum += mine[r];
    }
    // exclusive scan of the per-thread sums over the block
    uint64_t inc = tsum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t y = shfl_u64(inc, lane >= d ? lane - d : lane);
      if (lane >= d) inc += y;
    }
    if (lane == 63) s_wsum[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint64_t wbase = 0, total = 0;
    for (int i = 0; i < kWavesPerBlock; i++) {
      if (i < (int)(threadIdx.x >> 6)) wbase += s_wsum[i];
      total += s_wsum[i];
    }
    uint64_t run = wbase + inc - tsum;
#pragma unroll
    for (int r = 0; r < (int)(kHubBatch / kBlock); r++) {
      const uint32_t i = threadIdx.x * (kHubBatch / kBlock) + r;
      if (i < nb) s_pre[i] = run;
      run += mine[r];
    }
    if (threadIdx.x == 0) s_pre[nb] = total;
    __syncthreads();
    for (uint64_t j = (gw + W - coff % W) % W; j < total; j += W) {
      uint32_t lo = 0, hi = nb;  // last vertex with s_pre <= j
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_pre[mid] <= j)
          lo = mid;
        else
          hi = mid;
      }
      const uint32_t v = s_v[lo];
      const uint64_t e1 = s_e[lo];
      const uint64_t cb = s_b[lo] + (j - s_pre[lo]) * kPushChunk;
      const uint64_t ce = cb + kPushChunk < e1 ? cb + kPushChunk : e1;
      for (uint64_t e = cb + lane; e < ce; e += 128) {
        const bool okb = e + 64 < ce;
        const uint32_t da = stream_load(&a.nbr[e]);
        const uint32_t db = okb ? stream_load(&a.nbr[e + 64]) : 0;
        bfs_claim2(o, level, true, v, da, okb, v, db, cnt, edg, ind);
      }
      if (lane == 0) ex += ce - cb;
    }
    coff += total;
  }
  bfs_wave_flush(set_out, gw, cnt, edg, ind, exam, ex);
}

struct BfsPullArgs {
  const uint64_t *index;
  const uint32_t *nbr;
  const uint32_t *first_orig;  // original id of each vertex's first in-neighbour
  const uint64_t *fin;
  uint32_t V, light_max, ngroups;
  uint32_t start32;  // arcs < 2^32: the records hold the list starts
};

// Bottom-up step (bfs.cpp:60-76) over 64-vertex groups, a persistent grid:
// an unvisited vertex scans its in-neighbours until one is in the frontier.
// A wave takes its groups 64 at a time (lane j loads the bitmap words of
// group gw + j*W), lines up the unvisited vertices of all of them, and works
// through that sequence in batches of 64, one vertex per lane: so a sparse
// level (a few unvisited vertices per group) pays one dependent chain per 64
// vertices instead of one per group, and a dense level one per group as
// before.  In a batch, the first in-neighbour (the highest-degree one on a
// relabelled graph) decides most vertices without touching the list; its
// original id is read alongside, so the parent needs no further load.  The
// lists still open (up to light_max entries; longer ones are left to the
// chunk kernel, which runs after this one) are then scanned as one sequence
// spread evenly over the lanes, 64 entries per round, a hit handed to the
// list's owner lane through a per-wave LDS slot (any hit is a valid parent,
// as any in-neighbour in the frontier is in bfs.cpp:66-70); rounds stop once
// every open list has a hit.  Found vertices set their bit in their group's
// LDS word; the wave then writes each group's words of the output frontier
// and of visited (it owns them; the chunk step only ORs bits in later).
// W waves take part, blocks of kBlock threads; gw is this wave's index.
// kCount: the launch counts adjacency entries for the roofline's byte model
// (count_examined): a list's entries up to its first hit in list order, as a
// serial scan stops (bfs.cpp:66-67), so each round hands the owner its first
// hit (the lowest entry, an LDS atomic min) instead of any one.
// kStart32 (arcs < 2^32): a list's start comes with the vertex's record, so
// a scan needs no load of the offsets first.
template <bool kCount, bool kStart32>
__device__ __forceinline__ void bfs_pull_step(uint32_t level, const BfsOut &o,
                                              const BfsPullArgs &a, uint64_t *set_out,
                                              unsigned long long *exam, uint32_t W,
                                              uint32_t gw) {
  const int lane = lane_id();
  const int wv = (int)(threadIdx.x >> 6) % kWavesPerBlock;
  // a hit per owner lane: its vertex; counting: (entry << 32) | vertex
  using Hit = typename std::conditional<kCount, unsigned long long, uint32_t>::type;
  constexpr Hit kNoHit = (Hit)~(Hit)0;
  __shared__ Hit s_hit[kWavesPerBlock][64];
  __shared__ unsigned long long s_found[kWavesPerBlock][64];  // found bits per group
  s_hit[wv][lane] = kNoHit;
  s_found[wv][lane] = 0;
  wave_sync();
  uint64_t cnt = 0, edg = 0, ind = 0, examined = 0;
  for (uint64_t r0 = gw; r0 < a.ngroups; r0 += (uint64_t)W * 64) {
    const uint64_t gj = r0 + (uint64_t)lane * W;  // this lane's group
    const bool gvalid = gj < a.ngroups;
    uint64_t visw = ~0ULL, open = 0;
    if (gvalid) {
      visw = o.visited[gj] | a.fin[gj];
      open = ~visw;  // vertices without in-arcs start visited
      const uint64_t vb = gj * 64;
      if (vb + 64 > a.V) open &= (1ULL << (a.V - vb)) - 1;  // 0 < V - vb < 64
    }
    uint32_t total;
    const uint32_t gexcl = wave_excl_scan((uint32_t)__popcll(open), &total);
    for (uint32_t c0 = 0; c0 < total; c0 += 64) {
      const uint32_t c = c0 + lane;
      const bool ok = c < total;
      const int gk = wave_owner(gexcl, ok ? c : total - 1);  // the group of vertex c
      const uint64_t ok_open = shfl_u64(open, gk);
      const uint32_t rk = c - (uint32_t)__shfl((int)gexcl, gk);
      const uint32_t v = (uint32_t)((r0 + (uint64_t)gk * W) * 64 + (ok ? nth_set_bit(ok_open, rk) : 0));
      bool cand = ok;
      const uint4 vi = cand ? o.vinfo[v] : make_uint4(0, 0, 0, 0);
      const uint32_t deg = cand ? rec_ideg(vi.y, o.ideg, v) : 0;
      cand = cand && deg > 0;
      uint32_t found = kUnset;
      bool head = false;  // found at the first in-neighbour
      uint32_t head_orig = 0;
      if (cand) {
        head_orig = a.first_orig[v];
        head = test_bit(a.fin, vi.z);
        if (head) found = vi.z;
        examined += 1;
      }
      cand = cand && found == kUnset && deg > 1 && deg <= a.light_max;
      uint64_t b = 0;
      uint32_t len = 0;
      if (cand) {
        b = (kStart32 ? (uint64_t)vi.w : a.index[v]) + 1;
        len = deg - 1;
      }
      uint32_t ltotal;
      const uint32_t lexcl = wave_excl_scan(len, &ltotal);
      for (uint32_t j0 = 0; j0 < ltotal; j0 += 64) {
        const uint32_t j = j0 + lane;
        const bool jok = j < ltotal;
        const int k = wave_owner(lexcl, jok ? j : ltotal - 1);
        const uint64_t bk = shfl_u64(b, k);
        const uint32_t exk = (uint32_t)__shfl((int)lexcl, k);
        const bool open_k = (uint32_t)__shfl((int)found, k) == kUnset;
        bool hit = false;
        uint32_t u = 0;
        if (jok && open_k) {
          u = stream_load(&a.nbr[bk + (j - exk)]);
          hit = test_bit(a.fin, u);
        }
        if constexpr (kCount) {
          if (hit) atomicMin(&s_hit[wv][k], ((unsigned long long)j << 32) | u);
          wave_sync();
          const uint32_t first_k = (uint32_t)(s_hit[wv][k] >> 32);
          examined += (uint64_t)__popcll(__ballot(jok && open_k && first_k >= j)) * (lane == 0);
          wave_sync();
        } else {
          examined += (uint64_t)__popcll(__ballot(jok && open_k)) * (lane == 0);
          if (hit) s_hit[wv][k] = u;
          wave_sync();
        }
        const Hit h = s_hit[wv][lane];
        if (h != kNoHit) {
          if (found == kUnset) found = (uint32_t)h;
          s_hit[wv][lane] = kNoHit;
        }
        wave_sync();
        if (!__ballot(cand && found == kUnset)) break;
      }
      if (found != kUnset) {
        // both loads (the parent's id, a capped degree) before the writes;
        // the writes are non-temporal: nothing in this launch reads them, and
        // lines left dirty in L2 are written back at the kernel boundary
        const uint32_t pval = head ? head_orig : o.vinfo[found].x;
        edg += rec_odeg(vi.y, o.odeg, v);
        ind += deg;
        __builtin_nontemporal_store(pval, &o.parent[vi.x]);
        if (o.depth) __builtin_nontemporal_store(level, &o.depth[vi.x]);
        atomicOr(&s_found[wv][gk], 1ULL << (v & 63));
      }
    }
    wave_sync();
    if (gvalid) {
      const uint64_t m = s_found[wv][lane];
      o.fout[gj] = m;
      o.visited[gj] = visw;  // this level joins visited; m joins next iteration
      cnt += (uint64_t)__popcll(m);
      s_found[wv][lane] = 0;
    }
    wave_sync();
  }
  bfs_wave_flush(set_out, gw, cnt, edg, ind, kCount ? exam : nullptr, examined);
}

struct BfsHeavyArgs {
  const uint32_t *nbr;
  const uint32_t *chunk_v;
  const uint64_t *chunk_beg, *chunk_end;
  const uint64_t *fin;
  uint32_t n_chunks;
};

// In-lists above light_max, one wave per chunk (grid-stride); chunks of one
// vertex race and the compare-and-swap on its parent entry keeps one.  The
// chunks poll that entry to stop early once another chunk has won.
__device__ __forceinline__ void bfs_heavy_step(const BfsCtl *c, const BfsOut &o,
                                               const BfsHeavyArgs &a, uint64_t *set_out,
                                               unsigned long long *exam, uint32_t W,
                                               uint32_t gw) {
  const uint32_t level = c->level;
  const int lane = lane_id();
  uint64_t cnt = 0, edg = 0, ind = 0, examined = 0;
  // lane j tests chunk base + j*W (the chunks of one vertex stay spread over
  // the waves), then the wave scans the chunks whose vertex is still open
  for (uint64_t base = gw; base < a.n_chunks; base += (uint64_t)W * 64) {
    const uint64_t mine = base + (uint64_t)lane * W;
    uint32_t vl = 0;
    bool open = false;
    if (mine < a.n_chunks) {
      vl = a.chunk_v[mine];
      // found earlier, in this level's frontier, or at its head by the pull
      // kernel (or by another chunk) this level
      open = !(((o.visited[vl >> 6] | a.fin[vl >> 6] | o.fout[vl >> 6]) >> (vl & 63)) & 1ULL);
    }
    for (uint64_t todo = __ballot(open); todo; todo &= todo - 1) {
      const int j = __builtin_ctzll(todo);
      const uint64_t ch = base + (uint64_t)j * W;
      const uint32_t v = (uint32_t)__shfl((int)vl, j);
      const uint64_t bit = 1ULL << (v & 63);
      const uint4 vi = o.vinfo[v];
      const uint32_t vod = rec_odeg(vi.y, o.odeg, v), vid = rec_ideg(vi.y, o.ideg, v);
      uint32_t *pv = &o.parent[vi.x];
      const uint64_t cb = a.chunk_beg[ch], ce = a.chunk_end[ch];
      for (uint64_t e0 = cb; e0 < ce; e0 += 64) {
        if (__hip_atomic_load(pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != o.none) break;
        const uint64_t e = e0 + lane;
        const uint32_t u = e < ce ? stream_load(&a.nbr[e]) : 0;
        const bool hit = e < ce && test_bit(a.fin, u);
        const uint64_t hm = __ballot(hit);
        examined += hm ? (uint64_t)(__builtin_ctzll(hm) + 1) : (ce - e0 < 64 ? ce - e0 : 64);
        if (hm) {
          const uint32_t fu = __shfl(u, __builtin_ctzll(hm));
          if (lane == 0 && atomicCAS(pv, o.none, o.vinfo[fu].x) == o.none) {
            if (o.depth) o.depth[vi.x] = level;
            atomicOr((unsigned long long *)&o.fout[v >> 6], (unsigned long long)bit);
            cnt++;
            edg += vod;
            ind += vid;
          }
          break;
        }
      }
    }
  }
  if (lane != 0) examined = 0;  // counted once per wave
  bfs_wave_flush(set_out, gw, cnt, edg, ind, exam, examined);
}

// The two launches of an iteration.  A launch of the mode not chosen would
// return at once but still cost a kernel boundary, so each launch serves
// both modes:
//   step       -- every block: the decision of iteration n; then, pull: the
//                 bottom-up step; push: the frontier binning;
//   push_heavy -- push: the persistent push over the queues (push_blocks);
//                 pull: the heavy in-list chunks (heavy_blocks).  Block 0
//                 first clears what the decision read or what the next
//                 iteration's binning fills: the count slots and examined
//                 counters of iteration n's decision, and the queue counters
//                 of iteration n + 1.
// The step kernel is held to 80 scalar and 64 vector registers (8 waves per
// SIMD; the few values that do not fit spill outside the batch loop): with
// more scalar registers the bottom-up step ran 14-30% slower at the same
// vector count, as with fewer waves per SIMD.  kCount: the examined
// counters are kept (without, their pointers are null at compile time).
template <bool kCount, bool kStart32>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_num_sgpr(80), amdgpu_waves_per_eu(8, 8)))
void bfs_step_kernel(BfsCtl *c, uint32_t n, const uint64_t *set_in,
                     const unsigned long long *exam_in_, uint32_t *lv, BfsHostFlag *hf, BfsOut o,
                     BfsBinArgs b, BfsPullArgs a, uint64_t *set_out,
                     unsigned long long *exam_out_) {
  const unsigned long long *exam_in = kCount ? exam_in_ : nullptr;
  unsigned long long *exam_out = kCount ? exam_out_ : nullptr;
  uint32_t level;
  const uint32_t mode = bfs_decide_block(c, n, set_in, exam_in, lv, hf, &level);
  if (mode == kBfsPush)
    bfs_bin_step(b, o.visited, o.fout);
  else if (mode == kBfsPull)
    bfs_pull_step<kCount, kStart32>(level, o, a, set_out, exam_out, gridDim.x * kWavesPerBlock,
                                    global_wave());
}

// Blocks past the mode's count return at once.  kCount: the examined
// counters are kept (without, their pointers are null at compile time).  At
// most 80 scalar registers (the rest spill to vector lanes, within the 64
// vector registers of 8 waves per SIMD), as for the bottom-up kernel.
template <bool kCount>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_num_sgpr(80))) void
bfs_push_heavy_kernel(const BfsCtl *c, BfsOut o, BfsPushArgs p, BfsHeavyArgs h, uint64_t *set_out,
                      unsigned long long *exam_, uint32_t push_blocks, uint32_t heavy_blocks,
                      uint64_t *set_spent, unsigned long long *exam_spent_, Counters *qc_next) {
  unsigned long long *exam = kCount ? exam_ : nullptr;
  unsigned long long *exam_spent = kCount ? exam_spent_ : nullptr;
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    uint64_t *sp = set_spent + (size_t)threadIdx.x * kCountStride;  // kBfsSlots == 64
    sp[0] = 0;
    sp[1] = 0;
    sp[2] = 0;
    if (exam_spent) exam_spent[threadIdx.x * 16] = 0;
    if (threadIdx.x == 0) {
      qc_next->n[0] = qc_next->n[1] = qc_next->n[2] = 0;
      qc_next->total = 0;
      qc_next->edges = 0;
    }
  }
  const uint32_t mode = c->mode;
  if (mode == kBfsPush) {
    if (blockIdx.x < push_blocks)
      bfs_push_step(c, o, p, set_out, exam, push_blocks * kWavesPerBlock, global_wave());
  } else if (mode == kBfsPull && h.n_chunks && blockIdx.x < heavy_blocks) {
    bfs_heavy_step(c, o, h, set_out, exam, heavy_blocks * kWavesPerBlock, global_wave());
  }
}

// Device state of the BFS loop, kept in the traversal context.
struct BfsState {
  BfsCtl *ctl = nullptr;
  uint64_t *sets = nullptr;   // 2 x kBfsSetWords
  Counters *qc = nullptr;     // 2 (by iteration parity)
  unsigned long long *exam = nullptr;  // 2 x kExamWords
  uint32_t *lv = nullptr;
  uint32_t lv_cap = 0;
  BfsCtl *h_ctl = nullptr;      // pinned
  BfsHostFlag *h_flag = nullptr;  // pinned, coherent: written by the device
};

void bfs_state_free(void *p) {
  BfsState *s = (BfsState *)p;
  if (!s) return;
  device_free(s->ctl);
  device_free(s->sets);
  device_free(s->qc);
  device_free(s->exam);
  device_free(s->lv);
  if (s->h_ctl) (void)hipHostFree(s->h_ctl);
  if (s->h_flag) (void)hipHostFree(s->h_flag);
  delete s;
}

static int bfs_state(Work *w, uint32_t lv_need, BfsState **out) {
  BfsState *s = (BfsState *)w->bfs;
  if (!s) {
    s = new BfsState;
    w->bfs = s;
    GG_TRY(device_alloc((void **)&s->ctl, sizeof(BfsCtl)));
    GG_TRY(device_alloc((void **)&s->sets, 2 * kBfsSetWords * 8));
    GG_TRY(device_alloc((void **)&s->qc, 2 * sizeof(Counters)));
    GG_TRY(device_alloc((void **)&s->exam, 2 * kExamWords * 8));
    GG_HIP(hipHostMalloc((void **)&s->h_ctl, sizeof(BfsCtl), hipHostMallocCoherent));
    GG_HIP(hipHostMalloc((void **)&s->h_flag, sizeof(BfsHostFlag), hipHostMallocCoherent));
  }
  if (lv_need > s->lv_cap) {
    device_free(s->lv);
    s->lv = nullptr;
    s->lv_cap = 0;
    GG_TRY(device_alloc((void **)&s->lv, (size_t)lv_need * 4));
    s->lv_cap = lv_need;
  }
  *out = s;
  return GG_OK;
}

// The device buffers one traversal uses, wherever they live (the graph's
// traversal context, or the allocation-free variant's scratch array).
struct BfsBufs {
  BfsCtl *ctl;
  uint64_t *sets;  // 2 x kBfsSetWords
  Counters *qc;    // 2
  uint32_t *lv;
  uint32_t lv_cap;
  BfsHostFlag *hf;  // where the decision is published to the host
  unsigned long long *exam;  // 2 x kExamWords; null unless count_examined
  uint64_t *visited, *f[2];
  uint32_t *q[3];
  size_t words;  // V / 64 + 1
};

// Launch shapes and arguments fixed for one traversal.
struct BfsPlan {
  uint32_t V, r_int, root;
  uint32_t *parent, *depth;
  unsigned bin_blocks, pull_grid, push_blocks, heavy_blocks, grid_b;
  uint32_t ngroups;
};

static BfsPlan bfs_plan(const gg_graph *g, uint32_t root, uint32_t *parent, uint32_t *depth,
                        size_t words) {
  const Dir &in = in_dir(g);
  const Options &opt = options();
  BfsPlan p;
  p.V = g->vertices;
  p.root = root;
  p.r_int = g->to_internal(root);
  p.parent = parent;
  p.depth = depth;
  // one block per 1024 bitmap words (at most 1024 blocks): every block makes
  // the decision, so fewer blocks read the count slots
  p.bin_blocks = compact_grid(words);
  // the bottom-up step walks the groups that hold a vertex with in-arcs
  // (the ones after them can never be found) and the root's: its frontier
  // and visited words must be carried like any other
  p.ngroups = std::min<uint32_t>((p.V + 63) / 64, std::max<uint32_t>(in.live_groups, p.r_int / 64 + 1));
  const uint64_t pw = (uint64_t)opt.bfs_pull_waves;
  p.pull_grid = grid_waves(pw ? std::min<uint64_t>(p.ngroups, pw) : p.ngroups);
  p.push_blocks = grid_waves((uint64_t)opt.huge_grid_waves);
  // the heavy-chunk step finds little left after the bottom-up step's head
  // tests: half the pull grid (its launch cost grows with the grid)
  const uint64_t nch = std::max(in.n_chunks, 1u), hw = pw / 2;
  p.heavy_blocks = grid_waves(hw ? std::min<uint64_t>(nch, hw) : nch);
  p.grid_b = std::max(p.push_blocks, in.n_chunks ? p.heavy_blocks : 1u);
  return p;
}

// Output arrays and buffers to their initial state, the root, and the first
// decision.
static void bfs_issue_start(const gg_graph *g, const BfsPlan &p, const BfsBufs &b, hipStream_t st) {
  const Dir &in = in_dir(g);
  const Options &opt = options();
  const uint64_t threshold = g->edges / (uint64_t)opt.switch_divisor;
  hipLaunchKernelGGL(bfs_init_kernel, dim3(grid1(std::max<uint64_t>((uint64_t)p.V / 4, b.words))),
                     dim3(kBlock), 0, st, p.parent, p.depth,
                     p.V, b.visited, in.empty, b.f[0], b.f[1], b.words, b.sets, 2 * kBfsSetWords,
                     b.exam);
  const Dir &out = out_dir(g);
  hipLaunchKernelGGL(bfs_seed_kernel, dim3(1), dim3(kBlock), 0, st, b.ctl, b.sets, b.qc, b.exam,
                     b.lv, b.visited, b.f[0], out.index, out.nbr, p.r_int, p.root, p.parent,
                     p.depth, in.vinfo, out.degree, in.degree, p.V, (int32_t)opt.force_mode,
                     threshold, g->arcs, (uint64_t)opt.bfs_alpha, b.lv_cap);
}

// The two launches of iteration n (step, push_heavy).  Buffers by
// parity: the frontier bitmaps, the count slots (iteration n decides on
// sets[n & 1] and counts into sets[(n + 1) & 1]), the examined counters
// likewise, and the queue counters (qc[n & 1] binned and pushed in
// iteration n).
static void bfs_issue_iteration(const gg_graph *g, const BfsPlan &p, const BfsBufs &b, uint32_t n,
                                hipStream_t st, bool prof = true) {
  const Dir &out = out_dir(g), &in = in_dir(g);
  const Options &opt = options();
  uint64_t *fin = b.f[n & 1], *fout = b.f[(n + 1) & 1];
  uint64_t *set_in = b.sets + (size_t)(n & 1) * kBfsSetWords;
  uint64_t *set_out = b.sets + (size_t)((n + 1) & 1) * kBfsSetWords;
  unsigned long long *exam_in = b.exam ? b.exam + (n & 1) * kExamWords : nullptr;
  unsigned long long *exam_out = b.exam ? b.exam + ((n + 1) & 1) * kExamWords : nullptr;
  Counters *qc = b.qc + (n & 1), *qc_next = b.qc + ((n + 1) & 1);
  BfsOut o{p.parent, p.depth, in.vinfo, b.visited, fout, p.V, out.degree, in.degree};
  {
    ProfScope ps(prof ? "bfs_step" : nullptr);
    BfsBinArgs ba{fin, b.words, out.degree, (uint32_t)opt.push_small_max,
                  (uint32_t)opt.push_big_max, b.q[0], b.q[1], b.q[2], qc};
    BfsPullArgs pa{in.index, in.nbr, in.first_orig, fin, p.V, in.light_max, p.ngroups,
                   g->arcs < (1ULL << 32) && opt.bfs_record_starts ? 1u : 0u};
    auto kern = exam_out ? (pa.start32 ? bfs_step_kernel<true, true> : bfs_step_kernel<true, false>)
                         : (pa.start32 ? bfs_step_kernel<false, true> : bfs_step_kernel<false, false>);
    hipLaunchKernelGGL(kern, dim3(std::max(p.pull_grid, p.bin_blocks)), dim3(kBlock), 0, st, b.ctl,
                       n, set_in, exam_in, b.lv, b.hf, o, ba, pa, set_out, exam_out);
  }
  {
    ProfScope ps(prof ? "bfs_push_heavy" : nullptr);
    BfsPushArgs a{out.index, out.nbr, b.q[0], b.q[1], b.q[2], qc};
    BfsHeavyArgs h{in.nbr, in.chunk_v, in.chunk_beg, in.chunk_end, fin, in.n_chunks};
    hipLaunchKernelGGL(exam_out ? bfs_push_heavy_kernel<true> : bfs_push_heavy_kernel<false>,
                       dim3(p.grid_b), dim3(kBlock), 0, st, b.ctl, o, a, h, set_out, exam_out,
                       p.push_blocks, p.heavy_blocks, set_in, exam_in, qc_next);
  }
}

// Statistics from the control block read back after the traversal.
static void bfs_stats(const BfsCtl &c, double ms, gg_run_stats *stats) {
  gg_run_stats r;
  memset(&r, 0, sizeof(r));
  r.iterations = c.iter;
  r.push_iterations = c.push_iters;
  r.edges_examined = c.examined;
  r.m_reach = c.m_reach;
  r.reached = c.reached;
  r.device_ms = ms;
  for (uint32_t i = 0; i < std::min<uint32_t>(c.iter, GG_STATS_RECORDS); i++) {
    r.modes[i] = c.modes[i];
    r.it_active[i] = c.active[i];
    r.it_new[i] = c.fresh[i];
    r.it_examined[i] = c.exam[i];
    r.it_unvisited[i] = c.unvisited[i];
  }
  put_stats(stats, r);
}

// ------------------------------------------------- allocation-free variant
// gg_bfs_noalloc keeps every buffer of the traversal in `scratch`, one
// array per GPU that lives as long as the process (the loader places it
// with the code object), so the call itself allocates nothing: no
// hipMalloc, no pinned host memory, no events.  Its size bounds the graphs
// it takes (kNoAllocMaxVertices) and the level counts it records.  Calls on
// one GPU share the array: calls that overlap on a GPU race.
constexpr uint32_t kNoAllocMaxVertices = 1u << 20;
constexpr uint32_t kNoAllocMaxLevels = 1u << 16;
constexpr size_t kNoAllocWords = kNoAllocMaxVertices / 64 + 1;
constexpr size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }
// byte offsets of the buffers in the array
constexpr size_t kNaCtl = 0;
constexpr size_t kNaSets = na_align(kNaCtl + sizeof(BfsCtl));
constexpr size_t kNaQc = na_align(kNaSets + 2 * kBfsSetWords * 8);
constexpr size_t kNaFlag = na_align(kNaQc + 2 * sizeof(Counters));
constexpr size_t kNaExam = na_align(kNaFlag + sizeof(BfsHostFlag));
constexpr size_t kNaLv = na_align(kNaExam + 2 * kExamWords * 8);
constexpr size_t kNaVisited = na_align(kNaLv + (size_t)kNoAllocMaxLevels * 4);
constexpr size_t kNaF0 = na_align(kNaVisited + kNoAllocWords * 8);
constexpr size_t kNaF1 = na_align(kNaF0 + kNoAllocWords * 8);
constexpr size_t kNaQ0 = na_align(kNaF1 + kNoAllocWords * 8);
constexpr size_t kNaQ1 = na_align(kNaQ0 + (size_t)kNoAllocMaxVertices * 4);
constexpr size_t kNaQ2 = na_align(kNaQ1 + (size_t)kNoAllocMaxVertices * 4);
constexpr size_t kNoAllocBytes = na_align(kNaQ2 + (size_t)kNoAllocMaxVertices * 4);
static_assert(kNoAllocBytes < ((size_t)14 << 20), "scratch stays a modest 14 MiB");

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

// The array's device address on each GPU (the host-side name of `scratch`
// is not one), looked up once per device with that device current.
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

int bfs_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(GG_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return GG_OK;
  void *ptr = nullptr;
  GG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail(GG_EHIP, "scratch array has no device address");
  g_scratch_ptr[device] = ptr;
  return GG_OK;
}

static int scratch_bufs(uint32_t V, int cap, BfsBufs *b) {
  int dev = -1;
  GG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(GG_EINVAL, "device index out of range");
  uint8_t *base;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = (uint8_t *)g_scratch_ptr[dev];
  }
  if (!base) {  // initialised before this device was current: first call here
    GG_TRY(bfs_noalloc_setup(dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = (uint8_t *)g_scratch_ptr[dev];
  }
  b->ctl = (BfsCtl *)(base + kNaCtl);
  b->sets = (uint64_t *)(base + kNaSets);
  b->qc = (Counters *)(base + kNaQc);
  b->hf = (BfsHostFlag *)(base + kNaFlag);
  b->exam = options().count_examined ? (unsigned long long *)(base + kNaExam) : nullptr;
  b->lv = (uint32_t *)(base + kNaLv);
  b->lv_cap = (uint32_t)std::min<int64_t>(std::max(cap, 1), (int64_t)V + 1);
  b->visited = (uint64_t *)(base + kNaVisited);
  b->f[0] = (uint64_t *)(base + kNaF0);
  b->f[1] = (uint64_t *)(base + kNaF1);
  b->q[0] = (uint32_t *)(base + kNaQ0);
  b->q[1] = (uint32_t *)(base + kNaQ1);
  b->q[2] = (uint32_t *)(base + kNaQ2);
  b->words = ((size_t)V >> 6) + 1;
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_bfs(gg_graph *g, uint32_t root, uint32_t *d_parent, uint32_t *d_depth,
                      uint32_t *h_frontier, int cap, gg_run_stats *stats) {
  GG_TRY(check_graph(g));
  if (root >= g->vertices) return fail(GG_EINVAL, "root out of range");
  if (!d_parent) return fail(GG_EINVAL, "null output");
  Work *w;
  GG_TRY(work_get(g, &w));
  const uint32_t V = g->vertices;
  const uint32_t lv_need = (uint32_t)std::min<int64_t>(std::max(cap, 1), (int64_t)V + 1);
  BfsState *s;
  GG_TRY(bfs_state(w, lv_need, &s));
  const Options &opt = options();
  BfsBufs b{s->ctl, s->sets, s->qc, s->lv, s->lv_cap, s->h_flag,
            opt.count_examined ? s->exam : nullptr, w->visited, {w->f[0].bits, w->f[1].bits},
            {w->f[0].q[0], w->f[0].q[1], w->f[0].q[2]}, w->words};
  const BfsPlan p = bfs_plan(g, root, d_parent, d_depth, b.words);
  hipStream_t st = stream();
  volatile BfsHostFlag *hf = s->h_flag;
  hf->decided = 0;
  hf->done = 0;
  hf->final_copy = s->h_ctl;  // the device leaves the final control block there
  const auto t0 = std::chrono::steady_clock::now();
  bfs_issue_start(g, p, b, st);
  // iterations queued ahead of the device: before issuing iteration n, wait
  // until iteration n - ahead has been decided (iteration n - ahead - 1 has
  // finished).  Short levels take less device time than the host needs to
  // see a decision and issue two launches; each iteration queued past the
  // end runs as two launches that return at once.
  const uint32_t ahead = (uint32_t)std::max<int64_t>(1, opt.bfs_ahead);
  for (uint32_t n = 0;; n++) {
    if (n >= ahead) {
      const uint32_t need = n - ahead + 1;
      uint64_t spins = 0;
      while (!hf->done && hf->decided < need) {
        if ((++spins & 0xFFF) == 0 && hipStreamQuery(st) == hipSuccess && !hf->done &&
            hf->decided < need)
          return fail(GG_EHIP, "gg_bfs: device loop stopped without a decision");
      }
      if (hf->done) break;
    }
    if (n > V + 2) return fail(GG_EHIP, "gg_bfs: traversal did not terminate");
    bfs_issue_iteration(g, p, b, n, st);
    GG_HIP(hipGetLastError());
  }
  // done was published (release) after the device copied the control block
  // into h_ctl: no copy and no wait for the launches still queued, which
  // return at once
  std::atomic_thread_fence(std::memory_order_acquire);
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const BfsCtl &c = *s->h_ctl;
  if (h_frontier && cap > 0) {
    const uint32_t levels = std::min<uint32_t>(c.iter, (uint32_t)cap);
    if (levels) {
      GG_HIP(hipMemcpyAsync(h_frontier, s->lv, (size_t)levels * 4, hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
    }
  }
  bfs_stats(c, ms, stats);
  return GG_OK;
}

// No host flag to poll: the host issues iterations in rounds and reads the
// control block's `done` word back after each round (stream order: a copy
// into host memory it owns, no pinned buffer).  Iterations issued past the
// end return at once.
extern "C" int gg_bfs_noalloc(gg_graph *g, uint32_t root, uint32_t *d_parent, uint32_t *d_depth,
                              uint32_t *h_frontier, int cap, gg_run_stats *stats) {
  GG_TRY(check_graph(g));
  if (root >= g->vertices) return fail(GG_EINVAL, "root out of range");
  if (!d_parent) return fail(GG_EINVAL, "null output");
  const uint32_t V = g->vertices;
  if (V > kNoAllocMaxVertices)
    return fail(GG_EINVAL, "gg_bfs_noalloc: the graph has more than 2^20 vertices, the most its "
                           "scratch array holds (use gg_bfs)");
  if (cap > (int)kNoAllocMaxLevels)
    return fail(GG_EINVAL, "gg_bfs_noalloc: cap above 65536 level counts, the most its scratch "
                           "array records");
  BfsBufs b;
  GG_TRY(scratch_bufs(V, cap, &b));
  const BfsPlan p = bfs_plan(g, root, d_parent, d_depth, b.words);
  hipStream_t st = stream();
  const auto t0 = std::chrono::steady_clock::now();
  bfs_issue_start(g, p, b, st);
  constexpr uint32_t kRound = 4;  // iterations per round
  uint32_t done = 0;
  for (uint32_t n = 0; !done; n += kRound) {
    if (n > V + 2) return fail(GG_EHIP, "gg_bfs_noalloc: traversal did not terminate");
    for (uint32_t k = 0; k < kRound; k++) bfs_issue_iteration(g, p, b, n + k, st, false);
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(&done, &b.ctl->done, sizeof(done), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
  }
  BfsCtl c;
  GG_HIP(hipMemcpyAsync(&c, b.ctl, sizeof(BfsCtl), hipMemcpyDeviceToHost, st));
  GG_HIP(hipStreamSynchronize(st));
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (h_frontier && cap > 0) {
    const uint32_t levels = std::min<uint32_t>(c.iter, (uint32_t)cap);
    if (levels) {
      GG_HIP(hipMemcpyAsync(h_frontier, b.lv, (size_t)levels * 4, hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
    }
  }
  bfs_stats(c, ms, stats);
  return GG_OK;
}

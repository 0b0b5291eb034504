#!/usr/bin/env python3
"""Benchmark: BFS GTEPS on RMAT-24 (scale 24, edge factor 16, permuted ids),
16 seeded non-isolated roots per step, plain (single-source) method on one GPU.

One step = BFS from the 16 roots, inputs (CSR/CSC) resident in HBM.
value = sum over roots of m_reach / BFS time (Graph500-style GTEPS), where
m_reach = sum of out-degrees of the reached vertices.

Also reported on the same graph (not part of `value`): PageRank (20
iterations, FP64) GTEPS and HBM roofline fraction, and CC on the symmetrised
graph.  The CPU baseline is the oracle (C/OpenMP restatement of the
reference's engine) on a bounded sample of the same workload, timed on this
host's cores, with its depths checked against the GPU's.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s
# The BFS graph's relabeling: the 2^19 highest-out-degree vertices first, in
# degree order, the others in file order (the build option relabel_hubs).
# Consecutive internal ids then have ascending file ids, so the parents that
# a wave writes (indexed by file id) share lines; with every vertex in degree
# order they land on random lines (1.13x the BFS rate with the hybrid order
# on RMAT-24; PageRank, CC and batched BC prefer the full order, which their
# legs keep).
BFS_RELABEL_HUBS = 1 << 19
# The BFS leg's engine options, from a sweep on this graph (RMAT-24, the 16
# bench roots): push while the frontier's out-arcs are below |E|/3200 or
# below 1/32 of the unvisited vertices' in-arcs, small-degree push bin up to
# 32 arcs.  Divisor 800-6400 with alpha 24-48 lie within 1% of each other;
# alpha 16 (0.454 ms per BFS) and below push too long for the current
# bottom-up step, alpha 64 pulls too early.  Batched BFS and the plain BC run
# faster with the defaults (200 / 24 / 64), so only this leg uses them.
BFS_OPTIONS = {"switch_divisor": 3200, "bfs_alpha": 32, "push_small_max": 32}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pick_roots(out_degree, k, seed):
    cand = np.nonzero(out_degree)[0]
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(cand, size=min(k, len(cand)), replace=False)]


def bfs_model_bytes(st, V, no_in=0, found=8):
    """Algorithmic bytes of each process_edges call of one BFS:
    push: 20F + 8X + wN  (queue entry, offset pair, col, state probe, writes)
    pull: 16U + 4X + wN + 3V/8  (offsets of unvisited, col until exit, writes,
          visited / frontier-in / frontier-out bitmaps).
    w = `found` bytes written per vertex found: 4 for the parent alone (the
    reference's output), 8 with the depth array as well.
    U counts the unvisited vertices that have in-arcs: the `no_in` vertices
    without any (other than a root among them) can never be found, and the
    bottom-up step does not read their lists."""
    push = pull = 0
    for i, mode in enumerate(st["modes"]):
        F, N, X = st["it_active"][i], st["it_new"][i], st["it_examined"][i]
        U = max(0, st["it_unvisited"][i] - no_in)
        if mode == 0:
            push += 20 * F + 8 * X + found * N
        else:
            pull += 16 * U + 4 * X + found * N + 3 * V / 8
    return push, pull


def depths_from_parents(parent, root):
    """Level of every vertex in the tree the parents form (0xFFFFFFFF where
    parent == V); fails unless every reached vertex leads up to the root."""
    V = len(parent)
    depth = np.full(V, 0xFFFFFFFF, dtype=np.uint32)
    node = np.nonzero(parent < V)[0]
    cur = node.copy()
    d = np.zeros(len(node), dtype=np.int64)
    for _ in range(1024):
        up = np.nonzero(cur != root)[0]
        if not len(up):
            break
        cur[up] = parent[cur[up]]
        d[up] += 1
    assert (cur == root).all(), "parents do not form a tree rooted at the root"
    depth[node] = d
    return depth


def batch_extra(gg, seed, reorder):
    """RMAT-22 edge factor 32, 64 roots: one batched traversal against 64
    sequential single-root calls on the same GPU (BFS and BC), results checked
    identical (depths) / within 1e-9 (BC) for sampled roots."""
    scale, ef = 22, 32
    V = 1 << scale
    g = gg.Graph.rmat(scale, ef, seed=seed, permute=True, reorder=reorder)
    roots = pick_roots(g.out_degree(), 64, seed + 2)
    parent, depth = gg.DeviceArray(V, np.uint32), gg.DeviceArray(V, np.uint32)
    rows = gg.DeviceArray(V * 64, np.uint32)
    g.bfs_into(roots[0], parent, depth)
    gg.bfs_batch_into(g, roots, rows)
    gg.synchronize()
    t = time.perf_counter()
    for r in roots:
        g.bfs_into(r, parent, depth)
    gg.synchronize()
    seq = time.perf_counter() - t
    t = time.perf_counter()
    st = gg.bfs_batch_into(g, roots, rows)
    gg.synchronize()
    bat = time.perf_counter() - t
    host_rows = rows.numpy().reshape(V, 64)
    same = True
    for i in (0, 21, 63):
        g.bfs_into(roots[i], parent, depth)
        same = same and bool(np.array_equal(host_rows[:, i], depth.numpy()))
    out = {"workload": f"RMAT-{scale} edge factor {ef}, 64 roots", "bfs_sequential_ms":
           round(seq * 1e3, 2), "bfs_batched_ms": round(bat * 1e3, 2),
           "bfs_speedup": round(seq / bat, 2), "bfs_batched_queries_per_s": round(64 / bat, 1),
           "bfs_batched_gteps": round(st.m_reach / bat / 1e9, 2), "bfs_depths_identical": same}
    del host_rows
    rows.free()
    parent.free()
    depth.free()
    dep, inv = gg.DeviceArray(V, np.float64), gg.DeviceArray(V, np.float64)
    bdep, binv = gg.DeviceArray(V * 64, np.float64), gg.DeviceArray(V * 64, np.float64)
    g.bc_into(roots[0], dep, inv)
    gg.bc_batch_into(g, roots, bdep, binv)
    gg.synchronize()
    t = time.perf_counter()
    for r in roots:
        g.bc_into(r, dep, inv)
    gg.synchronize()
    seq = time.perf_counter() - t
    t = time.perf_counter()
    gg.bc_batch_into(g, roots, bdep, binv)
    gg.synchronize()
    bat = time.perf_counter() - t
    g.bc_into(roots[63], dep, inv)
    a = bdep.numpy().reshape(V, 64)[:, 63]
    b = dep.numpy()
    fin = np.isfinite(b)
    ok = bool(np.array_equal(np.isnan(a), np.isnan(b)) and
              np.all(np.abs(a[fin] - b[fin]) <= 1e-9 * np.maximum(np.abs(b[fin]), 1e-300)))
    out.update({"bc_sequential_ms": round(seq * 1e3, 2), "bc_batched_ms": round(bat * 1e3, 2),
                "bc_speedup": round(seq / bat, 2), "bc_batched_within_1e-9": ok})
    for x in (dep, inv, bdep, binv):
        x.free()
    del g
    return out


def serve_extra(gg, seed, reorder, seconds, rate=2000.0, workers=4, budget=200e9,
                bfs_workers=None):
    """Serving: a Poisson stream at `rate` queries/s, BFS 50% / SSSP 30% /
    BC 20%, uniformly random roots, over 4 resident weighted RMAT graphs
    (scales 18, 20, 22, 24, edge factor 16) admitted by a 200 GB budget; the
    server coalesces same-graph, same-kind queries into batches of up to 64
    on `workers` streams.  Latency = scheduled arrival to completion."""
    scales = (18, 20, 22, 24)
    graphs = [gg.Graph.rmat(s, 16, seed=seed + s, permute=True, weighted=True, reorder=reorder)
              for s in scales]
    rng = np.random.default_rng(seed)

    def trace(n):
        gi = rng.integers(0, len(scales), n)
        kind = rng.choice(3, size=n, p=[0.5, 0.3, 0.2])
        root = np.array([rng.integers(0, 1 << scales[i]) for i in gi], dtype=np.uint32)
        arrival = np.cumsum(rng.exponential(1.0 / rate, n))
        return arrival, gi, kind, root

    bw = gg.get_option("serve_bfs_workers") if bfs_workers is None else bfs_workers
    with gg.options(serve_bfs_workers=bw), gg.Server(int(budget), workers=workers) as s:
        ids = [s.add_graph(g) for g in graphs]
        # warm-up: every worker allocates its buffers for each (graph, kind)
        arr, gi, kind, root = trace(400)
        s.replay(arr * 0, [ids[i] for i in gi], kind, root)
        before = s.info()
        n = int(rate * seconds)
        arr, gi, kind, root = trace(n)
        lat, wall = s.replay(arr, [ids[i] for i in gi], kind, root)
        info = s.info()
    batches = info["batches"] - before["batches"]
    out = {"workload": f"Poisson {rate:.0f} q/s for {seconds} s, BFS/SSSP/BC 50/30/20, "
                       f"weighted RMAT scales {list(scales)} edge factor 16",
           "queries": n, "workers": workers, "bfs_workers": min(bw, workers - 1),
           "budget_GB": budget / 1e9,
           "queries_per_s": round(n / wall, 1), "offered_per_s": rate,
           "p50_ms": round(float(np.percentile(lat, 50)), 2),
           "p99_ms": round(float(np.percentile(lat, 99)), 2),
           "mean_ms": round(float(lat.mean()), 2),
           "batches": batches, "mean_batch": round(n / max(batches, 1), 2),
           "max_batch": info["max_batch"], "used_GB": round(info["used_bytes"] / 1e9, 2)}
    for k, name in enumerate(("bfs", "sssp", "bc")):
        m = kind == k
        out[f"{name}_p50_ms"] = round(float(np.percentile(lat[m], 50)), 2)
        out[f"{name}_p99_ms"] = round(float(np.percentile(lat[m], 99)), 2)
    # a backlog shows as latency growing over the trace: both halves by arrival
    half = arr < arr[-1] / 2
    for tag, m in (("first_half", half), ("second_half", ~half)):
        out[f"{tag}_p50_ms"] = round(float(np.percentile(lat[m], 50)), 2)
        out[f"{tag}_p99_ms"] = round(float(np.percentile(lat[m], 99)), 2)
    # steady-state throughput: completions per second over the middle 80% of
    # the trace (n / wall also counts the drain of the last batches)
    done_s = arr + lat / 1e3
    t0, t1 = 0.1 * arr[-1], 0.9 * arr[-1]
    out["completion_rate_per_s"] = round(float(((done_s >= t0) & (done_s < t1)).sum() / (t1 - t0)), 1)
    out["sustained"] = bool(out["completion_rate_per_s"] >= 0.99 * rate and
                            out["second_half_p99_ms"] <= 1.5 * out["first_half_p99_ms"])
    del graphs
    return out


def pagerank_extra(gg, g, V, E, iterations=20, runs=3):
    """PageRank (FP64, d = 0.85) on the bench graph: the first call (which
    builds the bin plan of a relabelled graph), wall time of `runs` calls,
    one call with per-launch events for the kernel time, and the plain pull
    (pr_hot_sources = 0) beside it with the largest relative difference."""
    pr = gg.DeviceArray(V, np.float64)
    t = time.perf_counter()
    g.pagerank_into(pr, iterations)
    gg.synchronize()
    first_s = time.perf_counter() - t
    t = time.perf_counter()
    for _ in range(runs):
        g.pagerank_into(pr, iterations)
    gg.synchronize()
    pr_s = (time.perf_counter() - t) / runs
    names = ("pr_pull", "pr_heavy", "pr_emit", "pr_bins", "pr_split_finish")
    gg.prof_reset()
    gg.prof_enable(True)
    g.pagerank_into(pr, iterations)
    gg.prof_enable(False)
    prk = {k: gg.prof_read(k) for k in names}
    pr_kernel_s = sum(v[0] for v in prk.values()) / 1e3
    ranks = pr.numpy()
    hot = gg.get_option("pr_hot_sources")
    with gg.options(pr_hot_sources=0):
        g.pagerank_into(pr, iterations)
        gg.synchronize()
        t = time.perf_counter()
        g.pagerank_into(pr, iterations)
        gg.synchronize()
        pull_s = time.perf_counter() - t
    pulled = pr.numpy()
    pr.free()
    rel = float(np.max(np.abs(ranks - pulled) / np.abs(pulled)))
    b_iter = 12 * E + 28 * V
    return {
        "timed_region": "ms: gg_pagerank calls after the first (the plan built), output "
                        "preallocated; first_call_ms: the first call, which builds the bin plan "
                        "(the reference-equivalent first compute(), pagerank.cpp:27-105)",
        "iterations": iterations,
        "ms": round(pr_s * 1e3, 3),
        "gteps": round(iterations * E / pr_s / 1e9, 2),
        "ms_per_iter": round(pr_s * 1e3 / iterations, 4),
        "model_bytes_per_iter": b_iter,
        "achieved_GBs": round(iterations * b_iter / pr_kernel_s / 1e9, 1),
        "frac": round(iterations * b_iter / pr_kernel_s / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms_per_iter": {k: round(v[0] / iterations, 4) for k, v in prk.items() if v[1]},
        "hot_sources": hot,
        "bin_bits": gg.get_option("pr_bin_bits"),
        "range_bits": gg.get_option("pr_range_bits"),
        "first_call_ms": round(first_s * 1e3, 1),
        "plan_build_ms": round((first_s - pr_s) * 1e3, 1),
        "pull_ms_per_iter": round(pull_s * 1e3 / iterations, 4),
        "max_rel_vs_pull": rel,
    }


def cc_extra(gg, scale, ef, seed, reorder):
    """CC (cc.cpp) on the bench topology made undirected by reverse-edge
    doubling: one warm call (builds the bin plan on a relabelled graph), then
    one timed call."""
    V, E = 1 << scale, (1 << scale) * ef
    gu = gg.Graph.rmat(scale, ef, seed=seed, permute=True, undirected=True, reorder=reorder)
    lab = gg.DeviceArray(V, np.uint32)
    gu.cc_into(lab)
    t = time.perf_counter()
    st = gu.cc_into(lab)
    gg.synchronize()
    cc_s = time.perf_counter() - t
    lab.free()
    gu.close()
    return {"timed_region": "one gg_cc call after a warm call (which builds the bin plan), "
                            "labels preallocated",
            "ms": round(cc_s * 1e3, 3), "iterations": st.iterations,
            "push_iterations": st.push_iterations, "gteps_2E": round(2 * E / cc_s / 1e9, 2)}


def rmat26_extra(gg, seed, scale=26, ef=16, n_roots=8, iterations=20):
    """The scale-26 case: SSSP (plain method) from 8 seeded non-isolated
    roots on the weighted graph (float32 weights in (0,1]), then PageRank
    (FP64, 20 iterations) on the same topology, with the plain pull beside
    it; both graphs degree-relabelled (results in original ids).  SSSP
    GTEPS = sum of m_reach / sum of times."""
    V = 1 << scale
    t = time.perf_counter()
    g = gg.Graph.rmat(scale, ef, seed=seed, permute=True, weighted=True, reorder=True)
    gg.synchronize()
    build_s = time.perf_counter() - t
    roots = pick_roots(g.out_degree(), n_roots, seed)
    dist = gg.DeviceArray(V, np.float32)
    g.sssp_into(roots[0], dist)
    times, iters, m = [], [], 0
    for r in roots:
        t = time.perf_counter()
        st = g.sssp_into(r, dist)
        gg.synchronize()
        times.append(time.perf_counter() - t)
        iters.append(st.iterations)
        m += st.m_reach
    dist.free()
    E = g.edges
    g.close()
    out = {"workload": f"RMAT scale {scale} edge factor {ef}, weighted (SSSP) and unweighted "
                       f"(PageRank), both relabelled, seed {seed}",
           "vertices": V, "edges": E, "build_s": round(build_s, 2),
           "sssp_roots": len(roots), "sssp_ms_mean": round(1e3 * float(np.mean(times)), 1),
           "sssp_ms_max": round(1e3 * float(np.max(times)), 1),
           "sssp_iterations": iters, "sssp_gteps": round(m / sum(times) / 1e9, 2)}
    # PageRank measured as on the RMAT-24 graph (pagerank_extra): wall time of
    # 3 calls after the plan-building first one, the roofline fraction from
    # the kernels' event time; the wall-time fraction beside it
    h = gg.Graph.rmat(scale, ef, seed=seed, permute=True, reorder=True)
    r = pagerank_extra(gg, h, V, E, iterations)
    h.close()
    b_iter = 12 * E + 28 * V
    out.update({"pagerank_ms_per_iter": r["ms_per_iter"],
                "pagerank_first_call_ms": r["first_call_ms"],
                "pagerank_gteps": r["gteps"],
                "pagerank_frac": r["frac"],
                "pagerank_frac_wall": round(b_iter / (r["ms_per_iter"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                "pagerank_kernel_ms_per_iter": r["kernel_ms_per_iter"],
                "pagerank_pull_ms_per_iter": r["pull_ms_per_iter"],
                "pagerank_max_rel_vs_pull": r["max_rel_vs_pull"]})
    return out


def facade_extra(gg, scale, ef, seed, roots, iterations=20):
    """The facade toolkits (geminigraph_amd/toolkits: the reference's BFS and
    PageRank compute() over Graph<Empty>::process_edges with device
    callbacks) run as the reference's programs are: on a binary edge-list
    file, 1 + 5 compute() calls, each printing its exec_time.  Beside them,
    the C-ABI calls on the same graph (file ids, no relabeling, as the facade
    keeps them), timed the same way."""
    import re
    import subprocess
    import tempfile
    V, E = 1 << scale, (1 << scale) * ef
    tk = os.path.join(ROOT, "geminigraph_amd", "toolkits")
    if not all(os.path.exists(os.path.join(tk, n)) for n in ("bfs", "pagerank")):
        return {"error": "facade toolkits not built (make -C geminigraph_amd toolkits)"}
    buf = gg.rmat_device(scale, E, seed, permute=True)
    host = np.empty(E * 8, dtype=np.uint8)
    gg.lib().gg_memcpy_d2h(host.ctypes.data, gg.ctypes.c_void_p(buf.ptr), E * 8)
    g = gg.Graph.build_device(buf, E, V, hub_first=True)  # the facade's own load flags
    buf.free()
    d = tempfile.mkdtemp(prefix="ggfacade", dir="/tmp")
    path = os.path.join(d, "rmat.bin")
    host.tofile(path)
    del host

    def run(name, arg):
        p = subprocess.run([os.path.join(tk, name), path, str(V), str(arg)], capture_output=True,
                           text=True, timeout=300, env=dict(os.environ, GG_RUNS="6"))
        if p.returncode != 0:
            raise RuntimeError(f"{name}: {p.stderr[-500:]}")
        t = [float(x) for x in re.findall(r"exec_time=([0-9.eE+-]+)\(s\)", p.stdout)]
        return t[1:]  # the first compute() is the warm-up, as pagerank.cpp:144-147

    out = {"workload": f"RMAT-{scale} edge factor {ef}, permuted ids, not relabelled (the facade "
                       f"keeps file ids; lists hub-first), {len(roots)} of the bench roots",
           "timed_region": "exec_time of compute() as the toolkits print it (allocation of the "
                           "vertex arrays and subsets included, load excluded), median of the "
                           "5 runs after the first; C-ABI: the same calls with per-call output "
                           "allocation"}
    try:
        bfs_t = [float(np.median(run("bfs", r))) for r in roots]
        pr_t = float(np.median(run("pagerank", iterations)))
    finally:
        os.remove(path)
        os.rmdir(d)
    cabi_bfs = []
    for r in roots:
        ts = []
        for _ in range(6):
            t = time.perf_counter()
            par = gg.DeviceArray(V, np.uint32)
            g.bfs_into(r, par)
            gg.synchronize()
            ts.append(time.perf_counter() - t)
            par.free()
        cabi_bfs.append(float(np.median(ts[1:])))
    ts = []
    for _ in range(6):
        t = time.perf_counter()
        pr = gg.DeviceArray(V, np.float64)
        g.pagerank_into(pr, iterations)
        gg.synchronize()
        ts.append(time.perf_counter() - t)
        pr.free()
    cabi_pr = float(np.median(ts[1:]))
    g.close()
    out.update({
        "bfs_facade_ms": round(1e3 * float(np.mean(bfs_t)), 3),
        "bfs_cabi_ms": round(1e3 * float(np.mean(cabi_bfs)), 3),
        "bfs_facade_over_cabi": round(float(np.mean(bfs_t)) / float(np.mean(cabi_bfs)), 2),
        "pagerank_facade_ms": round(1e3 * pr_t, 2),
        "pagerank_cabi_ms": round(1e3 * cabi_pr, 2),
        "pagerank_facade_over_cabi": round(pr_t / cabi_pr, 2)})
    return out


def cpu_threads():
    threads = os.environ.get("OMP_NUM_THREADS")
    if not threads:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(int(threads), 64))


CPU_KIND = ("the oracle: this build's C/OpenMP restatement of the reference's engine and "
            "toolkits (not the reference binary, which needs MPI), timed as pagerank.cpp:144-147 "
            "times compute(): one untimed run, then timed runs, graph load excluded")


def timed_runs(fn, runs):
    """One untimed run (the reference's first compute()), then `runs` timed
    ones; returns (median, min) seconds."""
    import oracle
    fn()
    ts = []
    for _ in range(runs):
        t = oracle.wall_time()
        fn()
        ts.append(oracle.wall_time() - t)
    return float(np.median(ts)), float(np.min(ts))


def cpu_extras(scale, ef, seed, gpu):
    """CPU columns beside the GPU figures of `extras`, on the same graphs:
    PageRank on RMAT-16 (the reference's own CPU config) and RMAT-24, CC on
    the symmetrised RMAT-24 graph, SSSP on the weighted RMAT-24 graph.
    Bounded samples keep the host work to about a minute."""
    import oracle
    threads = cpu_threads()
    oracle.set_threads(threads)
    out = {"cores": threads, "kind": "port", "what": CPU_KIND}
    # PageRank, RMAT-16 edge factor 16, 20 iterations: the whole config, 1 + 5 runs
    e16 = oracle.rmat_edges(16, 16, seed=seed, permute=True)
    o16 = oracle.Graph(e16, 1 << 16)
    med, mn = timed_runs(lambda: o16.pagerank(20), 5)
    out["pagerank_rmat16"] = {"ms_median": round(med * 1e3, 2), "ms_min": round(mn * 1e3, 2),
                              "gteps": round(20 * len(e16) / med / 1e9, 3),
                              "sample": "20 iterations, 1 untimed + 5 timed runs"}
    del o16, e16
    V = 1 << scale
    e = oracle.rmat_edges(scale, ef, seed=seed, permute=True)
    o = oracle.Graph(e, V)
    it = 4
    med, mn = timed_runs(lambda: o.pagerank(it), 5)
    out["pagerank"] = {"ms_per_iter_median": round(med * 1e3 / it, 2),
                       "ms_per_iter_min": round(mn * 1e3 / it, 2),
                       "gteps": round(it * len(e) / med / 1e9, 3),
                       "sample": f"{it} iterations per run (the per-iteration cost of the 20-"
                                 f"iteration run), 1 untimed + 5 timed runs, RMAT-{scale}"}
    del o
    ou = oracle.Graph(e, V, undirected=True)
    med, mn = timed_runs(lambda: ou.cc(), 3)
    out["cc"] = {"ms_median": round(med * 1e3, 1), "ms_min": round(mn * 1e3, 1),
                 "gteps_2E": round(2 * len(e) / med / 1e9, 3),
                 "sample": "the whole CC, 1 untimed + 3 timed runs"}
    del ou, e
    ew = oracle.rmat_edges(scale, ef, seed=seed + 3, permute=True, weighted=True)
    ow = oracle.Graph(ew, V)
    del ew
    roots = gpu.get("sssp_roots", [])[:3]
    od = ow.out_degree().astype(np.int64)
    ts, m = [], 0
    ow.sssp(roots[0])
    for r in roots:
        t = oracle.wall_time()
        dist, _ = ow.sssp(r)
        ts.append(oracle.wall_time() - t)
        m += int(od[dist < 1e9].sum())
    out["sssp"] = {"ms_mean": round(1e3 * float(np.mean(ts)), 1),
                   "gteps": round(m / sum(ts) / 1e9, 3),
                   "sample": f"{len(roots)} of the GPU leg's roots after one untimed run, "
                             f"weighted RMAT-{scale}"}
    del ow
    return out


def sssp24_extra(gg, scale, ef, seed, n_roots=8):
    """SSSP (plain method) from 8 seeded roots on the weighted RMAT-24 graph
    (relabelled, the weighted bin plan): the GPU figure the CPU column of the
    same graph stands beside."""
    V = 1 << scale
    g = gg.Graph.rmat(scale, ef, seed=seed, permute=True, weighted=True, reorder=True)
    roots = pick_roots(g.out_degree(), n_roots, seed)
    dist = gg.DeviceArray(V, np.float32)
    g.sssp_into(roots[0], dist)
    gg.synchronize()
    times, m = [], 0
    for r in roots:
        t = time.perf_counter()
        st = g.sssp_into(r, dist)
        gg.synchronize()
        times.append(time.perf_counter() - t)
        m += st.m_reach
    dist.free()
    g.close()
    return {"workload": f"SSSP from {n_roots} roots, weighted RMAT-{scale} edge factor {ef}, "
                        f"seed {seed}, relabelled",
            "ms_mean": round(1e3 * float(np.mean(times)), 2), "gteps": round(m / sum(times) / 1e9, 2),
            "sssp_roots": roots, "timed_region": "each gg_sssp call, outputs preallocated"}


def pagerank16_extra(gg, seed, iterations=20, runs=5):
    """PageRank on the reference's CPU config (RMAT-16 edge factor 16, 20
    iterations): the GPU's figure for the config itself, 1 + 5 runs as
    pagerank.cpp:144-147, each with its own output allocation."""
    g = gg.Graph.rmat(16, 16, seed=seed, permute=True, reorder=True)
    ts = []
    for i in range(runs + 1):
        t = time.perf_counter()
        pr = gg.DeviceArray(1 << 16, np.float64)
        g.pagerank_into(pr, iterations)
        gg.synchronize()
        el = time.perf_counter() - t
        pr.free()
        if i:
            ts.append(el)
    g.close()
    med = float(np.median(ts))
    return {"ms_median": round(med * 1e3, 3), "ms_min": round(min(ts) * 1e3, 3),
            "gteps": round(iterations * (1 << 20) / med / 1e9, 2),
            "timed_region": "compute() as pagerank.cpp:27-105: output allocation, the plan build "
                            "of the first call (run 0, untimed), 20 iterations"}


def cpu_baseline(scale, ef, seed, roots, gpu_depths, sample_roots, repeats):
    import oracle
    threads = cpu_threads()
    oracle.set_threads(threads)
    t0 = time.time()
    e = oracle.rmat_edges(scale, ef, seed=seed, permute=True)
    og = oracle.Graph(e, 1 << scale)
    del e
    log(f"cpu oracle graph ready in {time.time() - t0:.1f}s ({threads} threads)")
    od = og.out_degree().astype(np.int64)
    sample = roots[:sample_roots]
    # one untimed pass (warm-up, as pagerank.cpp:144-147, and the depth
    # parity check), then `repeats` timed passes over the sampled roots
    parity = True
    for i, r in enumerate(sample):
        _, depth, _ = og.bfs(r)
        if gpu_depths is not None and i < len(gpu_depths):
            parity = parity and bool(np.array_equal(depth, gpu_depths[i]))
    total_m, total_t = 0, 0.0
    for _ in range(repeats):
        for r in sample:
            t = oracle.wall_time()
            _, depth, _ = og.bfs(r)
            total_t += oracle.wall_time() - t
            total_m += int(od[depth != 0xFFFFFFFF].sum())
    log(f"cpu bfs: {repeats} x {len(sample)} roots in {total_t:.2f}s")
    return {
        "value": total_m / total_t / 1e9,
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"BFS from the first {len(sample)} of the {len(roots)} roots on the same "
                  f"RMAT-{scale} graph, {repeats} timed passes after one untimed pass "
                  f"({total_t:.1f} s of BFS), compute() only (graph build excluded, as the "
                  f"reference's exec_time excludes load)",
    }, parity


# the two launches of every BFS iteration: the step (decision, then the
# bottom-up step or the frontier binning), and the push (push iterations) or
# the heavy in-list chunks (pull iterations)
STEP_KERNELS = ("gg::bfs_step_kernel", "gg::bfs_push_heavy_kernel")


def kernel_name(row):
    """A profiler row's kernel name without return type, template arguments
    or parameters (gg::bfs_step_kernel<false, true>(...) -> gg::bfs_step_kernel)."""
    return row["Kernel_Name"].split("(")[0].replace("void ", "").split("<")[0]
COPY_KERNEL = "gg::gg_copy_kernel"
SEED_ARCS = 1024  # gg_bfs: a first push of at most this many arcs runs in the seed kernel
COPY_BYTES = 256 << 20


def bfs_graph(gg, args):
    """The bench's BFS graph: RMAT, permuted ids, relabelled with the hybrid
    order (relabel_hubs = args.relabel_hubs) unless --no-reorder."""
    with gg.options(relabel_hubs=args.relabel_hubs):
        return gg.Graph.rmat(args.scale, args.edge_factor, seed=args.seed, permute=True,
                             reorder=not args.no_reorder)


def pmc_child(args):
    """Workload of the counter passes: one step of the bench (the same graph,
    roots and options), then one calibration copy of COPY_BYTES."""
    import geminigraph_amd as gg
    gg.init(0)
    for k, v in BFS_OPTIONS.items():
        gg.set_option(k, v)
    g = bfs_graph(gg, args)
    V = 1 << args.scale
    roots = pick_roots(g.out_degree(), args.roots, args.seed + 1)
    parent = gg.DeviceArray(V, np.uint32)
    for r in roots:
        g.bfs_into(r, parent)
    src, dst = gg.DeviceArray(COPY_BYTES, np.uint8), gg.DeviceArray(COPY_BYTES, np.uint8)
    gg.device_copy(dst, src, COPY_BYTES)
    gg.synchronize()


def measure_traffic(args):
    """HBM bytes per iteration of the BFS step kernels, from two rocprofv3
    counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each
    scaled by what the same counter reads for the plain copy kernel moving
    COPY_BYTES each way.  Returns (bytes per launch or None, detail)."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, {"error": "rocprofv3 not found"}
    tmp = tempfile.mkdtemp(prefix="ggpmc", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    child = [sys.executable, os.path.abspath(__file__), "--pmc-child", "--scale", str(args.scale),
             "--edge-factor", str(args.edge_factor), "--roots", str(args.roots),
             "--seed", str(args.seed)] + (["--no-reorder"] if args.no_reorder else [])
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = os.path.join(tmp, counter)
        cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "-d", out, "-o", "run",
               "-f", "csv", "--"] + child
        log(f"counter pass: {counter}")
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
        path = None
        for root, _, files in os.walk(out):
            for f in files:
                if f.endswith("counter_collection.csv"):
                    path = os.path.join(root, f)
        if path is None:
            return None, {"error": f"{counter} pass produced no counters (exit {r.returncode})"}
        pull, pull_n, copy = 0.0, 0, None
        for row in csv.DictReader(open(path)):
            name = kernel_name(row)
            v = float(row["Counter_Value"])
            if name in STEP_KERNELS:
                pull += v
                pull_n += name == STEP_KERNELS[0]
            elif name == COPY_KERNEL:
                copy = v
        if not copy or not pull_n:
            return None, {"error": f"{counter}: kernels not found"}
        factor = COPY_BYTES / (copy * 1024.0)
        per[counter] = {"kib_per_launch": pull / pull_n, "factor": factor,
                        "bytes_per_launch": pull / pull_n * 1024.0 * factor}
    traffic = per["FETCH_SIZE"]["bytes_per_launch"] + per["WRITE_SIZE"]["bytes_per_launch"]
    return traffic, {"fetch": per["FETCH_SIZE"], "write": per["WRITE_SIZE"],
                     "launches": pull_n, "workload": "one bench step (same graph and roots)"}


LEVEL_PASSES = (
    "TCC_EA0_RDREQ_sum,TCC_EA0_RDREQ_32B_sum,TCC_EA0_WRREQ_sum,TCC_EA0_WRREQ_64B_sum",
    "TCC_EA0_RDREQ_64B_sum,TCC_EA0_RDREQ_128B_sum,TCC_HIT_sum,TCC_MISS_sum",
    "TCC_READ_sum,TCC_WRITE_sum,TCC_ATOMIC_sum,TCC_REQ_sum",
)


def level_child(args):
    """Workload of the per-level passes: the bench graph, one untimed BFS,
    BFS from the first bench root with the edge counters on (its per-level
    records, printed as JSON), then the same BFS as the bench runs it (no
    counting), the one the passes read (the last BFS of the run)."""
    import geminigraph_amd as gg
    gg.init(0)
    for k, v in BFS_OPTIONS.items():
        gg.set_option(k, v)
    g = bfs_graph(gg, args)
    roots = pick_roots(g.out_degree(), args.roots, args.seed + 1)
    parent = gg.DeviceArray(1 << args.scale, np.uint32)
    g.bfs_into(roots[1], parent)
    with gg.options(count_examined=1):
        st = g.bfs_into(roots[0], parent).as_dict()
    g.bfs_into(roots[0], parent)
    gg.synchronize()
    n = st["iterations"]
    out = {k: st[k][:n] for k in ("modes", "it_active", "it_new", "it_examined", "it_unvisited")}
    # level 0 runs inside the seed kernel (no iteration launches) when it is a
    # push of at most SEED_ARCS out-arcs, as gg_bfs decides it
    out["seed_level"] = int(st["modes"][0] == 0 and int(g.out_degree()[roots[0]]) <= SEED_ARCS)
    print(json.dumps(out), flush=True)


def _csv_rows(out, suffix):
    import csv
    for root, _, files in os.walk(out):
        for f in files:
            if f.endswith(suffix):
                return list(csv.DictReader(open(os.path.join(root, f))))
    return []


def _last_bfs(rows, key):
    """Dispatches of the last BFS (after its seed kernel) grouped per level:
    [{kernel: row}], a level starting at each step launch."""
    rows = sorted(rows, key=key)
    name = kernel_name
    seeds = [i for i, r in enumerate(rows) if name(r) == "gg::bfs_seed_kernel"]
    levels = []
    for r in rows[seeds[-1] + 1:]:
        k = name(r).replace("gg::bfs_", "").replace("_kernel", "")
        if k in ("init", "seed"):
            break
        if k == "step":
            levels.append({})
        if levels:
            levels[-1][k] = r
    return levels


def level_table(args):
    """Per-level table of one bench BFS: mode, frontier, kernel times (kernel
    trace) and HBM requests of the step and push/heavy kernels (three
    counter passes), printed as markdown."""
    import subprocess
    import tempfile
    prof = "rocprofv3"
    tmp = tempfile.mkdtemp(prefix="gglev", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    child = [sys.executable, os.path.abspath(__file__), "--level-child", "--scale", str(args.scale),
             "--edge-factor", str(args.edge_factor), "--roots", str(args.roots),
             "--seed", str(args.seed), "--relabel-hubs", str(args.relabel_hubs)] + \
        (["--no-reorder"] if args.no_reorder else [])
    runs = [("trace", ["--kernel-trace"])] + [(f"p{i}", ["--pmc", c]) for i, c in
                                               enumerate(LEVEL_PASSES)]
    stats, per = None, {}
    for tag, flags in runs:
        out = os.path.join(tmp, tag)
        r = subprocess.run(["timeout", "-s", "KILL", "120", prof] + flags +
                           ["-d", out, "-o", "run", "-f", "csv", "--"] + child, cwd="/tmp", env=env,
                           capture_output=True, text=True)
        if tag == "trace":
            stats = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
            per["trace"] = _last_bfs(_csv_rows(out, "kernel_trace.csv"),
                                     key=lambda x: int(x["Dispatch_Id"]))
        else:
            rows = _csv_rows(out, "counter_collection.csv")
            merged = {}
            for x in rows:
                d = merged.setdefault(int(x["Dispatch_Id"]), dict(x, counters={}))
                d["counters"][x["Counter_Name"]] = float(x["Counter_Value"])
            per[tag] = _last_bfs(list(merged.values()), key=lambda x: int(x["Dispatch_Id"]))
    print("| level | mode | frontier | new | examined | step us | push/heavy us | "
          "read MB (128B/64B/32B req) | write req (64B) | TCC req | hit/miss |")
    print("|---|---|---|---|---|---|---|---|---|---|---|")
    # (the trace also holds the iterations queued past the end, which return
    # at once: only the levels the BFS ran are listed; a level 0 run by the
    # seed kernel has no launches of its own)
    n0 = stats.get("seed_level", 0)
    for i in range(len(stats["modes"])):
        j = i - n0  # this level's iteration launches
        lv = per["trace"][j] if 0 <= j < len(per["trace"]) else {}
        us = {k: (int(v["End_Timestamp"]) - int(v["Start_Timestamp"])) / 1e3 for k, v in lv.items()}
        c = {}
        for tag in ("p0", "p1", "p2"):
            if 0 <= j < len(per[tag]):
                for k in ("step", "push_heavy"):
                    for n, v in per[tag][j].get(k, {}).get("counters", {}).items():
                        c[n] = c.get(n, 0.0) + v
        r128, r64, r32 = (c.get("TCC_EA0_RDREQ_128B_sum", 0), c.get("TCC_EA0_RDREQ_64B_sum", 0),
                          c.get("TCC_EA0_RDREQ_32B_sum", 0))
        read_mb = (128 * r128 + 64 * r64 + 32 * r32) / 1e6
        print(f"| {i} | {'push' if stats['modes'][i] == 0 else 'pull'} | {stats['it_active'][i]} | "
              f"{stats['it_new'][i]} | {stats['it_examined'][i]} | "
              f"{'(seed kernel)' if j < 0 else format(us.get('step', 0), '.1f')} | "
              f"{'-' if j < 0 else format(us.get('push_heavy', 0), '.1f')} | "
              f"{read_mb:.1f} ({int(r128)}/{int(r64)}/{int(r32)}) | "
              f"{int(c.get('TCC_EA0_WRREQ_sum', 0))} ({int(c.get('TCC_EA0_WRREQ_64B_sum', 0))}) | "
              f"{int(c.get('TCC_REQ_sum', 0))} | {int(c.get('TCC_HIT_sum', 0))}/"
              f"{int(c.get('TCC_MISS_sum', 0))} |", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--roots", type=int, default=16)
    ap.add_argument("--seed", type=int, default=24)
    ap.add_argument("--cpu-sample-roots", type=int, default=16)
    ap.add_argument("--cpu-repeats", type=int, default=6)
    ap.add_argument("--parity-roots", type=int, default=16,
                    help="roots whose parent-tree depths are checked against the oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-reorder", action="store_true",
                    help="keep file vertex ids (no degree-sorted relabeling)")
    ap.add_argument("--relabel-hubs", type=int, default=BFS_RELABEL_HUBS,
                    help="BFS graph: vertices kept in degree order before the file-order rest "
                         "(0 = all in degree order)")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the serving trace (0 skips the serving leg)")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the rocprofv3 counter passes behind roofline.traffic")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--level-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--level-table", action="store_true",
                    help="per-level kernel times and HBM requests of one bench BFS (rocprofv3)")
    ap.add_argument("--pagerank-only", action="store_true",
                    help="run only the PageRank leg (for profiling) and print its figures")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched BFS/BC leg (for profiling) and print its figures")
    ap.add_argument("--cc-only", action="store_true",
                    help="run only the CC leg (for profiling) and print its figures")
    ap.add_argument("--rmat26-only", action="store_true",
                    help="run only the scale-26 SSSP/PageRank leg (for profiling)")
    ap.add_argument("--facade-only", action="store_true",
                    help="run only the facade-toolkit leg and print its figures")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving leg and print its figures")
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-bfs-workers", type=int, default=None,
                    help="serving: workers reserved for BFS queries (default: the option's)")
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return
    if args.level_child:
        level_child(args)
        return
    if args.level_table:
        level_table(args)
        return
    if args.pagerank_only:
        import geminigraph_amd as gg
        gg.init(0)
        g = gg.Graph.rmat(args.scale, args.edge_factor, seed=args.seed, permute=True,
                          reorder=not args.no_reorder)
        V, E = 1 << args.scale, (1 << args.scale) * args.edge_factor
        log(f"PageRank: {pagerank_extra(gg, g, V, E)}")
        # the calibration copy, so counter passes over this run can be scaled
        src, dst = gg.DeviceArray(COPY_BYTES, np.uint8), gg.DeviceArray(COPY_BYTES, np.uint8)
        gg.device_copy(dst, src, COPY_BYTES)
        gg.synchronize()
        return
    if args.batched_only:
        import geminigraph_amd as gg
        gg.init(0)
        log(f"batched: {batch_extra(gg, args.seed + 7, not args.no_reorder)}")
        return
    if args.cc_only:
        import geminigraph_amd as gg
        gg.init(0)
        log(f"CC: {cc_extra(gg, args.scale, args.edge_factor, args.seed, not args.no_reorder)}")
        return
    if args.rmat26_only:
        import geminigraph_amd as gg
        gg.init(0)
        log(f"rmat26: {rmat26_extra(gg, args.seed + 2)}")
        return
    if args.facade_only:
        import geminigraph_amd as gg
        gg.init(0)
        g = gg.Graph.rmat(args.scale, args.edge_factor, seed=args.seed, permute=True)
        roots = pick_roots(g.out_degree(), args.roots, args.seed + 1)[:4]
        g.close()
        log(f"facade: {facade_extra(gg, args.scale, args.edge_factor, args.seed, roots)}")
        return
    if args.serve_only:
        import geminigraph_amd as gg
        gg.init(0)
        log(f"serving: {serve_extra(gg, args.seed + 11, not args.no_reorder, args.serve_seconds, workers=args.serve_workers, bfs_workers=args.serve_bfs_workers)}")
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one partition on one GPU)", file=sys.stderr)
        sys.exit(2)

    import geminigraph_amd as gg
    gg.init(0)
    scale, ef = args.scale, args.edge_factor
    V, E = 1 << scale, (1 << scale) * ef

    t0 = time.time()
    reorder = not args.no_reorder
    bfs_opts = gg.options(**BFS_OPTIONS)  # the BFS leg's options, restored before the extras
    bfs_opts.__enter__()
    g = bfs_graph(gg, args)
    gg.synchronize()
    build_s = time.time() - t0
    log(f"RMAT-{scale} ef{ef}: V={V} E={E} built on device in {build_s:.2f}s")
    roots = pick_roots(g.out_degree(), args.roots, args.seed + 1)

    parent = gg.DeviceArray(V, np.uint32)
    depth = gg.DeviceArray(V, np.uint32)

    # the reference's BFS computes parents only (bfs.cpp:26-94), so does the
    # timed step; depths are checked from the tree the parents form
    def step(with_depth=False):
        m = 0
        for r in roots:
            m += g.bfs_into(r, parent, depth if with_depth else None).m_reach
        return m

    for _ in range(args.warmup):
        step()
    gg.synchronize()
    t = time.perf_counter()
    m_total = 0
    for _ in range(args.steps):
        m_total += step()
    gg.synchronize()
    elapsed = time.perf_counter() - t
    gteps = m_total / elapsed / 1e9
    ms_per_step = elapsed * 1e3 / args.steps
    # one more step with per-launch HIP events (kept out of the timed region)
    gg.prof_reset()
    gg.prof_enable(True)
    step()
    gg.prof_enable(False)
    kernels = {k: gg.prof_read(k) for k in ("bfs_step", "bfs_push_heavy")}
    log(f"BFS x{len(roots)}: {ms_per_step:.2f} ms/step, {gteps:.2f} GTEPS; kernels {kernels}")

    # one instrumented pass (edge counters on) for the algorithmic byte model
    gpu_depths = []
    push_b = pull_b = 0.0
    indeg = g.in_degree()
    n_no_in = int(np.count_nonzero(indeg == 0))
    with gg.options(count_examined=1):
        for r in roots:
            st = g.bfs_into(r, parent).as_dict()
            pb, lb = bfs_model_bytes(st, V, n_no_in - int(indeg[r] == 0), found=4)
            push_b += pb
            pull_b += lb
            if len(gpu_depths) < args.parity_roots:
                gpu_depths.append(depths_from_parents(parent.numpy(), r))
    # Two launches per iteration: bfs_step_kernel (the push/pull decision,
    # then the bottom-up step of a pull or the frontier binning of a push)
    # and bfs_push_heavy_kernel (the push, or the heavy in-list chunks of a
    # pull).  The roofline covers both, every launch of every iteration; per
    # launch = per iteration (per bfs_step_kernel dispatch, as rocprof
    # counts); the bytes are the model's for every iteration, push and pull.
    step_ms = kernels["bfs_step"][0] + kernels["bfs_push_heavy"][0]
    step_launches = kernels["bfs_step"][1]
    step_b = push_b + pull_b
    achieved = step_b / (step_ms / 1e3) / 1e9 if step_ms > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": "bfs_step_kernel + bfs_push_heavy_kernel (both launches of every BFS "
                  "iteration: decision, then bottom-up or frontier binning; push or heavy "
                  "in-lists)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_launch": round(step_b / max(1, step_launches), 1),
        "avg_launch_ms": round(step_ms / max(1, step_launches), 4),
    }

    if not args.no_pmc:
        traffic, detail = measure_traffic(args)
        roofline["traffic"] = round(traffic, 1) if traffic else None
        roofline["traffic_detail"] = detail
        if traffic:
            roofline["traffic_over_algorithmic"] = round(traffic / roofline["bytes_per_launch"], 3)
        log(f"traffic: {roofline['traffic']} B/launch {detail}")

    extras = {"build_s": round(build_s, 2),
              "bfs_timed_region": "value: the 16 gg_bfs calls of each step, parents written into "
                                  "one preallocated array, the traversal buffers (frontiers, "
                                  "visited) kept by the graph from the first call on; graph "
                                  "build excluded as in the reference"}
    # the reference's region (bfs.cpp:23-86): compute() allocates the parent
    # array inside exec_time; here each call also allocates (and, after the
    # timer, frees) its own parent array
    ref_ts, ref_m = [], 0
    for r in roots:
        t = time.perf_counter()
        par = gg.DeviceArray(V, np.uint32)
        st = g.bfs_into(r, par)
        gg.synchronize()
        ref_ts.append(time.perf_counter() - t)
        ref_m += st.m_reach
        par.free()
    extras["bfs_reference_region"] = {
        "ms_per_bfs": round(1e3 * float(np.mean(ref_ts)), 4),
        "gteps": round(ref_m / sum(ref_ts) / 1e9, 2),
        "timed_region": "each call allocates its parent array (gg_malloc, which hands back the "
                        "buffer the previous call freed) inside the timer and synchronises at its "
                        "end, as bfs.cpp:23-29 allocates with alloc_vertex_array; the traversal "
                        "buffers, which "
                        "bfs.cpp also allocates per call (alloc_vertex_subset), stay with the graph"}
    log(f"BFS, reference region: {extras['bfs_reference_region']}")
    # the same roots with the depth array written as well (8 bytes per vertex
    # found instead of 4)
    step(with_depth=True)
    gg.synchronize()
    t = time.perf_counter()
    md = step(with_depth=True)
    gg.synchronize()
    el = time.perf_counter() - t
    extras["bfs_with_depth"] = {"gteps": round(md / el / 1e9, 2),
                                "ms_per_bfs": round(el * 1e3 / len(roots), 4)}
    log(f"BFS with depths: {extras['bfs_with_depth']}")
    # measured stream-copy bandwidth (plain copy kernel, 1 GiB read + 1 GiB write)
    nb = 1 << 30
    src, dst = gg.DeviceArray(nb, np.uint8), gg.DeviceArray(nb, np.uint8)
    gg.device_copy(dst, src, nb)
    gg.prof_reset()
    gg.prof_enable(True)
    for _ in range(3):
        gg.device_copy(dst, src, nb)
    gg.prof_enable(False)
    cms, cn = gg.prof_read("copy")
    extras["stream_copy_GBs"] = round(2 * nb * cn / (cms / 1e3) / 1e9, 1)
    roofline["frac_of_stream_copy"] = round(achieved / extras["stream_copy_GBs"], 4)
    src.free()
    dst.free()
    switch_divisor = gg.get_option("switch_divisor")
    bfs_opts.__exit__(None, None, None)
    if not args.no_extras:
        del g
        parent.free()
        depth.free()
        # PageRank on the same RMAT graph with every vertex in degree order
        # (the default relabeling)
        gp = gg.Graph.rmat(scale, ef, seed=args.seed, permute=True, reorder=reorder)
        extras["pagerank"] = pagerank_extra(gg, gp, V, E)
        log(f"PageRank: {extras['pagerank']}")
        del gp
        extras["facade"] = facade_extra(gg, scale, ef, args.seed, roots[:4])
        log(f"facade: {extras['facade']}")
        extras["pagerank_rmat16"] = pagerank16_extra(gg, args.seed)
        log(f"PageRank RMAT-16: {extras['pagerank_rmat16']}")
        extras["cc"] = cc_extra(gg, scale, ef, args.seed, reorder)
        log(f"CC: {extras['cc']}")
        extras["sssp"] = sssp24_extra(gg, scale, ef, args.seed + 3)
        log(f"SSSP: {extras['sssp']}")
        extras["batched"] = batch_extra(gg, args.seed + 7, reorder)
        log(f"batched: {extras['batched']}")
        extras["rmat26"] = rmat26_extra(gg, args.seed + 2)
        log(f"rmat26: {extras['rmat26']}")
        if args.serve_seconds > 0:
            extras["serving"] = serve_extra(gg, args.seed + 11, reorder, args.serve_seconds,
                                            workers=args.serve_workers,
                                            bfs_workers=args.serve_bfs_workers)
            log(f"serving: {extras['serving']}")

    cpu = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(scale, ef, args.seed, roots, gpu_depths,
                                   args.cpu_sample_roots, args.cpu_repeats)
        extras["depth_parity_vs_oracle"] = parity
        log(f"CPU baseline: {cpu}, parity={parity}")
        if not args.no_extras:
            extras["cpu"] = cpu_extras(scale, ef, args.seed, extras.get("sssp", {}))
            log(f"CPU columns: {extras['cpu']}")

    out = {
        "metric": "BFS GTEPS (RMAT-24, 16 roots, plain method)",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"BFS from {len(roots)} seeded non-isolated roots, RMAT scale "
                               f"{scale} edge factor {ef} (a=0.57 b=0.19 c=0.19), permuted ids",
                   "vertices": V, "edges": E, "roots": len(roots), "seed": args.seed,
                   "outputs": "parent[V] per root, as bfs.cpp:26-94 (depths checked from "
                              "the parent tree)",
                   "reorder": reorder,
                   "relabel_hubs": args.relabel_hubs if reorder else None,
                   "switch_divisor": switch_divisor, "options": BFS_OPTIONS},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

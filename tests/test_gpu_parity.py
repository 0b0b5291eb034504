"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Bar (from the reference's contract):
  * BFS depths, CC labels, reachability sets and per-iteration frontier sizes
    bit-exact; BFS parents valid;
  * SSSP float32 distances bit-exact (the north-star tolerance is 1 ulp);
  * PageRank and BC doubles within 1e-9 relative; BC NaN positions equal.
"""
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

REL = 1e-9  # FP64 tolerance for PageRank and BC


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def roots_of(out_degree, k, seed):
    cand = np.nonzero(out_degree)[0]
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(cand, size=min(k, len(cand)), replace=False)]


def assert_rel(got, ref, rel=REL):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape
    nan_g, nan_r = np.isnan(got), np.isnan(ref)
    assert np.array_equal(nan_g, nan_r), f"NaN mismatch at {np.nonzero(nan_g != nan_r)[0][:10]}"
    inf_g, inf_r = np.isinf(got), np.isinf(ref)
    assert np.array_equal(inf_g, inf_r)
    m = ~(nan_r | inf_r)
    if m.any():
        err = np.abs(got[m] - ref[m]) / np.maximum(np.abs(ref[m]), 1e-300)
        bad = err > rel
        # exact zeros must stay zero
        assert not bad.any(), (f"max rel err {err.max():.3e} at {np.nonzero(m)[0][np.argmax(err)]}")


def check_parents(e, V, root, parent, depth):
    src = e["src"].astype(np.int64)
    dst = e["dst"].astype(np.int64)
    reached = depth != 0xFFFFFFFF
    assert np.array_equal(parent < V, reached)
    assert parent[root] == root
    key = src * V + dst
    key.sort()
    vs = np.nonzero(reached)[0]
    vs = vs[vs != root]
    p = parent[vs].astype(np.int64)
    assert (depth[p] + 1 == depth[vs]).all()
    k = p * V + vs
    pos = np.searchsorted(key, k)
    assert (pos < len(key)).all() and (key[np.minimum(pos, len(key) - 1)] == k).all()


MODES = {"auto": 0, "push": 1, "pull": 2}


# ------------------------------------------------------------ generator/build

@pytest.mark.parametrize("scale,weighted,permute", [(10, False, True), (14, True, True),
                                                    (16, False, False)])
def test_rmat_device_matches_oracle(gg, scale, weighted, permute):
    ne = (1 << scale) * 16
    say("rmat", scale, weighted, permute)
    buf = gg.rmat_device(scale, ne, seed=77, permute=permute, weighted=weighted)
    got = buf.numpy()
    ref = oracle.rmat_edges(scale, seed=77, permute=permute, weighted=weighted)
    assert np.array_equal(got, ref.view(np.uint8))


def _segments_equal(idx_a, nbr_a, idx_b, nbr_b, w_a=None, w_b=None):
    assert np.array_equal(idx_a, idx_b)
    # the device build sorts every list by neighbour (ties in input order)
    V = len(idx_a) - 1
    keys_b = np.repeat(np.arange(V, dtype=np.int64), np.diff(idx_b).astype(np.int64))
    order = np.lexsort((np.arange(len(nbr_b)), nbr_b, keys_b))
    assert np.array_equal(nbr_a, nbr_b[order])
    if w_a is not None:
        assert np.array_equal(w_a, w_b[order])


@pytest.mark.parametrize("weighted,undirected", [(False, False), (True, False), (False, True)])
def test_build_matches_oracle(gg, weighted, undirected):
    scale = 12
    e = oracle.rmat_edges(scale, seed=5, weighted=weighted)
    V = 1 << scale
    say("build", weighted, undirected)
    g = gg.Graph.build(e, V, undirected=undirected)
    o = oracle.Graph(e, V, undirected=undirected)
    assert g.edges == o.edges and g.arcs == o.arcs
    assert np.array_equal(g.out_degree(), o.out_degree())
    assert np.array_equal(g.in_degree(), o.in_degree())
    _segments_equal(*g.csr()[:2], *o.csr()[:2], g.csr()[2], o.csr()[2])
    _segments_equal(*g.csc()[:2], *o.csc()[:2], g.csc()[2], o.csc()[2])


def test_load_file_matches_build(gg, tmp_path):
    e = oracle.rmat_edges(11, seed=9, weighted=True)
    p = tmp_path / "g.bin"
    e.tofile(p)
    V = 1 << 11
    a = gg.Graph.load_directed(p, V, weighted=True)
    b = gg.Graph.build(e, V)
    assert a.edges == len(e)
    for x, y in zip(a.csr(), b.csr()):
        assert np.array_equal(x, y)
    u = gg.Graph.load_undirected_from_directed(p, V, weighted=True)
    assert u.edges == len(e) and u.arcs == 2 * len(e)


def test_bad_inputs_raise(gg, tmp_path):
    e = oracle.edges_from_pairs([(0, 5)])
    with pytest.raises(gg.GGError):
        gg.Graph.build(e, 4)  # endpoint >= vertices
    with pytest.raises(gg.GGError):
        gg.Graph.load_directed(tmp_path / "missing.bin", 10)
    g = gg.Graph.build(oracle.edges_from_pairs([(0, 1)]), 2)
    with pytest.raises(gg.GGError):
        g.bfs(5)
    with pytest.raises(gg.GGError):
        g.sssp(0)  # unweighted


# ----------------------------------------------------------------- PageRank

@pytest.mark.parametrize("scale,ef,seed", [(12, 16, 1), (16, 16, 2024)])
def test_pagerank_matches_oracle(gg, scale, ef, seed):
    e = oracle.rmat_edges(scale, ef, seed=seed)
    V = 1 << scale
    say("pagerank", scale, ef)
    g = gg.Graph.build(e, V)
    pr, deltas, st = g.pagerank(20)
    ref, ref_d = oracle.Graph(e, V).pagerank(20)
    assert_rel(pr, ref)
    assert_rel(deltas[:20], ref_d[:20], rel=1e-8)
    assert st["iterations"] == 20


def test_pagerank_heavy_vertices(gg):
    # a hub with in-degree 5000 exercises the chunked heavy path
    n = 6000
    pairs = [(i, 0) for i in range(1, 5001)] + [(0, i) for i in range(1, 3000)] + \
            [(i, i + 1) for i in range(1, n - 1)]
    e = oracle.edges_from_pairs(pairs)
    g = gg.Graph.build(e, n)
    pr, _, _ = g.pagerank(10)
    ref, _ = oracle.Graph(e, n).pagerank(10)
    assert_rel(pr, ref)


@pytest.mark.parametrize("hot", [0, 1, 64, 1000, 30000, 1 << 20])
def test_pagerank_hot_cold_plan(gg, hot):
    # relabelled graph: in-arcs from the `hot` highest out-degree sources are
    # gathered in source order per destination bin, the others streamed by
    # propagation blocking (1 << 20: every source is hot, no cold arcs)
    scale = 15
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=61)
    g = gg.Graph.build(e, V, reorder=True)
    say("pagerank plan", hot)
    with gg.options(pr_hot_sources=hot):
        pr, deltas, _ = g.pagerank(20)
        pr2, _, _ = g.pagerank(20)  # the plan is reused
    ref, ref_d = oracle.Graph(e, V).pagerank(20)
    assert_rel(pr, ref)
    assert_rel(pr2, ref)
    assert_rel(deltas[:20], ref_d[:20], rel=1e-8)


@pytest.mark.parametrize("hot", [4096, 1 << 18])
def test_pagerank_plan_many_ranges(gg, hot):
    # 8 destination bins, several split bins, cold source ranges whose arcs
    # span several emit parts
    scale = 17
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=67)
    g = gg.Graph.build(e, V, reorder=True)
    say("pagerank plan many ranges", hot)
    with gg.options(pr_hot_sources=hot):
        pr, deltas, _ = g.pagerank(8)
    ref, ref_d = oracle.Graph(e, V).pagerank(8)
    assert_rel(pr, ref)
    assert_rel(deltas[:8], ref_d[:8], rel=1e-8)


@pytest.mark.parametrize("bins,ranges,hot", [(10, 9, 3000), (13, 14, 1 << 19), (14, 8, 700)])
def test_pagerank_plan_sizes(gg, bins, ranges, hot):
    # other bin and range sizes: many bins and ranges, hot ids above 2^18
    scale = 16
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=71)
    g = gg.Graph.build(e, V, reorder=True)
    say("pagerank plan sizes", bins, ranges, hot)
    with gg.options(pr_hot_sources=hot, pr_bin_bits=bins, pr_range_bits=ranges):
        pr, _, _ = g.pagerank(6)
    ref, _ = oracle.Graph(e, V).pagerank(6)
    assert_rel(pr, ref)


@pytest.mark.parametrize("hot", [2, 2500])
def test_pagerank_plan_heavy_prefixes(gg, hot):
    # a hub whose hot prefix alone exceeds the chunk limit, another whose
    # in-list is mostly cold, bins spanning several of the 8192-vertex ranges
    n = 20000
    pairs = [(i, 0) for i in range(1, 5001)] + [(0, i) for i in range(1, 3000)] + \
            [(i, 19999) for i in range(3000, 12000)] + [(i, i + 1) for i in range(1, n - 1)]
    e = oracle.edges_from_pairs(pairs)
    g = gg.Graph.build(e, n, reorder=True)
    with gg.options(pr_hot_sources=hot):
        pr, _, _ = g.pagerank(10)
    ref, _ = oracle.Graph(e, n).pagerank(10)
    assert_rel(pr, ref)


# ---------------------------------------------------------------------- BFS

@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
@pytest.mark.parametrize("scale,seed", [(12, 3), (16, 4)])
def test_bfs_matches_oracle(gg, scale, seed, mode):
    e = oracle.rmat_edges(scale, seed=seed)
    V = 1 << scale
    g = gg.Graph.build(e, V)
    o = oracle.Graph(e, V)
    for root in roots_of(o.out_degree(), 3, seed):
        say("bfs", scale, mode, root)
        with gg.options(force_mode=MODES[mode]):
            parent, depth, fr, st = g.bfs(root)
        _, ref_depth, ref_fr = o.bfs(root)
        assert np.array_equal(depth, ref_depth)
        assert np.array_equal(fr, ref_fr)
        check_parents(e, V, root, parent, depth)
        reached = depth != 0xFFFFFFFF
        assert st["m_reach"] == int(o.out_degree()[reached].astype(np.int64).sum())


def test_bfs_edge_cases(gg):
    # self-loops, duplicates, isolated vertices, V not a multiple of 64
    pairs = [(0, 0), (0, 1), (0, 1), (1, 2), (2, 0), (5, 6), (70, 71), (1, 70)]
    e = oracle.edges_from_pairs(pairs)
    V = 77
    g = gg.Graph.build(e, V)
    o = oracle.Graph(e, V)
    for root in (0, 5, 76):
        parent, depth, fr, _ = g.bfs(root)
        _, rd, rf = o.bfs(root)
        assert np.array_equal(depth, rd) and np.array_equal(fr, rf)
        check_parents(e, V, root, parent, depth)


@pytest.mark.parametrize("mode", ["push", "auto"])
@pytest.mark.parametrize("with_depth", [True, False])
def test_bfs_root_push_at_the_seed_limit(gg, with_depth, mode):
    # the first level is pushed by the one-block seed kernel up to 1024 root
    # out-arcs, by the grid above: roots with exactly 1024 and 1025 out-arcs
    # (duplicates and a self-loop among them), each leaf with two arcs to the
    # next level, and a chain behind that
    pairs = [(0, 0), (0, 1)] + [(0, i) for i in range(1, 1023)]  # 1024 arcs, 1022 leaves
    pairs += [(3000, i) for i in range(3001, 4026)]                # 1025 arcs
    for base in (1, 3001):
        for i in range(1000):
            pairs += [(base + i, 6000 + (base + i) % 1500), (base + i, 7600 + i % 7)]
    pairs += [(7600 + k, 7700 + k) for k in range(7)] + [(7700, 7701), (7701, 7702)]
    e = oracle.edges_from_pairs(pairs)
    V = 7800
    for reorder in (False, True):
        g = gg.Graph.build(e, V, reorder=reorder)
        o = oracle.Graph(e, V)
        for root in (0, 3000):
            say("bfs seed limit", with_depth, mode, reorder, root)
            _, ref_depth, ref_fr = o.bfs(root)
            parent = gg.DeviceArray(V, np.uint32)
            depth = gg.DeviceArray(V, np.uint32) if with_depth else None
            fr = np.zeros(64, dtype=np.uint32)
            with gg.options(force_mode=MODES[mode]):
                st = g.bfs_into(root, parent, depth, fr).as_dict()
            p = parent.numpy()
            d = depth.numpy() if with_depth else depths_from_parents(p, root, V)
            assert np.array_equal(d, ref_depth)
            assert np.array_equal(fr[:len(ref_fr)], ref_fr) and not fr[len(ref_fr):].any()
            assert st["iterations"] == len(ref_fr)
            assert st["it_active"][:len(ref_fr)] == [int(x) for x in ref_fr]
            check_parents(e, V, root, p, d)
            reached = ref_depth != 0xFFFFFFFF
            assert st["m_reach"] == int(o.out_degree()[reached].astype(np.int64).sum())
            parent.free()
            if depth is not None:
                depth.free()
        g.close()


def test_bfs_empty_graph(gg):
    e = oracle.edges_from_pairs(np.zeros((0, 2), dtype=np.uint32))
    g = gg.Graph.build(e, 10)
    parent, depth, fr, st = g.bfs(3)
    assert fr.tolist() == [1]
    assert depth[3] == 0 and (np.delete(depth, 3) == 0xFFFFFFFF).all()
    assert parent[3] == 3 and (np.delete(parent, 3) == 10).all()


def test_bfs_hub_bins(gg):
    # out-degree 20000 (huge push bin), 3000 (mid bin), in-degree 20000 (heavy pull)
    n = 40000
    pairs = [(0, i) for i in range(1, 20001)] + [(1, i) for i in range(20001, 23001)] + \
            [(i, 39999) for i in range(2, 20002)] + [(39999, 1)]
    e = oracle.edges_from_pairs(pairs)
    g = gg.Graph.build(e, n)
    o = oracle.Graph(e, n)
    for mode in ("push", "pull", "auto"):
        with gg.options(force_mode=MODES[mode]):
            parent, depth, fr, _ = g.bfs(0)
        _, rd, rf = o.bfs(0)
        assert np.array_equal(depth, rd) and np.array_equal(fr, rf), mode
        check_parents(e, n, 0, parent, depth)


@pytest.mark.parametrize("light_max", [None, 100000])
@pytest.mark.parametrize("reorder", [False, True])
def test_bfs_degrees_past_16_bits(gg, reorder, light_max):
    # the vertex records hold degrees in 16 bits; 65535 and above are read
    # from the degree arrays.  Hub 1 has out-degree 70000 (m_reach counts it)
    # and in-degree 70000, its one in-neighbour reachable from root 0 the
    # last entry of its file-ordered list (140009, behind 69999 sources with
    # no in-arcs): a bottom-up scan that stopped at 65535 entries would miss
    # it.  pull_light_max 100000 (set before the build) scans the hub inside
    # its group, the default leaves it to the heavy chunks.
    n = 140010
    pairs = [(0, 2), (2, 3), (3, 140009), (140009, 1)]
    pairs += [(1, 10 + i) for i in range(70000)]
    pairs += [(70010 + i, 1) for i in range(69999)]
    e = oracle.edges_from_pairs(pairs)
    opts = {} if light_max is None else {"pull_light_max": light_max}
    with gg.options(**opts):
        g = gg.Graph.build(e, n, reorder=reorder)
    o = oracle.Graph(e, n)
    for root in (0, 1, 70010):
        _, rd, rf = o.bfs(root)
        reached = rd != 0xFFFFFFFF
        for mode in ("push", "pull", "auto"):
            say("bfs 16-bit degrees", reorder, light_max, root, mode)
            with gg.options(force_mode=MODES[mode]):
                parent, depth, fr, st = g.bfs(root)
            assert np.array_equal(depth, rd) and np.array_equal(fr, rf), (root, mode)
            check_parents(e, n, root, parent, depth)
            assert st["m_reach"] == int(o.out_degree()[reached].astype(np.int64).sum())
    g.close()


@pytest.mark.parametrize("waves", [4, 64, 0])
def test_bfs_step_grids(gg, waves):
    # the step launch (decision, then the bottom-up step or the frontier
    # binning) runs on max(bottom-up grid, binning grid) blocks: a grid of 4
    # waves (one block, where the binning of 2^17 vertices takes three), 64,
    # and one wave per group (0); the heavy-chunk step takes half of it
    e = oracle.rmat_edges(17, seed=31)
    V = 1 << 17
    o = oracle.Graph(e, V)
    g = gg.Graph.build(e, V, reorder=True)
    for root in roots_of(o.out_degree(), 3, 31):
        _, rd, rf = o.bfs(root)
        reached = rd != 0xFFFFFFFF
        for mode in ("auto", "push", "pull"):
            say("bfs step grids", waves, root, mode)
            with gg.options(force_mode=MODES[mode], bfs_pull_waves=waves):
                parent, depth, fr, st = g.bfs(root)
            assert np.array_equal(depth, rd) and np.array_equal(fr, rf), (root, mode)
            check_parents(e, V, root, parent, depth)
            assert st["m_reach"] == int(o.out_degree()[reached].astype(np.int64).sum())
    g.close()


@pytest.mark.parametrize("ahead", [1, 2, 8])
def test_bfs_issue_ahead(gg, ahead):
    # the host queues up to `ahead` iterations past the last decided one;
    # those issued past the end return at once and must leave the output,
    # the level sizes and the next call's state alone (roots in a row)
    e = oracle.rmat_edges(15, seed=33)
    V = 1 << 15
    o = oracle.Graph(e, V)
    g = gg.Graph.build(e, V, reorder=True)
    for root in roots_of(o.out_degree(), 3, 33):
        _, rd, rf = o.bfs(root)
        for mode in ("auto", "push", "pull"):
            say("bfs issue ahead", ahead, root, mode)
            with gg.options(force_mode=MODES[mode], bfs_ahead=ahead):
                parent, depth, fr, st = g.bfs(root)
            assert np.array_equal(depth, rd) and np.array_equal(fr, rf), (root, mode)
            assert st["iterations"] == len(rf)
            check_parents(e, V, root, parent, depth)
    g.close()


@pytest.mark.parametrize("record_starts", [1, 0])
@pytest.mark.parametrize("reorder", [False, True])
def test_bfs_list_starts(gg, reorder, record_starts):
    # the bottom-up step takes each in-list's start from the vertex's record
    # (graphs below 2^32 arcs) or, with bfs_record_starts 0 as graphs with
    # more arcs do, from the offsets array: both give the oracle's levels
    e = oracle.rmat_edges(14, seed=21)
    V = 1 << 14
    g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    for root in roots_of(o.out_degree(), 3, 21):
        _, rd, rf = o.bfs(root)
        reached = rd != 0xFFFFFFFF
        for mode in ("pull", "auto"):
            say("bfs list starts", reorder, record_starts, root, mode)
            with gg.options(force_mode=MODES[mode], bfs_record_starts=record_starts):
                parent, depth, fr, st = g.bfs(root)
            assert np.array_equal(depth, rd) and np.array_equal(fr, rf), (root, mode)
            check_parents(e, V, root, parent, depth)
            assert st["m_reach"] == int(o.out_degree()[reached].astype(np.int64).sum())
    g.close()


def depths_from_parents(parent, root, V, max_levels=256):
    """Level of every vertex in the tree the parents form (0xFFFFFFFF where
    parent == V): each reached vertex walks up to the root."""
    depth = np.full(V, 0xFFFFFFFF, dtype=np.uint32)
    node = np.nonzero(parent < V)[0]
    d = np.zeros(len(node), dtype=np.int64)
    cur = node.copy()
    for _ in range(max_levels):
        up = cur != root
        if not up.any():
            break
        cur[up] = parent[cur[up]]
        d[up] += 1
    assert (cur == root).all(), "parents do not form a tree rooted at the root"
    depth[node] = d
    return depth


@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
def test_bfs_parents_only(gg, reorder, mode):
    # without a depth array (the reference's output: parents only) the claims
    # and writes change; the tree the parents form must have the oracle's
    # levels, its arcs must be graph arcs, and level sizes and m_reach match
    cases = [(oracle.rmat_edges(14, seed=9), 1 << 14)]
    n = 40000
    pairs = [(0, i) for i in range(1, 20001)] + [(1, i) for i in range(20001, 23001)] + \
            [(i, 39999) for i in range(2, 20002)] + [(39999, 1)]
    cases.append((oracle.edges_from_pairs(pairs), n))
    for e, V in cases:
        g = gg.Graph.build(e, V, reorder=reorder)
        o = oracle.Graph(e, V)
        parent = gg.DeviceArray(V, np.uint32)
        for root in roots_of(o.out_degree(), 3, 9):
            say("bfs parents only", reorder, mode, V, root)
            fr = np.zeros(64, dtype=np.uint32)
            with gg.options(force_mode=MODES[mode]):
                st = g.bfs_into(root, parent, None, fr)
            p = parent.numpy()
            _, ref_depth, ref_fr = o.bfs(root)
            depth = depths_from_parents(p, root, V)
            assert np.array_equal(depth, ref_depth)
            assert np.array_equal(fr[:len(ref_fr)], ref_fr) and not fr[len(ref_fr):].any()
            check_parents(e, V, root, p, depth)
            reached = ref_depth != 0xFFFFFFFF
            assert st.m_reach == int(o.out_degree()[reached].astype(np.int64).sum())
        parent.free()
        g.close()


# ----------------------------------------------------------------------- CC

@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
@pytest.mark.parametrize("scale,seed", [(12, 5), (16, 6)])
def test_cc_matches_oracle(gg, scale, seed, mode, reorder):
    # relabelled: dense steps on the destination-bin plan (min over in-labels)
    e = oracle.rmat_edges(scale, seed=seed)
    V = 1 << scale
    say("cc", scale, mode, reorder)
    g = gg.Graph.build(e, V, undirected=True, reorder=reorder)
    with gg.options(force_mode=MODES[mode]):
        label, fr, st = g.cc()
    ref, ref_fr = oracle.Graph(e, V, undirected=True).cc()
    assert np.array_equal(label, ref)
    assert np.array_equal(fr, ref_fr)


@pytest.mark.parametrize("bins,ranges,hot", [(9, 9, 500), (8, 12, 1 << 14), (13, 14, 0)])
def test_cc_plan_sizes(gg, bins, ranges, hot):
    # CC's dense steps on the bin plan with many bins, split bins, cold ranges
    # (hot = 0: the plain pull); labels and frontier sizes bit-exact
    scale = 15
    V = 1 << scale
    e = oracle.rmat_edges(scale, seed=7)
    g = gg.Graph.build(e, V, undirected=True, reorder=True)
    say("cc plan", bins, ranges, hot)
    with gg.options(force_mode=MODES["pull"], pr_bin_bits=bins, pr_range_bits=ranges,
                    pr_hot_sources=hot):
        label, fr, _ = g.cc()
    ref, ref_fr = oracle.Graph(e, V, undirected=True).cc()
    assert np.array_equal(label, ref)
    assert np.array_equal(fr, ref_fr)


# --------------------------------------------------------------------- SSSP

@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
@pytest.mark.parametrize("scale,seed", [(12, 7), (15, 8)])
def test_sssp_matches_oracle(gg, scale, seed, mode, reorder):
    # relabelled: dense steps on the destination-bin plan, whose arcs carry
    # their weights
    e = oracle.rmat_edges(scale, seed=seed, weighted=True)
    V = 1 << scale
    g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    for root in roots_of(o.out_degree(), 2, seed):
        say("sssp", scale, mode, root)
        with gg.options(force_mode=MODES[mode]):
            dist, fr, st = g.sssp(root)
        ref, ref_fr = o.sssp(root)
        assert np.array_equal(dist.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(fr, ref_fr)
        reached = ref < 1e9
        assert st["m_reach"] == int(o.out_degree()[reached].astype(np.int64).sum())


@pytest.mark.parametrize("bins,ranges,hot", [(9, 9, 500), (8, 12, 1 << 14), (13, 14, 1 << 20)])
def test_sssp_plan_sizes(gg, bins, ranges, hot):
    # SSSP's dense steps on the weighted bin plan: many bins, split bins,
    # cold ranges, every source hot; distances and frontier sizes bit-exact
    scale = 15
    V = 1 << scale
    e = oracle.rmat_edges(scale, seed=9, weighted=True)
    g = gg.Graph.build(e, V, reorder=True)
    o = oracle.Graph(e, V)
    say("sssp plan", bins, ranges, hot)
    with gg.options(force_mode=MODES["pull"], pr_bin_bits=bins, pr_range_bits=ranges,
                    pr_hot_sources=hot):
        for root in roots_of(o.out_degree(), 2, 9):
            dist, fr, _ = g.sssp(root)
            ref, ref_fr = o.sssp(root)
            assert np.array_equal(dist.view(np.uint32), ref.view(np.uint32))
            assert np.array_equal(fr, ref_fr)


def test_sssp_plan_hub_root(gg):
    # a root whose out-arcs (1000) lie between E/200 and E/50: the first
    # frontier is a push under the plan's divisor, so its queues must be built
    # under that divisor too
    scale = 13
    V = 1 << scale
    base = oracle.rmat_edges(scale, 12, seed=12, weighted=True)
    rng = np.random.default_rng(12)
    hub = V  # an extra vertex
    dst = rng.choice(V, 1000, replace=False)
    pairs = np.concatenate([np.stack([base["src"], base["dst"]], 1),
                            np.stack([np.full(1000, hub), dst], 1)]).astype(np.uint32)
    w = np.concatenate([base["w"], (np.arange(1000) % 7 + 1).astype(np.float32) / 8])
    e = oracle.edges_from_pairs(pairs, w)
    E = len(e)
    assert E // 200 < 1000 < E // 50
    g = gg.Graph.build(e, V + 1, reorder=True)
    o = oracle.Graph(e, V + 1)
    dist, fr, st = g.sssp(hub)
    ref, ref_fr = o.sssp(hub)
    assert np.array_equal(dist.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(fr, ref_fr)
    assert st["modes"][0] == 0  # pushed


# ----------------------------------------------------------------------- BC

@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
@pytest.mark.parametrize("scale,seed", [(12, 9), (15, 10)])
def test_bc_matches_oracle(gg, scale, seed, mode):
    e = oracle.rmat_edges(scale, seed=seed)
    V = 1 << scale
    g = gg.Graph.build(e, V)
    o = oracle.Graph(e, V)
    for root in roots_of(o.out_degree(), 2, seed):
        say("bc", scale, mode, root)
        with gg.options(force_mode=MODES[mode]):
            dep, inv, fr, st = g.bc(root)
        rdep, rinv, rfr = o.bc(root)
        assert np.array_equal(fr, rfr)
        assert_rel(inv, rinv)
        assert_rel(dep, rdep)


def test_bc_hub_graph(gg):
    n = 30000
    pairs = [(0, i) for i in range(1, 12001)] + [(i, 12001 + (i % 50)) for i in range(1, 12001)] + \
            [(12001 + j, 20000) for j in range(50)] + [(i, 20000) for i in range(1, 3000)]
    e = oracle.edges_from_pairs(pairs)
    g = gg.Graph.build(e, n)
    o = oracle.Graph(e, n)
    for mode in ("push", "pull", "auto"):
        with gg.options(force_mode=MODES[mode]):
            dep, inv, fr, _ = g.bc(0)
        rdep, rinv, rfr = o.bc(0)
        assert np.array_equal(fr, rfr), mode
        assert_rel(inv, rinv)
        assert_rel(dep, rdep)


@pytest.mark.parametrize("huge_waves", [1, 3, 6])
def test_push_huge_bin_small_grid(gg, huge_waves):
    # huge frontier vertices (12000 and 9000 out-arcs: 12 and 9 chunks) spread
    # over fewer huge-bin waves than chunks, in a grid rounded up to whole
    # blocks: every chunk must be expanded exactly once (BC adds per arc)
    n = 30000
    pairs = [(0, i) for i in range(1, 12001)] + [(i, 12001 + (i % 50)) for i in range(1, 12001)] + \
            [(12001 + j, 20000) for j in range(50)] + [(20000, 21000 + i) for i in range(9000)] + \
            [(i, 20000) for i in range(1, 3000)]
    w = (np.arange(len(pairs)) % 7 + 1) * np.float32(0.375)
    e = oracle.edges_from_pairs(pairs, weights=w)
    g = gg.Graph.build(e, n)
    o = oracle.Graph(e, n)
    say("push huge bin", huge_waves)
    with gg.options(force_mode=MODES["push"], huge_grid_waves=huge_waves, push_big_max=1024):
        dep, inv, fr, _ = g.bc(0)
        dist, sfr, _ = g.sssp(0)
        _, depth, bfr, _ = g.bfs(0)
    rdep, rinv, rfr = o.bc(0)
    assert np.array_equal(fr, rfr)
    assert_rel(inv, rinv)
    assert_rel(dep, rdep)
    rdist, rsfr = o.sssp(0)
    assert np.array_equal(dist.view(np.uint32), rdist.view(np.uint32))
    assert np.array_equal(sfr, rsfr)
    _, rdepth, rbfr = o.bfs(0)
    assert np.array_equal(depth, rdepth)
    assert np.array_equal(bfr, rbfr)


def test_repeat_runs_identical(gg):
    e = oracle.rmat_edges(14, seed=31)
    V = 1 << 14
    g = gg.Graph.build(e, V)
    root = roots_of(g.out_degree(), 1, 1)[0]
    d1 = g.bfs(root)[1]
    d2 = g.bfs(root)[1]
    assert np.array_equal(d1, d2)
    gu = gg.Graph.build(e, V, undirected=True)
    assert np.array_equal(gu.cc()[0], gu.cc()[0])


# ------------------------------------------------- degree-sorted relabeling

@pytest.mark.parametrize("undirected,weighted", [(False, False), (False, True), (True, False)])
def test_reordered_build_matches_oracle(gg, undirected, weighted):
    e = oracle.rmat_edges(11, seed=41, weighted=weighted)
    V = 1 << 11
    g = gg.Graph.build(e, V, undirected=undirected, reorder=True)
    assert g.reordered
    perm = g.perm()
    assert np.array_equal(np.sort(perm), np.arange(V))
    o = oracle.Graph(e, V, undirected=undirected)
    od = o.out_degree()
    # internal order = descending out-degree, ties by id
    iperm = np.argsort(perm)
    assert (np.diff(od[iperm].astype(np.int64)) <= 0).all()
    assert np.array_equal(g.out_degree(), od)
    assert np.array_equal(g.in_degree(), o.in_degree())
    _segments_equal(*g.csr()[:2], *o.csr()[:2], g.csr()[2], o.csr()[2])
    _segments_equal(*g.csc()[:2], *o.csc()[:2], g.csc()[2], o.csc()[2])


@pytest.mark.parametrize("undirected", [False, True])
def test_hybrid_relabel_order(gg, undirected):
    # relabel_hubs = K: the K highest-out-degree vertices first, in the full
    # degree order; then the others in file order, those without in-arcs last
    e = oracle.rmat_edges(11, seed=44)
    V, K = 1 << 11, 150
    with gg.options(relabel_hubs=K):
        g = gg.Graph.build(e, V, undirected=undirected, reorder=True)
    full = gg.Graph.build(e, V, undirected=undirected, reorder=True)
    iperm, iperm_full = np.argsort(g.perm()), np.argsort(full.perm())
    assert np.array_equal(iperm[:K], iperm_full[:K])
    o = oracle.Graph(e, V, undirected=undirected)
    indeg = o.out_degree() if undirected else o.in_degree()
    rest = np.setdiff1d(np.arange(V), iperm_full[:K])
    want = np.concatenate([rest[indeg[rest] > 0], rest[indeg[rest] == 0]])
    assert np.array_equal(iperm[K:], want)
    assert np.array_equal(g.out_degree(), o.out_degree())
    _segments_equal(*g.csr()[:2], *o.csr()[:2], g.csr()[2], o.csr()[2])
    _segments_equal(*g.csc()[:2], *o.csc()[:2], g.csc()[2], o.csc()[2])


@pytest.mark.parametrize("hubs", [0, 700])
@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
def test_reordered_toolkits_match_oracle(gg, mode, hubs):
    # hubs > 0: the hybrid relabeling (the top vertices in degree order, the
    # rest in file order), the BFS bench graph's
    scale = 13
    V = 1 << scale
    e = oracle.rmat_edges(scale, seed=43)
    ew = oracle.rmat_edges(scale, seed=43, weighted=True)
    with gg.options(relabel_hubs=hubs):
        g = gg.Graph.build(e, V, reorder=True)
        gu = gg.Graph.build(e, V, undirected=True, reorder=True)
        gw = gg.Graph.build(ew, V, reorder=True)
    o, ou, ow = oracle.Graph(e, V), oracle.Graph(e, V, undirected=True), oracle.Graph(ew, V)
    roots = roots_of(o.out_degree(), 2, 43)
    say("reordered", mode)
    with gg.options(force_mode=MODES[mode]):
        if mode != "push":
            pr, deltas, _ = g.pagerank(20)
            rpr, rdeltas = o.pagerank(20)
            assert_rel(pr, rpr)
            assert_rel(deltas[:20], rdeltas[:20], rel=1e-8)
        label, fr, _ = gu.cc()
        rl, rfr = ou.cc()
        assert np.array_equal(label, rl) and np.array_equal(fr, rfr)
        for r in roots:
            parent, depth, fr, st = g.bfs(r)
            _, rd, rf = o.bfs(r)
            assert np.array_equal(depth, rd) and np.array_equal(fr, rf)
            check_parents(e, V, r, parent, depth)
            dist, fr, _ = gw.sssp(r)
            rdist, rfr = ow.sssp(r)
            assert np.array_equal(dist.view(np.uint32), rdist.view(np.uint32))
            assert np.array_equal(fr, rfr)
            dep, inv, fr, _ = g.bc(r)
            rdep, rinv, rfr = o.bc(r)
            assert np.array_equal(fr, rfr)
            assert_rel(inv, rinv)
            assert_rel(dep, rdep)


def test_reordered_subset_uses_original_ids(gg):
    e = oracle.rmat_edges(10, seed=3)
    g = gg.Graph.build(e, 1 << 10, reorder=True)
    s = gg.VertexSubset(g)
    s.set_bit(5)
    assert s.get_bit(5) == 1 and s.count() == 1
    internal = int(g.perm()[5])
    words = s.words()
    assert (int(words[internal >> 6]) >> (internal & 63)) & 1 == 1


# ------------------------------------------------------- batched 64-root BFS

@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
def test_bfs_batch_matches_plain(gg, reorder, mode):
    scale = 13
    V = 1 << scale
    e = oracle.rmat_edges(scale, 32, seed=71)
    g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    roots = roots_of(o.out_degree(), 64, 71)
    say("bfs_batch", reorder, mode)
    with gg.options(force_mode=MODES[mode]):
        depths, st = gg.bfs_batch(g, roots)
    m = 0
    od = o.out_degree().astype(np.int64)
    for i, r in enumerate(roots):
        _, rd, _ = o.bfs(r)
        assert np.array_equal(depths[i], rd), (i, r)
        m += int(od[rd != 0xFFFFFFFF].sum())
    assert st["m_reach"] == m


@pytest.mark.parametrize("bins,ranges,hot", [(9, 9, 500), (8, 12, 1 << 14), (13, 14, 1 << 19)])
def test_bfs_batch_bin_plan(gg, bins, ranges, hot):
    # dense steps of the batched BFS on the destination-bin plan (relabelled
    # graph): many bins and cold ranges, split bins, 40 roots (a partial mask)
    scale = 14
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=73)
    g = gg.Graph.build(e, V, reorder=True)
    o = oracle.Graph(e, V)
    roots = roots_of(o.out_degree(), 40, 73)
    say("bfs_batch plan", bins, ranges, hot)
    with gg.options(force_mode=MODES["pull"], pr_bin_bits=bins, pr_range_bits=ranges,
                    pr_hot_sources=hot):
        depths, _ = gg.bfs_batch(g, roots)
    for i, r in enumerate(roots):
        assert np.array_equal(depths[i], o.bfs(r)[1]), (i, r)


@pytest.mark.parametrize("reorder,mode", [(False, "auto"), (False, "pull"), (True, "auto"),
                                          (True, "push")])
def test_bfs_batch_level_records(gg, reorder, mode):
    # the device-driven batched loop: per-level frontier sizes equal what the
    # oracle's depths give (vertices some root first reaches at that level),
    # the level count is the deepest level + 1, push iterations are the
    # levels run in push mode; without relabelling every in-list longer than
    # 32 entries is split into chunks (heavy kernel of the dense step)
    scale = 12
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=75)
    with gg.options(pull_light_max=32):
        g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    roots = roots_of(o.out_degree(), 50, 75)
    roots = roots + roots[:3]  # repeated roots count once in the first frontier
    say("bfs_batch records", reorder, mode)
    with gg.options(force_mode=MODES[mode]):
        depths, st = gg.bfs_batch(g, roots)
    ref = np.stack([o.bfs(r)[1] for r in roots])
    assert np.array_equal(depths, ref)
    reached = ref != 0xFFFFFFFF
    levels = int(ref[reached].max()) + 1
    assert st["iterations"] == levels
    for i in range(levels):
        fresh = int(((ref == i + 1) & reached).any(axis=0).sum())
        active = int(((ref == i) & reached).any(axis=0).sum())
        assert st["it_active"][i] == active, i
        assert st["it_new"][i] == fresh, i
    modes = st["modes"][:levels]
    assert st["push_iterations"] == sum(1 for m in modes if m == 0)
    if mode == "push":
        assert all(m == 0 for m in modes)
    if mode == "pull":
        assert all(m == 1 for m in modes)


def test_bfs_batch_graph_replays(gg):
    # levels after the first replay captured graphs (one per parity): calls
    # in a row with other root counts, with plain BFS work and syncs between
    # them, after a plan rebuild (other bin sizes) and with the graphs off,
    # each checked against the oracle
    scale = 13
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=76)
    g = gg.Graph.build(e, V, reorder=True)
    o = oracle.Graph(e, V)
    cand = roots_of(o.out_degree(), 64, 76)
    ref = {r: o.bfs(r)[1] for r in cand}
    cases = [(64, {}), (64, {}), (17, {}), (64, {"pr_bin_bits": 12}), (5, {"batch_graphs": 0}),
             (64, {}), (33, {"level_masks": 0}), (64, {})]
    for k, (n, opts) in enumerate(cases):
        roots = cand[:n]
        say("bfs_batch replay", k, n, opts)
        with gg.options(**{"batch_graphs": 1, **opts}):
            depths, _ = gg.bfs_batch(g, roots)
        for i, r in enumerate(roots):
            assert np.array_equal(depths[i], ref[r]), (k, i)
        g.bfs(cand[k])  # eager work between the replays
        gg.synchronize()


def test_bfs_batch_partial_and_repeated_roots(gg):
    e = oracle.rmat_edges(11, seed=72)
    V = 1 << 11
    g = gg.Graph.build(e, V)
    o = oracle.Graph(e, V)
    r = roots_of(o.out_degree(), 3, 72)
    roots = [r[0], r[1], r[0], r[2], 0]  # repeats and a possibly isolated vertex
    depths, _ = gg.bfs_batch(g, roots)
    for i, root in enumerate(roots):
        assert np.array_equal(depths[i], o.bfs(root)[1])
    with pytest.raises(ValueError):
        gg.bfs_batch(g, list(range(65)))


# --------------------------------------------------- hub-first lists (file ids)

def _check_hub_first(idx, nbr, ref_idx, ref_nbr, od, noin):
    """Same lists as the oracle's (as multisets), each ordered by descending
    out-degree of the neighbour, then with-in-arcs first, then id."""
    assert np.array_equal(idx, ref_idx)
    V = len(idx) - 1
    owner = np.repeat(np.arange(V), np.diff(idx).astype(np.int64))
    a = np.lexsort((nbr, owner))
    b = np.lexsort((ref_nbr, owner))
    assert np.array_equal(nbr[a], ref_nbr[b])
    same = owner[1:] == owner[:-1]
    x, y = nbr[:-1][same].astype(np.int64), nbr[1:][same].astype(np.int64)
    ok = (od[x] > od[y]) | ((od[x] == od[y]) & ((noin[x] < noin[y]) |
                                                ((noin[x] == noin[y]) & (x <= y))))
    assert ok.all()


@pytest.mark.parametrize("scale,seed", [(12, 31), (15, 32)])
def test_hub_first_build_and_toolkits(gg, scale, seed):
    e = oracle.rmat_edges(scale, seed=seed, weighted=True)
    V = 1 << scale
    g = gg.Graph.build(e, V, hub_first=True)
    assert not g.reordered
    o = oracle.Graph(e, V)
    od = o.out_degree().astype(np.int64)
    noin = (o.in_degree() == 0).astype(np.int64)
    idx, nbr, _ = g.csc()
    ridx, rnbr, _ = o.csc()
    _check_hub_first(idx, nbr, ridx, rnbr, od, noin)
    idx, nbr, _ = g.csr()
    ridx, rnbr, _ = o.csr()
    _check_hub_first(idx, nbr, ridx, rnbr, od, noin)
    for root in roots_of(o.out_degree(), 2, seed):
        say("hub-first", scale, root)
        parent, depth, fr, _ = g.bfs(root)
        _, rd, rf = o.bfs(root)
        assert np.array_equal(depth, rd) and np.array_equal(fr, rf)
        check_parents(e, V, root, parent, depth)
        dist, fr, _ = g.sssp(root)
        rdist, rfr = o.sssp(root)
        assert np.array_equal(dist.view(np.uint32), rdist.view(np.uint32))
        assert np.array_equal(fr, rfr)
        dep, inv, fr, _ = g.bc(root)
        rdep, rinv, rfr = o.bc(root)
        assert np.array_equal(fr, rfr)
        assert_rel(inv, rinv)
        assert_rel(dep, rdep)
    pr, _, _ = g.pagerank(20)
    rpr, _ = o.pagerank(20)
    assert_rel(pr, rpr)
    gu = gg.Graph.build(e, V, undirected=True, hub_first=True)
    label, fr, _ = gu.cc()
    ref, ref_fr = oracle.Graph(e, V, undirected=True).cc()
    assert np.array_equal(label, ref) and np.array_equal(fr, ref_fr)


def test_into_wrappers_refuse_short_or_mistyped_buffers(gg):
    # the kernels write V (or V * 64) elements: a shorter buffer is refused
    # before the call, as is one of another element type
    e = oracle.rmat_edges(10, seed=33, weighted=True)
    V = 1 << 10
    g = gg.Graph.build(e, V)
    short = gg.DeviceArray(V - 1, np.uint32)
    rows_short = gg.DeviceArray(V, np.uint32)
    f64 = gg.DeviceArray(V, np.float64)
    with pytest.raises(ValueError):
        g.bfs_into(0, short)
    with pytest.raises(ValueError):
        g.sssp_into(0, f64)
    with pytest.raises(ValueError):
        g.cc_into(f64)
    with pytest.raises(ValueError):
        gg.bfs_batch_into(g, [0, 1], rows_short)
    with pytest.raises(ValueError):
        g.pagerank_into(short)
    for x in (short, rows_short, f64):
        x.free()

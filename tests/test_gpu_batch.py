"""Batched queries (up to 64 roots per traversal): every root's result must
equal the single-root method's -- the oracle's, bit for bit for SSSP
distances, within 1e-9 relative for BC."""
import numpy as np
import pytest

import oracle
from test_gpu_parity import MODES, assert_rel, roots_of, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
def test_sssp_batch_matches_oracle(gg, reorder, mode):
    scale = 12
    V = 1 << scale
    e = oracle.rmat_edges(scale, seed=81, weighted=True)
    g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    roots = roots_of(o.out_degree(), 40, 81) + [0]
    say("sssp_batch", reorder, mode)
    with gg.options(force_mode=MODES[mode]):
        dist, _ = gg.sssp_batch(g, roots)
    for i, r in enumerate(roots):
        ref, _ = o.sssp(r)
        assert np.array_equal(dist[i].view(np.uint32), ref.view(np.uint32)), (i, r)


@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
def test_batch_device_loop_matches_host_loop(gg, mode):
    # the device-driven loops (mode decided on the device, no per-iteration
    # host sync) against the host-driven ones: identical distances, depths
    # and BFS per-level records (the batched SSSP relaxes asynchronously
    # within an iteration, so only its fixed point is fixed, not which
    # vertices change in which iteration); in-lists over 32 entries go to
    # the chunked (gated) heavy kernels
    scale = 12
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=83, weighted=True)
    with gg.options(pull_light_max=32):
        g = gg.Graph.build(e, V)
    o = oracle.Graph(e, V)
    roots = roots_of(o.out_degree(), 48, 83)
    keys = ("iterations", "push_iterations", "modes", "it_active", "it_new")
    say("batch device vs host loop", mode)
    out = {}
    for loop in (1, 0):
        with gg.options(force_mode=MODES[mode], batch_device_loop=loop):
            dist, sst = gg.sssp_batch(g, roots)
            depth, bst = gg.bfs_batch(g, roots)
        out[loop] = (dist, depth, {k: sst[k] for k in keys}, {k: bst[k] for k in keys})
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][3] == out[1][3]
    for i in (0, 17, 47):
        assert np.array_equal(out[1][0][i].view(np.uint32), o.sssp(roots[i])[0].view(np.uint32))
        assert np.array_equal(out[1][1][i], o.bfs(roots[i])[1])


@pytest.mark.parametrize("case", ["one_vertex", "no_edges", "v65", "loops_dups"])
@pytest.mark.parametrize("reorder", [False, True])
def test_batch_device_loop_small_graphs(gg, case, reorder):
    # degenerate inputs on the device-driven loops: a single vertex, no edges,
    # V = 65 (one vertex past a bitmap word), self-loops and duplicate arcs;
    # one root, the same root 64 times, and mixed roots
    if case == "one_vertex":
        V, pairs = 1, [(0, 0)]
    elif case == "no_edges":
        V, pairs = 5, []
    elif case == "v65":
        V, pairs = 65, [(i, i + 1) for i in range(64)] + [(64, 0), (3, 64)]
    else:
        V, pairs = 40, [(i, (i * 7) % 40) for i in range(40)] * 3 + [(5, 5), (6, 6), (0, 39)]
    w = (np.arange(max(len(pairs), 1))[:len(pairs)] % 13 + 1).astype(np.float32) / 13
    e = oracle.edges_from_pairs(pairs, w) if pairs else oracle.edges_from_pairs(
        np.zeros((0, 2), np.uint32), np.zeros(0, np.float32))
    g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    say("batch small graph", case, reorder)
    for roots in ([V - 1], [0] * 64, [i % V for i in range(0, 3 * V, 3)][:64]):
        depth, _ = gg.bfs_batch(g, roots)
        dist, _ = gg.sssp_batch(g, roots)
        for i, r in enumerate(roots):
            assert np.array_equal(depth[i], o.bfs(r)[1]), (case, roots, i)
            assert np.array_equal(dist[i].view(np.uint32), o.sssp(r)[0].view(np.uint32)), (case, i)


def test_sssp_batch_hub_graph(gg):
    # out-degree 20000 root (grid-shared push chunks), heavy in-degree sink
    n = 30000
    pairs = [(0, i) for i in range(1, 20001)] + [(i, 29999) for i in range(1, 12001)] + \
            [(29999, 1)] + [(i, i + 1) for i in range(20001, 29998)]
    w = (np.arange(len(pairs)) % 97 + 1).astype(np.float32) / 97
    e = oracle.edges_from_pairs(pairs, w)
    g = gg.Graph.build(e, n)
    o = oracle.Graph(e, n)
    roots = [0, 1, 29999, 20001]
    for mode in ("push", "pull", "auto"):
        with gg.options(force_mode=MODES[mode]):
            dist, _ = gg.sssp_batch(g, roots)
        for i, r in enumerate(roots):
            assert np.array_equal(dist[i].view(np.uint32), o.sssp(r)[0].view(np.uint32)), mode


@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("mode", ["auto", "push", "pull"])
def test_bc_batch_matches_oracle(gg, reorder, mode):
    scale = 12
    V = 1 << scale
    e = oracle.rmat_edges(scale, seed=82)
    g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    roots = roots_of(o.out_degree(), 24, 82)
    say("bc_batch", reorder, mode)
    with gg.options(force_mode=MODES[mode]):
        dep, inv, _ = gg.bc_batch(g, roots)
    for i, r in enumerate(roots):
        rdep, rinv, _ = o.bc(r)
        assert_rel(inv[i], rinv)
        assert_rel(dep[i], rdep)


def test_bc_batch_hub_graph(gg):
    n = 30000
    pairs = [(0, i) for i in range(1, 12001)] + [(i, 12001 + (i % 50)) for i in range(1, 12001)] + \
            [(12001 + j, 20000) for j in range(50)] + [(i, 20000) for i in range(1, 3000)]
    e = oracle.edges_from_pairs(pairs)
    g = gg.Graph.build(e, n)
    o = oracle.Graph(e, n)
    roots = [0, 5, 12001, 20000]
    dep, inv, _ = gg.bc_batch(g, roots)
    for i, r in enumerate(roots):
        rdep, rinv, _ = o.bc(r)
        assert_rel(inv[i], rinv)
        assert_rel(dep[i], rdep)


def test_bc_batch_depth_row_fallback(gg):
    # backward arcs selected from the depth rows (the path taken when the
    # per-level root masks would not fit) give the same dependencies
    scale = 11
    V = 1 << scale
    e = oracle.rmat_edges(scale, seed=83)
    g = gg.Graph.build(e, V, reorder=True)
    o = oracle.Graph(e, V)
    roots = roots_of(o.out_degree(), 17, 83)
    with gg.options(level_masks=0):
        dep, inv, _ = gg.bc_batch(g, roots)
    for i, r in enumerate(roots):
        rdep, rinv, _ = o.bc(r)
        assert_rel(inv[i], rinv)
        assert_rel(dep[i], rdep)


@pytest.mark.parametrize("case", ["rows", "overflow"])
def test_bfs_batch_depth_row_paths(gg, case):
    # depth rows written level by level (level masks off), and a mask store
    # too small for the traversal (rows of the stored levels are built from
    # it, then written directly): same depths as the oracle
    n = 6000
    pairs = [(i, i + 1) for i in range(n - 1)] + [(0, 3000), (3000, 5999), (17, 4)]
    e = oracle.edges_from_pairs(pairs)
    g = gg.Graph.build(e, n, reorder=True)
    o = oracle.Graph(e, n)
    roots = [0, 5, 17, 2999, 3000, 5998]
    opts = {"level_masks": 0} if case == "rows" else {"level_mask_bytes": 40 * n * 8}
    with gg.options(**opts):
        depths, _ = gg.bfs_batch(g, roots)
    for i, r in enumerate(roots):
        assert np.array_equal(depths[i], o.bfs(r)[1]), (case, r)

"""The reference's push/pull switch, reproduced call for call.

core/graph.hpp:1503-1509 decides every process_edges call by
`active_edges < edges / 20` (active_edges = the out-degree sum of the active
vertices, edges = the file's edge count).  With `switch_divisor` and
`plan_switch_divisor` at 20 and BFS's extra bound off (`bfs_alpha` 0), the
engine must take the same mode as the oracle's restatement of that rule in
every call: BFS, CC (doubled degrees, file edge count), SSSP, BC's forward
pass and its backward pass over the transposed graph (transposed out-degree
= in-degree, with the leading no-op call over the empty last level), and
PageRank (always dense).  The results themselves are checked elsewhere; here
the per-call mode sequence and the active-vertex counts are compared.
"""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

REFERENCE = dict(switch_divisor=20, plan_switch_divisor=20, bfs_alpha=0, force_mode=0)


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def roots_of(out_degree, k, seed):
    cand = np.nonzero(out_degree)[0]
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(cand, size=min(k, len(cand)), replace=False)]


def gpu_calls(st, n=None):
    n = st["iterations"] if n is None else n
    assert n <= 256, "more process_edges calls than the stats record"
    return np.asarray(st["modes"][:n], dtype=np.int64)


def ref_calls(tr):
    modes, active = tr.result()
    return modes.astype(np.int64), active.astype(np.int64)


def hub_edges():
    # a hub pushing to 12000 vertices that funnel into 50, then into one
    # vertex of in-degree 3049: levels cross the |E|/20 threshold both ways
    pairs = [(0, i) for i in range(1, 12001)] + [(i, 12001 + (i % 50)) for i in range(1, 12001)] + \
            [(12001 + j, 20000) for j in range(50)] + [(i, 20000) for i in range(1, 3000)] + \
            [(20000, 20001 + i) for i in range(400)]
    return oracle.edges_from_pairs(pairs), 30000


GRAPHS = [("rmat", 12, 21), ("rmat", 15, 22), ("hub", 0, 0)]


def make(kind, scale, seed, weighted=False):
    if kind == "hub":
        e, V = hub_edges()
        if weighted:
            w = ((np.arange(len(e)) % 13) + 1).astype(np.float32) / 16
            e = oracle.edges_from_pairs(np.stack([e["src"], e["dst"]], 1), w)
        return e, V
    return oracle.rmat_edges(scale, seed=seed, weighted=weighted), 1 << scale


@pytest.mark.parametrize("kind,scale,seed", GRAPHS)
@pytest.mark.parametrize("reorder", [False, True])
def test_bfs_switch_matches_reference(gg, kind, scale, seed, reorder):
    e, V = make(kind, scale, seed)
    g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    roots = [0] if kind == "hub" else roots_of(o.out_degree(), 3, seed)
    for root in roots:
        say("bfs switch", kind, scale, reorder, root)
        with gg.options(**REFERENCE):
            _, _, fr, st = g.bfs(root)
        tr = oracle.Trace()
        _, _, rfr = o.bfs(root, trace=tr)
        modes, active = ref_calls(tr)
        assert np.array_equal(fr, rfr)
        assert np.array_equal(gpu_calls(st), modes), (gpu_calls(st), modes)
        assert np.array_equal(np.asarray(st["it_active"][:st["iterations"]]), active)
        if kind == "hub":
            assert 0 in modes and 1 in modes  # the case crosses the threshold


@pytest.mark.parametrize("kind,scale,seed", GRAPHS)
@pytest.mark.parametrize("reorder", [False, True])
def test_cc_switch_matches_reference(gg, kind, scale, seed, reorder):
    # relabelled graphs run their dense steps on the bin plan, under
    # plan_switch_divisor
    e, V = make(kind, scale, seed)
    say("cc switch", kind, scale, reorder)
    g = gg.Graph.build(e, V, undirected=True, reorder=reorder)
    with gg.options(**REFERENCE):
        _, fr, st = g.cc()
    tr = oracle.Trace()
    _, rfr = oracle.Graph(e, V, undirected=True).cc(trace=tr)
    modes, active = ref_calls(tr)
    assert np.array_equal(fr, rfr)
    assert np.array_equal(gpu_calls(st), modes), (gpu_calls(st), modes)
    assert np.array_equal(np.asarray(st["it_active"][:st["iterations"]]), active)


@pytest.mark.parametrize("kind,scale,seed", GRAPHS)
@pytest.mark.parametrize("reorder", [False, True])
def test_sssp_switch_matches_reference(gg, kind, scale, seed, reorder):
    e, V = make(kind, scale, seed, weighted=True)
    g = gg.Graph.build(e, V, reorder=reorder)
    o = oracle.Graph(e, V)
    roots = [0] if kind == "hub" else roots_of(o.out_degree(), 2, seed)
    for root in roots:
        say("sssp switch", kind, scale, reorder, root)
        with gg.options(**REFERENCE):
            _, fr, st = g.sssp(root)
        tr = oracle.Trace()
        _, rfr = o.sssp(root, trace=tr)
        modes, active = ref_calls(tr)
        assert np.array_equal(fr, rfr)
        assert np.array_equal(gpu_calls(st), modes), (gpu_calls(st), modes)
        assert np.array_equal(np.asarray(st["it_active"][:st["iterations"]]), active)


@pytest.mark.parametrize("kind,scale,seed", GRAPHS)
def test_bc_switch_matches_reference(gg, kind, scale, seed):
    # forward levels, then the backward calls over the transposed graph
    e, V = make(kind, scale, seed)
    g = gg.Graph.build(e, V)
    o = oracle.Graph(e, V)
    roots = [0] if kind == "hub" else roots_of(o.out_degree(), 2, seed)
    for root in roots:
        say("bc switch", kind, scale, root)
        with gg.options(**REFERENCE):
            _, _, fr, st = g.bc(root)
        tr = oracle.Trace()
        _, _, rfr = o.bc(root, trace=tr)
        modes, _ = ref_calls(tr)
        assert np.array_equal(fr, rfr)
        assert len(modes) == 2 * len(rfr)  # forward calls + backward calls
        assert np.array_equal(gpu_calls(st), modes), (gpu_calls(st), modes)


def test_pagerank_always_dense(gg):
    e, V = make("rmat", 12, 23)
    g = gg.Graph.build(e, V)
    with gg.options(**REFERENCE):
        _, _, st = g.pagerank(10)
    tr = oracle.Trace()
    oracle.Graph(e, V).pagerank(10, trace=tr)
    modes, _ = ref_calls(tr)
    assert modes.tolist() == [1] * 10
    assert np.array_equal(gpu_calls(st), modes)


def test_long_chain_records_every_iteration(gg):
    # 200 levels: the per-call records (256 of them) hold every iteration
    n = 200
    e = oracle.edges_from_pairs([(i, i + 1) for i in range(n - 1)])
    g = gg.Graph.build(e, n)
    with gg.options(**REFERENCE):
        _, depth, fr, st = g.bfs(0)
    tr = oracle.Trace()
    _, rdepth, rfr = oracle.Graph(e, n).bfs(0, trace=tr)
    modes, active = ref_calls(tr)
    assert np.array_equal(depth, rdepth) and np.array_equal(fr, rfr)
    assert st["iterations"] == len(modes) > 64
    assert len(st["modes"]) == len(modes)
    assert np.array_equal(gpu_calls(st), modes)
    assert np.array_equal(np.asarray(st["it_active"]), active)

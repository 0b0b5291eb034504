"""Query server: several resident graphs, continuous batching across worker
streams, device-memory admission.  Every query's result must be the same
served alone and served in a mix (same-graph queries coalesced into batches,
batches of different graphs overlapping on the device), and equal to the
oracle's single-root result: bit for bit for BFS depths and SSSP distances,
within 1e-9 relative for BC dependencies."""
import numpy as np
import pytest

import oracle
from test_gpu_parity import assert_rel, roots_of, say

pytestmark = pytest.mark.gpu

GRAPHS = [  # (scale, edge factor, seed, reorder)
    (12, 16, 91, False),
    (13, 8, 92, True),
    (11, 16, 93, True),
]


@pytest.fixture(scope="module")
def resident(gg):
    out = []
    for scale, ef, seed, reorder in GRAPHS:
        V = 1 << scale
        e = oracle.rmat_edges(scale, ef, seed=seed, weighted=True)
        out.append((gg.Graph.build(e, V, reorder=reorder), oracle.Graph(e, V)))
    return out


def _trace(resident, n, seed):
    rng = np.random.default_rng(seed)
    q = []
    for _ in range(n):
        gi = int(rng.integers(len(resident)))
        kind = int(rng.choice(3, p=[0.5, 0.3, 0.2]))
        cand = roots_of(resident[gi][1].out_degree(), 40, seed + gi)
        q.append((gi, kind, int(cand[int(rng.integers(len(cand)))])))
    return q


def _check_equal(kind, got, ref):
    if kind == 2:
        assert_rel(got, ref)
    else:
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_serve_same_alone_and_in_mix(gg, resident):
    q = _trace(resident, 120, 7)
    budget = 64 << 30
    # alone: one worker, each query submitted only after the previous finished
    alone = []
    with gg.Server(budget, workers=1) as s:
        ids = [s.add_graph(g) for g, _ in resident]
        for gi, kind, root in q:
            res, _ = s.wait(s.submit(ids[gi], kind, root))
            alone.append(res)
        info = s.info()
        assert info["max_batch"] == 1 and info["completed"] == len(q)
    say("serve alone done")
    # in the mix: everything queued at once on four workers
    with gg.Server(budget, workers=4) as s:
        ids = [s.add_graph(g) for g, _ in resident]
        tickets = [s.submit(ids[gi], kind, root) for gi, kind, root in q]
        mixed = [s.wait(t)[0] for t in tickets]
        info = s.info()
    say("serve mix", info)
    assert info["completed"] == len(q)
    assert info["max_batch"] > 1, "no coalescing happened"
    assert info["batches"] < len(q)
    for (gi, kind, root), a, m in zip(q, alone, mixed):
        _check_equal(kind, m, a)
        o = resident[gi][1]
        if kind == 0:
            assert np.array_equal(m, o.bfs(root)[1]), (gi, root)
        elif kind == 1:
            assert np.array_equal(m.view(np.uint32), o.sssp(root)[0].view(np.uint32)), (gi, root)
        else:
            assert_rel(m, o.bc(root)[0])


def test_serve_matches_plain_calls(gg, resident):
    g, o = resident[1]
    roots = roots_of(o.out_degree(), 10, 5)
    with gg.Server(64 << 30, workers=2) as s:
        gid = s.add_graph(g)
        tb = [s.submit(gid, gg.QUERY_BFS, r) for r in roots]
        ts = [s.submit(gid, gg.QUERY_SSSP, r) for r in roots]
        tc = [s.submit(gid, gg.QUERY_BC, r) for r in roots]
        for r, t in zip(roots, tb):
            _, depth, _, _ = g.bfs(r)
            assert np.array_equal(s.wait(t)[0], depth)
        for r, t in zip(roots, ts):
            dist, _, _ = g.sssp(r)
            assert np.array_equal(s.wait(t)[0].view(np.uint32), dist.view(np.uint32))
        for r, t in zip(roots, tc):
            dep, _, _, _ = g.bc(r)
            assert_rel(s.wait(t)[0], dep)


def test_serve_replay_and_budget(gg, resident):
    q = _trace(resident, 200, 11)
    arrival = np.cumsum(np.random.default_rng(3).exponential(1 / 4000.0, len(q)))
    with gg.Server(64 << 30, workers=3) as s:
        ids = [s.add_graph(g) for g, _ in resident]
        lat, wall = s.replay(arrival, [ids[a] for a, _, _ in q], [k for _, k, _ in q],
                             [r for _, _, r in q])
        info = s.info()
    say("replay", wall, np.percentile(lat, [50, 99]), info)
    assert info["completed"] == len(q)
    assert np.all(lat > 0) and wall >= arrival[-1]
    assert info["used_bytes"] <= info["budget_bytes"]
    # admission: a budget below one graph's footprint is refused up front
    with gg.Server(1 << 20, workers=1) as s:
        with pytest.raises(gg.GGError) as ei:
            s.add_graph(resident[0][0])
        assert ei.value.code == -6


def test_serve_budget_evicts_idle_buffers(gg, resident):
    # room for every graph but for only the largest graph's query buffers: a
    # worker drops the buffers of the graph it served before.
    big = 1  # the scale-13 graph
    with gg.Server(64 << 30, workers=1) as s:
        ids = [s.add_graph(g) for g, _ in resident]
        graphs_bytes = s.info()["used_bytes"]
        g, o = resident[big]
        s.wait(s.submit(ids[big], gg.QUERY_BC, roots_of(o.out_degree(), 1, 3)[0]))
        per_worker = s.info()["used_bytes"] - graphs_bytes
    with gg.Server(graphs_bytes + per_worker + (64 << 10), workers=1) as s:
        ids = [s.add_graph(g) for g, _ in resident]
        for gi in (0, 1, 2, 0):
            g, o = resident[gi]
            r = roots_of(o.out_degree(), 1, gi)[0]
            res, _ = s.wait(s.submit(ids[gi], gg.QUERY_BC, r))
            assert_rel(res, o.bc(r)[0])
            info = s.info()
            assert info["used_bytes"] <= info["budget_bytes"]
            if gi == big:  # the other graphs' buffers were dropped
                assert info["used_bytes"] == graphs_bytes + per_worker


def test_serve_bad_requests(gg, resident):
    g, _ = resident[0]
    with gg.Server(64 << 30, workers=1) as s:
        gid = s.add_graph(g)
        with pytest.raises(gg.GGError):
            s.add_graph(g)
        with pytest.raises(gg.GGError):
            s.submit(gid + 5, gg.QUERY_BFS, 0)
        with pytest.raises(gg.GGError):
            s.submit(gid, 7, 0)
        with pytest.raises(gg.GGError):
            s.submit(gid, gg.QUERY_BFS, g.vertices)
    with pytest.raises(gg.GGError):
        gg.Server(1 << 30, workers=0)


@pytest.mark.parametrize("bfs_workers", [0, 3])
def test_serve_reserved_bfs_workers(gg, resident, bfs_workers):
    # 0: every worker takes every kind; 3 of 4: SSSP and BC all go to the one
    # general worker, BFS to any; results as the oracle's either way
    q = _trace(resident, 60, 13)
    with gg.options(serve_bfs_workers=bfs_workers), gg.Server(64 << 30, workers=4) as s:
        ids = [s.add_graph(g) for g, _ in resident]
        tickets = [s.submit(ids[gi], kind, root) for gi, kind, root in q]
        res = [s.wait(t)[0] for t in tickets]
        info = s.info()
    assert info["completed"] == len(q)
    for (gi, kind, root), m in zip(q, res):
        o = resident[gi][1]
        if kind == 0:
            assert np.array_equal(m, o.bfs(root)[1]), (gi, root)
        elif kind == 1:
            assert np.array_equal(m.view(np.uint32), o.sssp(root)[0].view(np.uint32)), (gi, root)
        else:
            assert_rel(m, o.bc(root)[0])

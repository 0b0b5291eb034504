"""gg_bfs_noalloc: the allocation-free BFS entry point against gg_bfs (and the
oracle) on two calls in a row, the second on a smaller graph, then the first
again (the shared scratch array must not carry anything over); its limits
are refused with GG_EINVAL.  Depths and level sizes must be identical to
gg_bfs's and the oracle's (bfs.cpp:22-109 semantics), parents valid."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def roots_of(out_degree, k, seed):
    cand = np.nonzero(out_degree)[0]
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(cand, size=min(k, len(cand)), replace=False)]


def check_parents(e, V, root, parent, depth):
    reached = depth != 0xFFFFFFFF
    assert np.array_equal(parent < V, reached)
    assert parent[root] == root
    key = np.sort(e["src"].astype(np.int64) * V + e["dst"].astype(np.int64))
    vs = np.nonzero(reached)[0]
    vs = vs[vs != root]
    p = parent[vs].astype(np.int64)
    assert (depth[p] + 1 == depth[vs]).all()
    k = p * V + vs
    pos = np.minimum(np.searchsorted(key, k), len(key) - 1)
    assert (key[pos] == k).all()


def run_both(gg, g, e, o, root, with_depth=True):
    V = g.vertices
    par_a, dep_a = gg.DeviceArray(V, np.uint32), gg.DeviceArray(V, np.uint32)
    par_b, dep_b = gg.DeviceArray(V, np.uint32), gg.DeviceArray(V, np.uint32)
    fr_a = np.zeros(1024, dtype=np.uint32)
    fr_b = np.zeros(1024, dtype=np.uint32)
    st_a = g.bfs_into(root, par_a, dep_a if with_depth else None, fr_a).as_dict()
    st_b = g.bfs_noalloc_into(root, par_b, dep_b if with_depth else None, fr_b).as_dict()
    _, ref_depth, ref_fr = o.bfs(root)
    n = st_a["iterations"]
    assert st_b["iterations"] == n == len(ref_fr)
    assert np.array_equal(fr_b, fr_a) and np.array_equal(fr_b[:n], ref_fr)
    for k in ("m_reach", "reached", "push_iterations"):
        assert st_b[k] == st_a[k], k
    assert st_b["modes"][:n] == st_a["modes"][:n]
    pb = par_b.numpy()
    if with_depth:
        db = dep_b.numpy()
        assert np.array_equal(db, dep_a.numpy())
        assert np.array_equal(db, ref_depth)
        check_parents(e, V, root, pb, db)
    else:
        check_parents(e, V, root, pb, ref_depth)
    for x in (par_a, dep_a, par_b, dep_b):
        x.free()


def test_noalloc_matches_bfs_two_calls_then_smaller(gg):
    big_e = oracle.rmat_edges(18, 16, seed=41)
    big_V = 1 << 18
    big = gg.Graph.build(big_e, big_V, reorder=True)
    big_o = oracle.Graph(big_e, big_V)
    small_e = oracle.rmat_edges(12, 8, seed=42)
    small_V = 1 << 12
    small = gg.Graph.build(small_e, small_V)
    small_o = oracle.Graph(small_e, small_V)
    for root in roots_of(big_o.out_degree(), 2, 41):
        say("noalloc big", root)
        run_both(gg, big, big_e, big_o, root)
    for root in roots_of(small_o.out_degree(), 2, 42):
        say("noalloc small", root)
        run_both(gg, small, small_e, small_o, root)
        run_both(gg, small, small_e, small_o, root, with_depth=False)
    # the large graph again, after the smaller one left its state in scratch
    root = roots_of(big_o.out_degree(), 3, 43)[-1]
    say("noalloc big again", root)
    run_both(gg, big, big_e, big_o, root, with_depth=False)
    big.close()
    small.close()


def test_noalloc_degrees_past_16_bits(gg):
    # vertex records keep degrees in 16 bits (65535 and above read from the
    # degree arrays): a hub with 70000 out-arcs and 70000 in-arcs, its one
    # reachable in-neighbour last in its list; roots before and at the hub
    n = 140010
    pairs = [(0, 2), (2, 3), (3, 140009), (140009, 1)]
    pairs += [(1, 10 + i) for i in range(70000)]
    pairs += [(70010 + i, 1) for i in range(69999)]
    e = oracle.edges_from_pairs(pairs)
    g = gg.Graph.build(e, n)
    o = oracle.Graph(e, n)
    for root in (0, 1):
        for mode in (0, 1, 2):
            say("noalloc 16-bit degrees", root, mode)
            with gg.options(force_mode=mode):
                run_both(gg, g, e, o, root)
    g.close()


def test_noalloc_forced_modes(gg):
    e = oracle.rmat_edges(14, 16, seed=44)
    V = 1 << 14
    g = gg.Graph.build(e, V, reorder=True)
    o = oracle.Graph(e, V)
    root = roots_of(o.out_degree(), 1, 44)[0]
    for mode in (1, 2):
        say("noalloc forced mode", mode)
        with gg.options(force_mode=mode):
            run_both(gg, g, e, o, root)
    g.close()


def test_noalloc_limits(gg):
    # more vertices than the scratch array holds, and too many level records
    V = (1 << 20) + 1
    e = oracle.edges_from_pairs([(0, 1), (1, V - 1)])
    g = gg.Graph.build(e, V)
    par = gg.DeviceArray(V, np.uint32)
    with pytest.raises(gg.GGError) as ex:
        g.bfs_noalloc_into(0, par)
    assert ex.value.code == -1 and "2^20" in str(ex.value)
    par.free()
    g.close()
    small = gg.Graph.build(oracle.edges_from_pairs([(0, 1)]), 4)
    par = gg.DeviceArray(4, np.uint32)
    with pytest.raises(gg.GGError):
        small.bfs_noalloc_into(0, par, None, np.zeros((1 << 16) + 1, dtype=np.uint32))
    st = small.bfs_noalloc_into(0, par)
    assert st.iterations == 2 and par.numpy().tolist() == [0, 0, 4, 4]
    par.free()
    small.close()

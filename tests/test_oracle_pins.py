"""Pins for the CPU oracle: each oracle function is checked against something
that does not restate it (closed forms, independent library routines, brute
force), chosen so that a dropped term, a wrong index or a transposed operand
fails at least one check.  The reference ships no tests, golden vectors or
fixtures for this path (it has no test directory at all), so these checks and
the committed fixtures under tests/golden/ carry parity."""
import heapq

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

import oracle

V_INF = np.uint32(0xFFFFFFFF)


# ---------------------------------------------------------------- generator

def _mix64(z):
    M = (1 << 64) - 1
    z &= M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def _py_rmat_edge(i, scale, seed, a, b, c, weighted):
    """The generator's definition, evaluated with Python integers."""
    M = (1 << 64) - 1
    key = _mix64(seed ^ 0xD1B54A32D192ED03)
    word = lambda j: _mix64((key + ((16 * i + j + 1) * 0x9E3779B97F4A7C15)) & M)
    ta, tab, tabc = int(a * 2**32), int((a + b) * 2**32), int((a + b + c) * 2**32)
    s = d = 0
    for l in range(scale):
        w = word(l >> 1)
        r = (w >> 32) if (l & 1) else (w & 0xFFFFFFFF)
        bit = 1 << (scale - 1 - l)
        if r < ta:
            pass
        elif r < tab:
            d |= bit
        elif r < tabc:
            s |= bit
        else:
            s |= bit
            d |= bit
    wt = None
    if weighted:
        wt = np.float32(((word(15) >> 40) + 1) / 2**24)
    return s, d, wt


def test_rmat_matches_definition():
    e = oracle.rmat_edges(12, num_edges=64, seed=7, permute=False, weighted=True)
    for i in range(64):
        s, d, w = _py_rmat_edge(i, 12, 7, 0.57, 0.19, 0.19, True)
        assert (e["src"][i], e["dst"][i]) == (s, d)
        assert e["w"][i] == w


def test_rmat_deterministic_and_seeded():
    a = oracle.rmat_edges(10, seed=5)
    b = oracle.rmat_edges(10, seed=5)
    c = oracle.rmat_edges(10, seed=6)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a["src"].max() < 1024 and a["dst"].max() < 1024


@pytest.mark.parametrize("scale", [1, 2, 3, 7, 16, 20])
def test_permutation_is_bijection(scale):
    vals = np.array([oracle.permute(x, scale, 11) for x in range(1 << min(scale, 12))],
                    dtype=np.int64)
    assert len(np.unique(vals)) == len(vals)
    assert vals.max() < (1 << scale)
    if scale <= 12:
        assert np.array_equal(np.sort(vals), np.arange(1 << scale))


def test_rmat_quadrant_frequencies():
    # scale 1, no permutation: (src, dst) = (0,0)/(0,1)/(1,0)/(1,1) with a/b/c/d
    e = oracle.rmat_edges(1, num_edges=200000, seed=3, permute=False)
    q = e["src"].astype(np.int64) * 2 + e["dst"]
    f = np.bincount(q, minlength=4) / len(q)
    assert np.allclose(f, [0.57, 0.19, 0.19, 0.05], atol=0.005)


def test_rmat_weights_in_unit_interval():
    e = oracle.rmat_edges(8, seed=2, weighted=True)
    w = e["w"].astype(np.float64)
    assert (w > 0).all() and (w <= 1).all()
    k = w * 2**24
    assert np.array_equal(k, np.round(k))


# -------------------------------------------------------------------- build

def _bruteforce_csr(src, dst, V, w=None):
    order = np.lexsort((np.arange(len(src)), src))  # stable by key
    idx = np.zeros(V + 1, dtype=np.uint64)
    idx[1:] = np.cumsum(np.bincount(src, minlength=V))
    return idx, dst[order], (w[order] if w is not None else None)


@pytest.mark.parametrize("weighted", [False, True])
def test_build_directed_matches_bruteforce(weighted):
    e = oracle.rmat_edges(9, seed=4, weighted=weighted)
    V = 1 << 9
    g = oracle.Graph(e, V)
    assert g.edges == len(e) and g.arcs == len(e)
    s, d = e["src"], e["dst"]
    w = e["w"] if weighted else None
    assert np.array_equal(g.out_degree(), np.bincount(s, minlength=V))
    assert np.array_equal(g.in_degree(), np.bincount(d, minlength=V))
    for got, (key, nb) in ((g.csr(), (s, d)), (g.csc(), (d, s))):
        idx, nbr, ww = _bruteforce_csr(key, nb, V, w)
        assert np.array_equal(got[0], idx)
        assert np.array_equal(got[1], nbr)
        if weighted:
            assert np.array_equal(got[2], ww)


def test_build_undirected_doubles_edges():
    e = oracle.rmat_edges(9, seed=4)
    V = 1 << 9
    g = oracle.Graph(e, V, undirected=True)
    assert g.edges == len(e)          # switch threshold uses the file count
    assert g.arcs == 2 * len(e)
    s = np.concatenate([e["src"], e["dst"]])
    d = np.concatenate([e["dst"], e["src"]])
    deg = np.bincount(s, minlength=V)
    assert np.array_equal(g.out_degree(), deg)
    assert np.array_equal(g.in_degree(), deg)
    idx, nbr, _ = g.csr()
    for v in range(V):
        assert sorted(nbr[idx[v]:idx[v + 1]].tolist()) == sorted(d[s == v].tolist())
    cidx, cnbr, _ = g.csc()
    assert np.array_equal(cidx, idx) and np.array_equal(cnbr, nbr)


# ----------------------------------------------------------------- pagerank

def _pr_dense(src, dst, V, iters, d=0.85):
    A = np.zeros((V, V))
    np.add.at(A, (dst, src), 1.0)  # A[v, u] = multiplicity of u -> v
    od = np.bincount(src, minlength=V).astype(np.float64)
    x = np.where(od > 0, 1.0 / np.maximum(od, 1), 1.0)
    for it in range(iters):
        y = (1 - d) + d * (A @ x)
        if it != iters - 1:
            y = np.where(od > 0, y / np.maximum(od, 1), y)
        x = y
    return x


def test_pagerank_cycle_is_one():
    n = 17
    e = oracle.edges_from_pairs([(i, (i + 1) % n) for i in range(n)])
    pr, deltas = oracle.Graph(e, n).pagerank(20)
    assert np.allclose(pr, 1.0, rtol=0, atol=1e-15)
    assert deltas[0] == 1.0 and np.allclose(deltas[1:20], 0.0)


def test_pagerank_complete_graph_is_one():
    n = 9
    e = oracle.edges_from_pairs([(i, j) for i in range(n) for j in range(n) if i != j])
    pr, _ = oracle.Graph(e, n).pagerank(20)
    assert np.allclose(pr, 1.0, rtol=1e-14)


def test_pagerank_star_closed_form():
    # leaves -> centre; the centre has no out-edges and keeps curr = next
    n = 11
    e = oracle.edges_from_pairs([(i, 0) for i in range(1, n)])
    pr, _ = oracle.Graph(e, n).pagerank(20)
    assert np.allclose(pr[1:], 0.15, rtol=1e-14)
    assert abs(pr[0] - (0.15 + 0.85 * (n - 1) * 0.15)) < 1e-12
    pr1, _ = oracle.Graph(e, n).pagerank(1)
    assert abs(pr1[0] - (0.15 + 0.85 * (n - 1))) < 1e-12


def test_pagerank_matches_dense_recurrence_with_duplicates():
    e = oracle.rmat_edges(8, seed=9)
    V = 1 << 8
    pr, _ = oracle.Graph(e, V).pagerank(20)
    ref = _pr_dense(e["src"].astype(np.int64), e["dst"].astype(np.int64), V, 20)
    assert np.allclose(pr, ref, rtol=1e-12, atol=0)


def test_pagerank_delta_matches_definition():
    e = oracle.rmat_edges(8, seed=9)
    V = 1 << 8
    g = oracle.Graph(e, V)
    od = g.out_degree().astype(np.float64)
    prev = None
    for it in (1, 2, 3):
        x, deltas = g.pagerank(it)
        if prev is not None:
            # delta(it) = sum |next - curr| * outdeg / V over non-final iterations
            xd = np.where(od > 0, x / np.maximum(od, 1), x)  # next as stored
            xd_prev = prev
            expect = (np.abs(xd - xd_prev) * np.where(od > 0, od, 1)).sum() / V
            x_full, deltas_full = g.pagerank(it + 1)
            assert abs(deltas_full[it] - expect) <= 1e-12 * max(1.0, expect)
        prev = np.where(od > 0, x / np.maximum(od, 1), x)


# ---------------------------------------------------------------------- bfs

def _roots(g, k, seed=0):
    od = g.out_degree()
    cand = np.nonzero(od)[0]
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(k, len(cand)), replace=False)


def check_bfs(e, V, root, parent, depth, frontier):
    src, dst = e["src"].astype(np.int64), e["dst"].astype(np.int64)
    A = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(V, V))
    dist = csgraph.shortest_path(A, directed=True, unweighted=True, indices=int(root))
    reached = np.isfinite(dist)
    assert np.array_equal(depth != V_INF, reached)
    assert np.array_equal(depth[reached].astype(np.int64), dist[reached].astype(np.int64))
    assert np.array_equal((parent < V), reached)
    assert parent[root] == root
    # parent valid: edge parent -> v exists and depth[parent] = depth[v] - 1
    ok = set(zip(src.tolist(), dst.tolist()))
    for v in np.nonzero(reached)[0]:
        if v == root:
            continue
        p = int(parent[v])
        assert (p, int(v)) in ok and depth[p] + 1 == depth[v]
    levels = np.bincount(depth[reached].astype(np.int64))
    assert np.array_equal(frontier, levels)


@pytest.mark.parametrize("scale,seed", [(9, 1), (11, 2), (12, 3)])
def test_bfs_matches_scipy(scale, seed):
    e = oracle.rmat_edges(scale, seed=seed)
    V = 1 << scale
    g = oracle.Graph(e, V)
    for root in _roots(g, 3, seed):
        parent, depth, frontier = g.bfs(int(root))
        check_bfs(e, V, int(root), parent, depth, frontier)


def test_bfs_path_graph():
    n = 40
    e = oracle.edges_from_pairs([(i, i + 1) for i in range(n - 1)])
    parent, depth, fr = oracle.Graph(e, n).bfs(0)
    assert np.array_equal(depth, np.arange(n))
    assert np.array_equal(parent[1:], np.arange(n - 1))
    assert np.array_equal(fr, np.ones(n))


# ----------------------------------------------------------------------- cc

def check_cc(e, V, label, frontier):
    src, dst = e["src"].astype(np.int64), e["dst"].astype(np.int64)
    A = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(V, V))
    ncomp, comp = csgraph.connected_components(A, directed=True, connection="weak")
    minid = np.full(ncomp, V, dtype=np.int64)
    np.minimum.at(minid, comp, np.arange(V))
    assert np.array_equal(label.astype(np.int64), minid[comp])
    # BSP (Jacobi) label propagation: frontier = vertices lowered per round
    s2 = np.concatenate([src, dst])
    d2 = np.concatenate([dst, src])
    lab = np.arange(V, dtype=np.int64)
    sizes = [V]
    while True:
        nl = lab.copy()
        np.minimum.at(nl, d2, lab[s2])
        low = nl < lab
        lab = nl
        if not low.any():
            break
        sizes.append(int(low.sum()))
    assert np.array_equal(frontier, np.array(sizes))


@pytest.mark.parametrize("scale,seed", [(9, 1), (12, 5)])
def test_cc_matches_scipy(scale, seed):
    e = oracle.rmat_edges(scale, seed=seed)
    V = 1 << scale
    label, frontier = oracle.Graph(e, V, undirected=True).cc()
    check_cc(e, V, label, frontier)


def test_cc_two_components_and_isolated():
    e = oracle.edges_from_pairs([(5, 3), (3, 4), (1, 2), (2, 1)])
    label, _ = oracle.Graph(e, 7, undirected=True).cc()
    assert label.tolist() == [0, 1, 1, 3, 3, 3, 6]


# --------------------------------------------------------------------- sssp

def f32_dijkstra(e, V, root):
    adj = [[] for _ in range(V)]
    for s, d, w in zip(e["src"].tolist(), e["dst"].tolist(), e["w"]):
        adj[s].append((d, np.float32(w)))
    dist = np.full(V, np.float32(1e9), dtype=np.float32)
    dist[root] = np.float32(0)
    h = [(0.0, root)]
    done = np.zeros(V, bool)
    while h:
        _, u = heapq.heappop(h)
        if done[u]:
            continue
        done[u] = True
        for v, w in adj[u]:
            nd = np.float32(dist[u] + w)
            if nd < dist[v]:
                dist[v] = nd
                heapq.heappush(h, (float(nd), v))
    return dist


def bsp_bellman_ford_sizes(e, V, root):
    src, dst, w = e["src"].astype(np.int64), e["dst"].astype(np.int64), e["w"]
    dist = np.full(V, np.float32(1e9), dtype=np.float32)
    dist[root] = 0
    sizes = [1]
    while True:
        cand = (dist[src] + w).astype(np.float32)
        nd = dist.copy()
        np.minimum.at(nd, dst, cand)
        low = nd < dist
        dist = nd
        if not low.any():
            break
        sizes.append(int(low.sum()))
    return dist, np.array(sizes)


@pytest.mark.parametrize("scale,seed", [(8, 1), (10, 2)])
def test_sssp_bitexact_vs_f32_dijkstra(scale, seed):
    e = oracle.rmat_edges(scale, seed=seed, weighted=True)
    V = 1 << scale
    g = oracle.Graph(e, V)
    for root in _roots(g, 2, seed):
        dist, frontier = g.sssp(int(root))
        ref = f32_dijkstra(e, V, int(root))
        assert np.array_equal(dist.view(np.uint32), ref.view(np.uint32))
        d2, sizes = bsp_bellman_ford_sizes(e, V, int(root))
        assert np.array_equal(d2.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(frontier, sizes)


def test_sssp_fixed_point_checker():
    # the full-size checker: zero on Dijkstra's distances, and each planted
    # mistake (a distance one ulp too high or too low, a reached vertex
    # marked unreached, an unreached one given a distance, the root off 0)
    # is caught
    scale, seed = 10, 4
    e = oracle.rmat_edges(scale, seed=seed, weighted=True)
    V = 1 << scale
    idx, src, w = oracle.Graph(e, V).csc()
    root = int(_roots(oracle.Graph(e, V), 1, seed)[0])
    ref = f32_dijkstra(e, V, root)
    assert oracle.sssp_fixed_point_violations(idx, src, w, ref, root) == 0
    reached = np.nonzero((ref < 1e9) & (np.arange(V) != root))[0]
    unreached = np.nonzero(ref >= 1e9)[0]
    assert len(reached) > 10 and len(unreached) > 0
    v = int(reached[len(reached) // 2])
    for bad_value in (np.nextafter(ref[v], np.float32(np.inf)),
                      np.nextafter(ref[v], np.float32(0)), np.float32(1e9)):
        d = ref.copy()
        d[v] = bad_value
        assert oracle.sssp_fixed_point_violations(idx, src, w, d, root) >= 1, bad_value
    d = ref.copy()
    d[int(unreached[0])] = np.float32(0.5)
    assert oracle.sssp_fixed_point_violations(idx, src, w, d, root) >= 1
    d = ref.copy()
    d[root] = np.float32(2 ** -24)
    assert oracle.sssp_fixed_point_violations(idx, src, w, d, root) >= 1


# ----------------------------------------------------------------------- bc

def test_bc_matches_networkx_brandes():
    nx = pytest.importorskip("networkx")
    from networkx.algorithms.centrality.betweenness import _single_source_shortest_path_basic
    e = oracle.rmat_edges(9, seed=8)
    V = 1 << 9
    pairs = np.unique(np.stack([e["src"], e["dst"]], 1), axis=0)  # simple digraph
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    e2 = oracle.edges_from_pairs(pairs)
    g = oracle.Graph(e2, V)
    G = nx.DiGraph()
    G.add_nodes_from(range(V))
    G.add_edges_from(pairs.tolist())
    for root in _roots(g, 3, 8):
        root = int(root)
        dep, inv, frontier = g.bc(root)
        _, _, sigma, D = _single_source_shortest_path_basic(G, root)
        bc = nx.betweenness_centrality_subset(G, [root], list(range(V)), normalized=False)
        reached = np.array([v in D for v in range(V)])
        for v in range(V):
            if not reached[v]:
                assert np.isnan(dep[v]) and np.isinf(inv[v])
                continue
            assert abs(1.0 / inv[v] - sigma[v]) <= 1e-9 * sigma[v]
            if v == root:
                assert abs(dep[v] - (reached.sum() - 1)) <= 1e-9 * reached.sum()
            else:
                assert abs(dep[v] - bc[v]) <= 1e-9 * max(1.0, abs(bc[v]))
        lv = np.bincount(np.array([D[v] for v in D]))
        assert np.array_equal(frontier, lv)


def test_bc_multigraph_hand_case():
    # 0->1 twice, 1->2: sigma = [1, 2, 2]; delta(0) = 2 (reached - 1),
    # delta(1) = sigma1/sigma2 * (1 + delta2) = 1, delta(2) = 0; 3 unreached.
    e = oracle.edges_from_pairs([(0, 1), (0, 1), (1, 2)])
    dep, inv, fr = oracle.Graph(e, 4).bc(0)
    assert np.allclose(1 / inv[:3], [1, 2, 2])
    assert np.allclose(dep[:3], [2, 1, 0])
    assert np.isnan(dep[3])
    assert fr.tolist() == [1, 1, 1]


def test_oracle_thread_count_invariance():
    e = oracle.rmat_edges(10, seed=12)
    V = 1 << 10
    g = oracle.Graph(e, V)
    root = int(_roots(g, 1, 1)[0])
    prev = oracle.get_threads()
    try:
        oracle.set_threads(1)
        d1 = g.bfs(root)[1]
        p1 = g.pagerank(5)[0]
        oracle.set_threads(4)
        d4 = g.bfs(root)[1]
        p4 = g.pagerank(5)[0]
    finally:
        oracle.set_threads(prev)
    assert np.array_equal(d1, d4)
    assert np.array_equal(p1, p4)  # one partial per vertex: summation order is fixed

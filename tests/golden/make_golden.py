"""Writes the golden fixtures in this directory: small graphs and the results
the reference's toolkits define on them, computed by routines independent of
the oracle and of the engine (numpy / scipy / a float32 heap Dijkstra /
textbook Brandes), so that both can be checked against them.

    python tests/golden/make_golden.py

Definitions (toolkits/*.cpp of the reference):
  * BFS depth = unweighted hop distance (0xFFFFFFFF unreached); level sizes =
    counts per depth (bfs.cpp's "active(i)" frontier sizes);
  * SSSP = float32 shortest distances with float32 left-fold sums, 1e9
    unreached (sssp.cpp:44-71);
  * CC = the smallest vertex id of each weakly connected component
    (cc.cpp:22-109 on the symmetrised graph);
  * PageRank = 20 iterations of 0.15 + 0.85 * sum(curr[u]) with curr
    divided by the out-degree except after the last iteration, curr0 =
    1/outdeg or 1 (pagerank.cpp:35-97), duplicates counted;
  * BC (one root) = Brandes dependencies over arcs with multiplicity, the
    root's own = reached - 1, NaN where unreached; sigma = shortest-path
    counts (bc.cpp:24-196).
"""
import heapq
import os

import numpy as np
from scipy.sparse import coo_matrix, csgraph

HERE = os.path.dirname(os.path.abspath(__file__))
UNSET = np.uint32(0xFFFFFFFF)


def rmat_edges(scale, ef, seed, a=0.57, b=0.19, c=0.19):
    """RMAT by recursive quadrant descent with numpy's generator (not the
    engine's or the oracle's counter-based one), ids permuted, float32
    weights on the (k+1) * 2^-24 lattice."""
    rng = np.random.default_rng(seed)
    n = (1 << scale) * ef
    src = np.zeros(n, dtype=np.int64)
    dst = np.zeros(n, dtype=np.int64)
    for _ in range(scale):
        r = rng.random(n)
        sb = (r >= a + b).astype(np.int64)                       # quadrants c, d
        db = ((r >= a) & (r < a + b)) | (r >= a + b + c)         # quadrants b, d
        src = src * 2 + sb
        dst = dst * 2 + db.astype(np.int64)
    perm = rng.permutation(1 << scale)
    w = ((rng.integers(0, 1 << 24, n) + 1) * 2.0 ** -24).astype(np.float32)
    return perm[src].astype(np.uint32), perm[dst].astype(np.uint32), w


def hub_graph():
    """Out-degree 1500 (above every push bin), in-degree 1500 (a heavy pull
    list), a tail, a second component and isolated vertices."""
    pairs = [(0, i) for i in range(1, 1501)] + [(i, 1501) for i in range(1, 1501)] + \
            [(1501, i) for i in range(1502, 1601)] + [(1600, 0)] + \
            [(1610 + i, 1610 + (i + 1) % 30) for i in range(30)]
    p = np.array(pairs, dtype=np.uint32)
    w = (((np.arange(len(p)) % 11) + 1) / 16.0).astype(np.float32)
    return p[:, 0], p[:, 1], w, 1700


def edge_case_graph():
    """Self-loops, duplicates, isolated vertices, V not a multiple of 64."""
    p = np.array([(0, 0), (0, 1), (0, 1), (1, 2), (2, 0), (5, 6), (70, 71), (1, 70)],
                 dtype=np.uint32)
    w = np.array([0.5, 0.25, 0.75, 1.0, 0.125, 0.5, 1.0, 0.375], dtype=np.float32)
    return p[:, 0], p[:, 1], w, 77


def bfs_depth(src, dst, V, root):
    A = coo_matrix((np.ones(len(src)), (src, dst)), shape=(V, V)).tocsr()
    d = csgraph.shortest_path(A, directed=True, unweighted=True, indices=root)
    out = np.full(V, UNSET, dtype=np.uint32)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint32)
    return out


def f32_dijkstra(src, dst, w, V, root):
    adj = [[] for _ in range(V)]
    for s, t, x in zip(src.tolist(), dst.tolist(), w):
        adj[s].append((t, np.float32(x)))
    dist = np.full(V, np.float32(1e9), dtype=np.float32)
    dist[root] = np.float32(0)
    heap, done = [(0.0, root)], np.zeros(V, bool)
    while heap:
        _, u = heapq.heappop(heap)
        if done[u]:
            continue
        done[u] = True
        for v, x in adj[u]:
            nd = np.float32(dist[u] + x)
            if nd < dist[v]:
                dist[v] = nd
                heapq.heappush(heap, (float(nd), v))
    return dist


def weak_components(src, dst, V):
    A = coo_matrix((np.ones(len(src)), (src, dst)), shape=(V, V))
    _, lab = csgraph.connected_components(A, directed=True, connection="weak")
    first = np.full(lab.max() + 1, V, dtype=np.int64)
    np.minimum.at(first, lab, np.arange(V))
    return first[lab].astype(np.uint32)


def pagerank(src, dst, V, iters=20, d=0.85):
    A = np.zeros((V, V))
    np.add.at(A, (dst.astype(np.int64), src.astype(np.int64)), 1.0)
    od = np.bincount(src, minlength=V).astype(np.float64)
    x = np.where(od > 0, 1.0 / np.maximum(od, 1), 1.0)
    for it in range(iters):
        y = (1 - d) + d * (A @ x)
        if it != iters - 1:
            y = np.where(od > 0, y / np.maximum(od, 1), y)
        x = y
    return x


def brandes(src, dst, V, root, depth):
    """Single-source Brandes over arcs (a duplicate arc is two shortest-path
    steps): sigma by levels, then delta(v) = sum over arcs v->w one level down
    of sigma(v)/sigma(w) * (1 + delta(w))."""
    sigma = np.zeros(V)
    sigma[root] = 1.0
    reached = depth != UNSET
    dv = depth.astype(np.int64)
    s, t = src.astype(np.int64), dst.astype(np.int64)
    step = reached[s] & reached[t] & (dv[t] == dv[s] + 1)
    s, t = s[step], t[step]
    top = int(dv[reached].max())
    for L in range(1, top + 1):
        m = dv[t] == L
        np.add.at(sigma, t[m], sigma[s[m]])
    delta = np.zeros(V)
    for L in range(top, 0, -1):
        m = dv[t] == L
        np.add.at(delta, s[m], sigma[s[m]] / sigma[t[m]] * (1.0 + delta[t[m]]))
    dep = np.where(reached, delta, np.nan)
    return dep, sigma


def fixture(name, src, dst, w, V, roots):
    out = {"src": src, "dst": dst, "w": w, "V": np.array([V]), "roots": np.array(roots, np.uint32)}
    depth = np.stack([bfs_depth(src, dst, V, r) for r in roots])
    levels = [np.bincount(d[d != UNSET]).astype(np.uint32) for d in depth]
    out["bfs_depth"] = depth
    out["bfs_levels"] = np.concatenate(levels)
    out["bfs_levels_off"] = np.cumsum([0] + [len(x) for x in levels]).astype(np.int64)
    out["sssp_dist"] = np.stack([f32_dijkstra(src, dst, w, V, r) for r in roots])
    dep, sig = zip(*[brandes(src, dst, V, r, d) for r, d in zip(roots, depth)])
    out["bc_dep"] = np.stack(dep)
    out["bc_sigma"] = np.stack(sig)
    out["cc_label"] = weak_components(src, dst, V)
    out["pagerank"] = pagerank(src, dst, V)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "V", V, "E", len(src), "roots", roots)


def main():
    s, d, w = rmat_edges(10, 8, seed=2024)
    V = 1 << 10
    od = np.bincount(s, minlength=V)
    rng = np.random.default_rng(7)
    roots = [int(x) for x in rng.choice(np.nonzero(od)[0], 3, replace=False)]
    fixture("rmat10", s, d, w, V, roots)
    s, d, w, V = hub_graph()
    fixture("hub", s, d, w, V, [0, 1501, 1615])
    s, d, w, V = edge_case_graph()
    fixture("edge_cases", s, d, w, V, [0, 5, 76])


if __name__ == "__main__":
    main()

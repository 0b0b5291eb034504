"""Full-size case: SSSP from 8 seeded roots on a weighted RMAT-26 graph
(67,108,864 vertices, 1,073,741,824 arcs, float32 weights in (0,1]) and
PageRank (20 iterations, FP64) on the same topology.  The oracle cannot run
at this size, so parity rests on properties that hold at any size:

* SSSP (sssp.cpp:22-105): the distances are a fixed point of the float32
  relaxation the reference iterates -- no arc u->v has fl(d[u] + w) < d[v],
  and every reached vertex other than the root has an in-arc with
  fl(d[u] + w) == d[v]; the root is at 0 (the oracle's checker over the
  CSC, with OpenMP); forced push and forced pull (different kernels) give
  bit-identical distances; m_reach is the out-degree sum of the reached;
* PageRank (pagerank.cpp:24-129): the destination-bin plan and the plain
  pull (independent kernels) agree within 1e-9 relative.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SCALE = 26
SEED = 26


def say(*a):
    print("[gpu]", *a, flush=True)


def test_sssp_rmat26_fixed_point(gg):
    V = 1 << SCALE
    g = gg.Graph.rmat(SCALE, 16, seed=SEED, permute=True, weighted=True)
    assert g.arcs == 16 * V
    od = g.out_degree().astype(np.int64)
    rng = np.random.default_rng(SEED)
    roots = [int(x) for x in rng.choice(np.nonzero(od)[0], 8, replace=False)]
    idx, src, w = g.csc()  # not relabelled: the device arrays as they are
    dist = gg.DeviceArray(V, np.float32)
    first = None
    for r in roots:
        say("sssp rmat26 root", r)
        st = g.sssp_into(r, dist)
        d = dist.numpy()
        if first is None:
            first = d.view(np.uint32).copy()
        assert oracle.sssp_fixed_point_violations(idx, src, w, d, r) == 0, r
        assert st.m_reach == int(od[d < 1e9].sum())
    for mode in (1, 2):
        say("sssp rmat26 forced mode", mode)
        with gg.options(force_mode=mode):
            g.sssp_into(roots[0], dist)
        assert np.array_equal(dist.numpy().view(np.uint32), first), mode
    dist.free()
    g.close()


def test_pagerank_rmat26_plan_vs_pull(gg):
    V = 1 << SCALE
    g = gg.Graph.rmat(SCALE, 16, seed=SEED, permute=True, reorder=True)
    pr = gg.DeviceArray(V, np.float64)
    say("pagerank rmat26 plan")
    st = g.pagerank_into(pr, 20)
    assert st.iterations == 20
    plan = pr.numpy()
    say("pagerank rmat26 pull")
    with gg.options(pr_hot_sources=0):
        g.pagerank_into(pr, 20)
    pull = pr.numpy()
    assert np.all(np.isfinite(plan)) and np.all(plan > 0)
    assert np.max(np.abs(plan - pull) / np.abs(pull)) <= 1e-9
    pr.free()
    g.close()

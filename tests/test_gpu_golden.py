"""The HIP path (through the C ABI) against the committed golden fixtures
(tests/golden), computed by routines independent of both the engine and the
oracle; file ids and relabelled graphs."""
import numpy as np
import pytest

from golden_common import NAMES, close, levels, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(gg, name, reorder):
    f = load(name)
    V = f["V"]
    print("[gpu] golden", name, reorder, flush=True)
    g = gg.Graph.build(f["edges"], V, reorder=reorder)
    gu = gg.Graph.build(f["edges_unweighted"], V, reorder=reorder)
    for i, r in enumerate(f["roots"].tolist()):
        _, depth, fr, _ = g.bfs(r)
        assert np.array_equal(depth, f["bfs_depth"][i])
        assert np.array_equal(fr, levels(f, i))
        dist, _, _ = g.sssp(r)
        assert np.array_equal(dist.view(np.uint32), f["sssp_dist"][i].view(np.uint32))
        dep, inv, _, _ = gu.bc(r)
        close(dep, f["bc_dep"][i])
        reached = f["bfs_depth"][i] != 0xFFFFFFFF
        close(1.0 / inv[reached], f["bc_sigma"][i][reached])
    pr, _, _ = gu.pagerank(20)
    close(pr, f["pagerank"])
    label, _, _ = gg.Graph.build(f["edges_unweighted"], V, undirected=True, reorder=reorder).cc()
    assert np.array_equal(label, f["cc_label"])
    # batched: every root's row equals its fixture
    d, _ = gg.bfs_batch(g, f["roots"])
    assert np.array_equal(d, f["bfs_depth"])
    dd, _ = gg.sssp_batch(g, f["roots"])
    assert np.array_equal(dd.view(np.uint32), f["sssp_dist"].view(np.uint32))

"""The oracle against the committed golden fixtures (tests/golden), which
were computed by routines independent of it (numpy / scipy / float32
Dijkstra / textbook Brandes)."""
import numpy as np
import pytest

import oracle
from golden_common import NAMES, close, levels, load


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_golden(name):
    f = load(name)
    V = f["V"]
    o = oracle.Graph(f["edges"], V)
    for i, r in enumerate(f["roots"].tolist()):
        _, depth, fr = o.bfs(r)
        assert np.array_equal(depth, f["bfs_depth"][i])
        assert np.array_equal(fr, levels(f, i))
        dist, _ = o.sssp(r)
        assert np.array_equal(dist.view(np.uint32), f["sssp_dist"][i].view(np.uint32))
    ou = oracle.Graph(f["edges_unweighted"], V)
    for i, r in enumerate(f["roots"].tolist()):
        dep, inv, _ = ou.bc(r)
        close(dep, f["bc_dep"][i])
        reached = f["bfs_depth"][i] != 0xFFFFFFFF
        close(1.0 / inv[reached], f["bc_sigma"][i][reached])
    pr, _ = ou.pagerank(20)
    close(pr, f["pagerank"])
    label, _ = oracle.Graph(f["edges_unweighted"], V, undirected=True).cc()
    assert np.array_equal(label, f["cc_label"])


def test_planted_mistakes_fail_the_fixture():
    # a changed depth, distance bit or label must not pass the comparison
    f = load("rmat10")
    d = f["bfs_depth"][0].copy()
    v = int(np.nonzero(d == 2)[0][0])
    d[v] = 3
    assert not np.array_equal(d, f["bfs_depth"][0])
    x = f["sssp_dist"][0].copy()
    x[v] = np.nextafter(x[v], np.float32(2))
    assert not np.array_equal(x.view(np.uint32), f["sssp_dist"][0].view(np.uint32))
    with pytest.raises(AssertionError):
        y = f["pagerank"].copy()
        y[v] *= 1 + 1e-7
        close(y, f["pagerank"])

"""The threshold autotuner (gg_tune): chosen values come from the candidate
lists, results are the same before and after tuning (no result depends on
the options), apply=0 leaves the options as they were, bad arguments are
refused."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

BFS_KNOBS = {"switch_divisor": {20, 50, 100, 200, 400, 1000}, "bfs_alpha": {0, 8, 16, 24, 32, 64},
             "push_big_max": {256, 1024, 4096}, "push_small_max": {16, 32, 64},
             "bfs_pull_waves": {4096, 8192, 16384}}
PR_KNOBS = {"pr_hot_sources": {0, 1 << 17, 1 << 18, 1 << 19}, "pr_bin_bits": {12, 13, 14},
            "pr_range_bits": {12, 13, 14}}


@pytest.fixture
def saved_options(gg):
    names = list(BFS_KNOBS) + list(PR_KNOBS)
    before = {n: gg.get_option(n) for n in names}
    yield before
    for n, v in before.items():
        gg.set_option(n, v)


def test_tune_bfs(gg, saved_options):
    scale = 14
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=81)
    g = gg.Graph.build(e, V, reorder=True)
    o = oracle.Graph(e, V)
    root = int(np.argmax(o.out_degree()))
    _, ref, ref_fr = o.bfs(root)
    r = gg.tune(g, "bfs", samples=4, apply=False)
    assert set(r["options"]) == set(BFS_KNOBS)
    for n, v in r["options"].items():
        assert v in BFS_KNOBS[n] or v == saved_options[n], (n, v)
        assert gg.get_option(n) == saved_options[n]  # not applied
    assert r["trials"] >= 1 + sum(len(s) for s in BFS_KNOBS.values()) - len(BFS_KNOBS)
    assert 0 < r["tuned_ms"] <= r["baseline_ms"]
    r = gg.tune(g, "bfs", samples=4, apply=True)
    for n, v in r["options"].items():
        assert gg.get_option(n) == v
    _, depth, fr, _ = g.bfs(root)
    assert np.array_equal(depth, ref)
    assert np.array_equal(fr, ref_fr)


def test_tune_pagerank(gg, saved_options):
    scale = 13
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=82)
    g = gg.Graph.build(e, V, reorder=True)
    r = gg.tune(g, "pagerank", samples=2, apply=True)
    assert set(r["options"]) == set(PR_KNOBS)
    for n, v in r["options"].items():
        assert gg.get_option(n) == v
        assert v in PR_KNOBS[n] or v == saved_options[n], (n, v)
    pr, _, _ = g.pagerank(10)
    ref, _ = oracle.Graph(e, V).pagerank(10)
    assert np.allclose(pr, ref, rtol=1e-9, atol=0)


def test_tune_bad_arguments(gg, saved_options):
    g = gg.Graph.build(oracle.edges_from_pairs([(0, 1), (1, 2)]), 3)
    with pytest.raises(gg.GGError):
        gg.tune(g, 7)
    with pytest.raises(gg.GGError):
        gg.tune(g, "bfs", samples=0)
    with pytest.raises(gg.GGError):
        gg.tune(g, "bfs", samples=65)
    for n, v in saved_options.items():
        assert gg.get_option(n) == v

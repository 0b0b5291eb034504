"""The five toolkits written against the C++ facade (geminigraph_amd/core,
Graph<EdgeData> with process_edges / process_vertices over device callbacks)
run as separate programs on a binary edge-list file, as the reference's do
(README: `./toolkits/X [path] [vertices] [iterations|root]`).  Their dumped
results are checked against the oracle with the same bar as the C-ABI path."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
from test_gpu_parity import assert_rel, check_parents, roots_of

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TK = os.path.join(ROOT, "geminigraph_amd", "toolkits")


def run_toolkit(name, args, dump_dir):
    exe = os.path.join(TK, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built: make -C geminigraph_amd toolkits")
    env = dict(os.environ, GG_DUMP=str(dump_dir), GG_RUNS="1")
    print(f"[gpu] toolkit {name} {args}", flush=True)
    p = subprocess.run([exe] + [str(a) for a in args], env=env, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return p.stdout


def load(dump_dir, name, dtype):
    return np.fromfile(os.path.join(dump_dir, name + ".bin"), dtype=dtype)


def active_lines(out):
    return [int(m.group(1)) for m in re.finditer(r"active\(\d+\)>=(\d+)", out)]


SCALE = 12
V = 1 << SCALE


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("graphs")
    e = oracle.rmat_edges(SCALE, seed=61)
    ew = oracle.rmat_edges(SCALE, seed=61, weighted=True)
    e.tofile(d / "g.bin")
    ew.tofile(d / "gw.bin")
    return d, e, ew


def test_pagerank_toolkit(files, tmp_path):
    d, e, _ = files
    out = run_toolkit("pagerank", [d / "g.bin", V, 20], tmp_path)
    pr = load(tmp_path, "pagerank", np.float64)
    ref, deltas = oracle.Graph(e, V).pagerank(20)
    assert_rel(pr, ref)
    printed = [float(x) for x in re.findall(r"delta\(\d+\)=([0-9.eE+-]+)", out)]
    assert np.allclose(printed, deltas[:20], rtol=1e-5, atol=1e-6)


def test_bfs_toolkit(files, tmp_path):
    d, e, _ = files
    o = oracle.Graph(e, V)
    for root in roots_of(o.out_degree(), 2, 61):
        out = run_toolkit("bfs", [d / "g.bin", V, root], tmp_path)
        parent = load(tmp_path, "parent", np.uint32)
        _, depth, fr = o.bfs(root)
        assert active_lines(out) == fr.tolist()
        assert np.array_equal(parent < V, depth != 0xFFFFFFFF)
        check_parents(e, V, root, parent, depth)
        assert f"found_vertices = {int((depth != 0xFFFFFFFF).sum())}" in out


def test_cc_toolkit(files, tmp_path):
    d, e, _ = files
    out = run_toolkit("cc", [d / "g.bin", V], tmp_path)
    label = load(tmp_path, "label", np.uint32)
    ref, _ = oracle.Graph(e, V, undirected=True).cc()
    assert np.array_equal(label, ref)
    assert f"components = {len(np.unique(ref))}" in out


def test_sssp_toolkit(files, tmp_path):
    d, _, ew = files
    o = oracle.Graph(ew, V)
    root = roots_of(o.out_degree(), 1, 62)[0]
    run_toolkit("sssp", [d / "gw.bin", V, root], tmp_path)
    dist = load(tmp_path, "distance", np.float32)
    ref, _ = o.sssp(root)
    assert np.array_equal(dist.view(np.uint32), ref.view(np.uint32))


def test_bc_toolkit(files, tmp_path):
    d, e, _ = files
    o = oracle.Graph(e, V)
    root = roots_of(o.out_degree(), 1, 63)[0]
    out = run_toolkit("bc", [d / "g.bin", V, root], tmp_path)
    dep = load(tmp_path, "dependencies", np.float64)
    inv = load(tmp_path, "inv_num_paths", np.float64)
    rdep, rinv, rfr = o.bc(root)
    assert active_lines(out) == rfr.tolist()
    assert_rel(inv, rinv)
    assert_rel(dep, rdep)


# A graph whose lists take the facade's long-list paths: vertex HUB has 5000
# out-arcs (sparse slots above 2048 entries, spread over the grid), SINK has
# 5000 in-arcs (dense signals in parts of 2048), joined to an RMAT part.
HV = 1 << 14
HUB, SINK = 8192, 14000


@pytest.fixture(scope="module")
def hub_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("hub")
    r = oracle.rmat_edges(13, seed=64)
    pairs = np.stack([r["src"], r["dst"]], axis=1).astype(np.int64).tolist()
    pairs += [(0, HUB)] + [(HUB, HUB + 1 + i) for i in range(5000)]
    pairs += [(HUB + 1 + i, SINK) for i in range(5000)] + [(SINK, SINK + 1 + i) for i in range(99)]
    pairs += [(SINK + 50, 3)]
    e = oracle.edges_from_pairs(pairs)
    w = ((np.arange(len(e)) % 13) + 1).astype(np.float32) / 16
    ew = oracle.edges_from_pairs(np.stack([e["src"], e["dst"]], 1), w)
    e.tofile(d / "h.bin")
    ew.tofile(d / "hw.bin")
    return d, e, ew


@pytest.mark.parametrize("root", [0, HUB])
def test_bfs_toolkit_long_lists(hub_files, tmp_path, root):
    d, e, _ = hub_files
    o = oracle.Graph(e, HV)
    out = run_toolkit("bfs", [d / "h.bin", HV, root], tmp_path)
    parent = load(tmp_path, "parent", np.uint32)
    _, depth, fr = o.bfs(root)
    assert active_lines(out) == fr.tolist()
    check_parents(e, HV, root, parent, depth)


def test_pagerank_sssp_bc_toolkits_long_lists(hub_files, tmp_path):
    d, e, ew = hub_files
    run_toolkit("pagerank", [d / "h.bin", HV, 20], tmp_path)
    ref, _ = oracle.Graph(e, HV).pagerank(20)
    assert_rel(load(tmp_path, "pagerank", np.float64), ref)
    run_toolkit("sssp", [d / "hw.bin", HV, 0], tmp_path)
    rdist, _ = oracle.Graph(ew, HV).sssp(0)
    assert np.array_equal(load(tmp_path, "distance", np.float32).view(np.uint32),
                          rdist.view(np.uint32))
    out = run_toolkit("bc", [d / "h.bin", HV, 0], tmp_path)
    rdep, rinv, rfr = oracle.Graph(e, HV).bc(0)
    assert active_lines(out) == rfr.tolist()
    assert_rel(load(tmp_path, "inv_num_paths", np.float64), rinv)
    assert_rel(load(tmp_path, "dependencies", np.float64), rdep)
    out = run_toolkit("cc", [d / "h.bin", HV], tmp_path)
    ref_l, _ = oracle.Graph(e, HV, undirected=True).cc()
    assert np.array_equal(load(tmp_path, "label", np.uint32), ref_l)

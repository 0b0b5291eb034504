"""Vertex arrays to and from files (dump / restore / gather,
core/graph.hpp:253-324 with one partition) and the cached binary CSR/CSC
(gg_graph_save / gg_graph_open): round trips are exact, the opened graph
gives the same results as the built one, bad files are refused."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _same_graph(a, b):
    assert (a.vertices, a.edges, a.arcs, a.weighted, a.symmetric) == \
        (b.vertices, b.edges, b.arcs, b.weighted, b.symmetric)
    assert a.reordered == b.reordered
    for x, y in zip(a.csr(), b.csr()):
        assert (x is None and y is None) or np.array_equal(x, y)
    for x, y in zip(a.csc(), b.csc()):
        assert (x is None and y is None) or np.array_equal(x, y)
    assert np.array_equal(a.out_degree(), b.out_degree())
    assert np.array_equal(a.in_degree(), b.in_degree())


@pytest.mark.parametrize("reorder", [False, True])
def test_graph_save_open_directed(gg, tmp_path, reorder):
    scale = 13
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=91, weighted=True)
    g = gg.Graph.build(e, V, reorder=reorder)
    p = tmp_path / "g.ggcsr"
    g.save(p)
    h = gg.Graph.open(p)
    _same_graph(g, h)
    root = int(np.argmax(g.out_degree()))
    assert np.array_equal(g.bfs(root)[1], h.bfs(root)[1])
    assert np.array_equal(g.sssp(root)[0].view(np.uint32), h.sssp(root)[0].view(np.uint32))
    # the bin plan sums with LDS atomics: FP64 order varies from run to run
    assert np.allclose(g.pagerank(10)[0], h.pagerank(10)[0], rtol=1e-12, atol=0)
    ref, _ = oracle.Graph(e, V).pagerank(10)
    assert np.allclose(h.pagerank(10)[0], ref, rtol=1e-9, atol=0)


def test_graph_save_open_undirected_and_transposed(gg, tmp_path):
    e = oracle.rmat_edges(12, 8, seed=92)
    V = 1 << 12
    u = gg.Graph.build(e, V, undirected=True)
    u.save(tmp_path / "u.ggcsr")
    uu = gg.Graph.open(tmp_path / "u.ggcsr")
    _same_graph(u, uu)
    assert np.array_equal(u.cc()[0], uu.cc()[0])
    d = gg.Graph.build(e, V)
    d.transpose()
    d.save(tmp_path / "t.ggcsr")
    t = gg.Graph.open(tmp_path / "t.ggcsr")
    assert t.info().transposed == 1
    _same_graph(d, t)
    root = int(np.argmax(d.out_degree()))
    assert np.array_equal(d.bfs(root)[1], t.bfs(root)[1])


def test_graph_open_refuses_bad_files(gg, tmp_path):
    e = oracle.rmat_edges(10, 8, seed=93)
    g = gg.Graph.build(e, 1 << 10)
    p = tmp_path / "g.ggcsr"
    g.save(p)
    data = p.read_bytes()
    (tmp_path / "short").write_bytes(data[:-4])
    (tmp_path / "magic").write_bytes(b"XXXXXXXX" + data[8:])
    # a neighbour id past the vertex count at the end of the file (the last
    # array is the in-direction's neighbour list of this unweighted graph)
    bad = bytearray(data)
    bad[-4:] = (5000).to_bytes(4, "little")
    (tmp_path / "nbr").write_bytes(bytes(bad))
    # offsets that decrease (the out-direction index follows the header and degrees)
    off = 48 + 4 * (1 << 10) + 8 * 5
    bad = bytearray(data)
    bad[off:off + 8] = (1 << 40).to_bytes(8, "little")
    (tmp_path / "index").write_bytes(bytes(bad))
    for name in ("short", "magic", "nbr", "index"):
        with pytest.raises(gg.GGError):
            gg.Graph.open(tmp_path / name)
    with pytest.raises(gg.GGError):
        gg.Graph.open(tmp_path / "missing")


def test_vertex_array_dump_restore_gather(gg, tmp_path):
    scale = 12
    V = 1 << scale
    e = oracle.rmat_edges(scale, 16, seed=94)
    g = gg.Graph.build(e, V, reorder=True)
    pr = gg.DeviceArray(V, np.float64)
    g.pagerank_into(pr, 5)
    host = g.gather_vertex_array(pr)
    assert np.array_equal(host, pr.numpy())
    p = tmp_path / "pr.bin"
    g.dump_vertex_array(pr, p)
    assert p.stat().st_size == 8 * V
    assert np.array_equal(np.fromfile(p, dtype=np.float64), host)
    back = gg.DeviceArray(V, np.float64)
    g.restore_vertex_array(back, p)
    assert np.array_equal(back.numpy(), host)
    # a file of another size is refused (the reference asserts)
    small = gg.DeviceArray(V, np.uint32)
    with pytest.raises(gg.GGError):
        g.restore_vertex_array(small, p)
    with pytest.raises(ValueError):
        g.dump_vertex_array(gg.DeviceArray(V - 1, np.uint32), tmp_path / "x.bin")


def test_device_arrays_reused_by_size(gg):
    # gg_free keeps a buffer for the next gg_malloc of the same size (up to
    # alloc_cache_bytes): the same address comes back (a size no other test
    # allocates, so no other kept buffer matches), other sizes get their own,
    # and a kept buffer carries new contents correctly
    V = (1 << 16) + 4321
    a = gg.DeviceArray(V, np.uint32)
    pa = a.ptr
    a.upload(np.arange(V, dtype=np.uint32))
    a.free()
    b = gg.DeviceArray(V, np.uint32)
    assert b.ptr == pa
    c = gg.DeviceArray(V + 1, np.uint32)
    assert c.ptr != pa
    want = np.arange(V, dtype=np.uint32)[::-1].copy()
    b.upload(want)
    assert np.array_equal(b.numpy(), want)
    b.free()
    c.free()
    with gg.options(alloc_cache_bytes=0):
        d = gg.DeviceArray(V + 2, np.uint32)
        d.free()  # released at once, not kept
        e = gg.DeviceArray(V + 2, np.uint32)
        assert np.array_equal(e.upload(np.ones(V + 2, np.uint32)).numpy(), np.ones(V + 2, np.uint32))
        e.free()
    # BFS into a reused output array matches BFS into a fresh one
    pairs = [(i, i + 1) for i in range(999)] + [(0, 500)]
    edges = oracle.edges_from_pairs(pairs)
    g = gg.Graph.build(edges, 1000)
    out1 = gg.DeviceArray(1000, np.uint32)
    g.bfs_into(0, out1)
    r1 = out1.numpy()
    out1.free()
    out2 = gg.DeviceArray(1000, np.uint32)
    g.bfs_into(0, out2)
    assert np.array_equal(out2.numpy(), r1)
    out2.free()
    g.close()

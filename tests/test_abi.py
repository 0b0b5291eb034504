"""The C-ABI library loads and exports every symbol the public header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

import geminigraph_amd as gg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(gg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("gg_graph_build", "gg_graph_load", "gg_pagerank", "gg_bfs", "gg_cc", "gg_sssp",
                 "gg_bc", "gg_rmat_generate", "gg_subset_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gg.LIB_PATH):
        pytest.fail("libgg.so is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(gg.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    assert declared_functions() <= set(gg.SIGNATURES), \
        sorted(declared_functions() - set(gg.SIGNATURES))


def test_last_error_and_options_without_gpu():
    L = gg.lib()
    assert L.gg_last_error() is not None
    # options are host state: settable without a device
    old = gg.get_option("switch_divisor")
    gg.set_option("switch_divisor", 25)
    assert gg.get_option("switch_divisor") == 25
    gg.set_option("switch_divisor", old)
    with pytest.raises(gg.GGError):
        gg.set_option("no_such_option", 1)
    with pytest.raises(gg.GGError):
        gg.set_option("force_mode", 7)
    # the relabeling's hub count: 0 (every vertex in degree order) up to 2^32 - 1
    assert gg.get_option("relabel_hubs") == 0
    with gg.options(relabel_hubs=1 << 19):
        assert gg.get_option("relabel_hubs") == 1 << 19
    assert gg.get_option("relabel_hubs") == 0
    for bad in (-1, 1 << 32):
        with pytest.raises(gg.GGError):
            gg.set_option("relabel_hubs", bad)
    assert gg.get_option("serve_bfs_workers") == 2
    for bad in (-1, 32):
        with pytest.raises(gg.GGError):
            gg.set_option("serve_bfs_workers", bad)
    # BFS list starts from the vertex records: a flag, on by default
    assert gg.get_option("bfs_record_starts") == 1
    with gg.options(bfs_record_starts=0):
        assert gg.get_option("bfs_record_starts") == 0
    assert gg.get_option("bfs_record_starts") == 1
    for bad in (-1, 2):
        with pytest.raises(gg.GGError):
            gg.set_option("bfs_record_starts", bad)
    # freed gg_malloc buffers kept for reuse: bytes, 0 disables, 4 GiB by default
    assert gg.get_option("alloc_cache_bytes") == 4 << 30
    with gg.options(alloc_cache_bytes=0):
        assert gg.get_option("alloc_cache_bytes") == 0
    with pytest.raises(gg.GGError):
        gg.set_option("alloc_cache_bytes", -1)


@pytest.mark.parametrize("bad", [-0.5, float("nan"), -float("inf")])
def test_negative_or_nan_weight_refused_without_gpu(bad):
    """A weighted host edge list with a negative or NaN weight is refused
    with GG_EINVAL before any device work: the SSSP kernels order f32
    distances by their bit patterns, which needs weights >= 0."""
    import oracle
    e = np.zeros(4, dtype=oracle.WEDGE_DTYPE)
    e["src"] = [0, 1, 2, 3]
    e["dst"] = [1, 2, 3, 0]
    e["w"] = [0.5, 1.0, bad, 0.25]
    with pytest.raises(gg.GGError) as ex:
        gg.Graph.build(e, 4)
    assert ex.value.code == -1 and "weight" in str(ex.value)


def test_weight_lattice_accepted_by_host_check():
    """Zero, -0.0, denormals and +inf pass the host check (they order
    correctly as bit patterns); the call then fails only for want of a GPU
    here, never with the weight message."""
    import oracle
    e = np.zeros(4, dtype=oracle.WEDGE_DTYPE)
    e["src"] = [0, 1, 2, 3]
    e["dst"] = [1, 2, 3, 0]
    e["w"] = [0.0, -0.0, 1e-45, float("inf")]
    try:
        gg.Graph.build(e, 4).close()
    except gg.GGError as ex:
        assert "weight" not in str(ex)

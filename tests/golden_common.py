"""Loads the golden fixtures of tests/golden (written by
tests/golden/make_golden.py from independent routines) and compares results
with them: depths, level sizes and labels bit-exact, SSSP float32 bit
patterns exact, PageRank and BC within 1e-9 relative (NaN positions equal)."""
import os

import numpy as np

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("rmat10", "hub", "edge_cases")


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"))  # plain arrays only
    f = {k: z[k] for k in z.files}
    f["V"] = int(f["V"][0])
    f["edges"] = oracle.edges_from_pairs(np.stack([f["src"], f["dst"]], 1), f["w"])
    f["edges_unweighted"] = oracle.edges_from_pairs(np.stack([f["src"], f["dst"]], 1))
    return f


def levels(f, i):
    a, b = f["bfs_levels_off"][i], f["bfs_levels_off"][i + 1]
    return f["bfs_levels"][a:b]


def close(got, ref, rel=1e-9):
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    m = ~np.isnan(ref)
    err = np.abs(got[m] - ref[m]) / np.maximum(np.abs(ref[m]), 1e-300)
    assert not len(err) or err.max() <= rel, err.max()

"""Shared inputs and comparisons for the parity tests (test/gpu-helpers.js is the model)."""
import numpy as np


def make_nodes(n, dims, size=200.0, seed=12345):
    """LCG nodes with f32-exact coordinates (test/gpu-helpers.js:42-53)."""
    s = seed
    out = np.zeros((n, 3))
    for i in range(n):
        for a in range(dims):
            s = (s * 1664525 + 1013904223) % 4294967296
            out[i, a] = np.float32((np.float32(s / 4294967296) - 0.5) * size)
    return out


LEVELS = {1: 24, 2: 24, 3: 20}  # depth of the tree's deepest cells (barnesHut.js:56)


def deepest_cells(pos, dims):
    """Each point's cell at the tree's deepest level, in d3's cube (floor(min), then a
    power-of-two extent holding the max: barnesHut.js:169-181), and that cell's width."""
    p = np.asarray(pos, np.float64)[:, :dims]
    lo = np.floor(p.min(0))
    size = 1.0
    while np.any(p.max(0) >= lo + size):
        size *= 2.0
    width = size / 2.0 ** LEVELS[dims]
    return np.floor((p - lo) / width).astype(np.int64), width


def separate_cells(pos, dims):
    """Moves every point that shares a deepest-level cell with an earlier one by a cell along
    x until none do.  The tree stops splitting at that level (as the reference's does,
    barnesHut.js:44-46) where d3 splits on; in 1D, where each level up scales a cell's
    strength by 1.41, one such pair changes every cell above it."""
    out = np.array(pos, np.float64)
    for _ in range(1000):
        q, width = deepest_cells(out, dims)
        _, first = np.unique(q, axis=0, return_index=True)
        if len(first) == len(out):
            return out
        dup = np.setdiff1d(np.arange(len(out)), first)
        out[dup, 0] = (out[dup, 0] + width).astype(np.float32)
    raise RuntimeError("could not separate the points")


def make_nodes_fast(n, dims, size=200.0, seed=12345):
    """Same distribution as make_nodes for large n (numpy PCG), f32-exact, every point in a
    deepest-level cell of its own (so without coincident points: those take the hash
    jiggle, which the FP64 oracle does not model)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 3))
    out[:, :dims] = ((rng.random((n, dims)) - 0.5) * size).astype(np.float32)
    return separate_cells(out, dims)


PARITY = []  # (name, rel-L2, extra) measured by the GPU parity tests of this session


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / (den if den > 0 else 1.0))


def close_count(actual, expected, rtol=1e-4, atol=1e-4):
    """Nodes that miss atol + rtol * max(1, |a|, |e|) on any field (gpu-helpers.js:63-79)."""
    tol = atol + rtol * np.maximum(1.0, np.maximum(np.abs(actual), np.abs(expected)))
    miss = np.any(np.abs(actual - expected) > tol, axis=1)
    assert np.all(np.isfinite(actual))
    return int(miss.sum())


def tree_outliers(n):
    return max(1, int(np.ceil(n / 1000)))

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


def make_nodes_fast(n, dims, size=200.0, seed=12345):
    """Same distribution as make_nodes for large n (numpy PCG), f32-exact and without
    coincident points (those take the hash jiggle, which the FP64 oracle does not model)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 3))
    out[:, :dims] = ((rng.random((n, dims)) - 0.5) * size).astype(np.float32)
    for _ in range(100):
        _, first = np.unique(out, axis=0, return_index=True)
        if len(first) == n:
            break
        dup = np.setdiff1d(np.arange(n), first)
        out[dup, 0] = np.nextafter(out[dup, 0].astype(np.float32), np.float32(np.inf)).astype(np.float64)
    return out


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

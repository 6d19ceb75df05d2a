"""Passes that share a launch leave the state they leave apart, bit for bit: a link pass of one
iteration right before a many-body pass runs in the tree build's first launch -- the
one-workgroup build up to 1,024 nodes, the box kernel (with the center mean) up to 32,768 --
in 1D, 2D and 3D, over ticks of link + manyBody + center."""
import numpy as np
import pytest

import d3fgpu
from bench import alphas, link_defaults, make_graph
from helpers import make_nodes

pytestmark = pytest.mark.gpu


def state_after(monkeypatch, fuse, pos, dims, forces, ticks):
    monkeypatch.setenv("D3FGPU_FUSE_LINK", "1" if fuse else "0")
    e = d3fgpu.Engine(0)
    try:
        e.set_nodes(pos, ndim=dims)
        e.set_forces(forces)
        a, _ = alphas(1.0, ticks)
        e.enqueue(a, 0.6)
        e.sync()
        assert e.stats().walk_errors == 0
        return e.readback()
    finally:
        e.close()


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("n", [2, 300, 1024, 1025, 5000, 32768, 32769])
def test_link_in_the_tree_build_launch_changes_nothing(monkeypatch, dims, n):
    pos = make_nodes(n, dims)
    src, dst = make_graph(n, 42)
    bias, strengths, distances = link_defaults(n, src, dst)
    forces = [d3fgpu.link(src, dst, strengths, distances, bias, iterations=1),
              d3fgpu.many_body(np.full(n, -30.0)), d3fgpu.center()]
    p0, v0 = state_after(monkeypatch, False, pos, dims, forces, 12)
    p1, v1 = state_after(monkeypatch, True, pos, dims, forces, 12)
    assert np.all(np.isfinite(p1))
    assert np.array_equal(p0, p1)
    assert np.array_equal(v0, v1)

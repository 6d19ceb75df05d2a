"""The many-body tree on the GPU against the bit-exact restatement of the reference GPU
algorithm (oracle/gputree.c): cube, Morton halves, sorted order, Karras topology, leaf
depths and d3 children bit for bit; centres and strengths against FP64 sums over the same
tree (barnesHut.js:14-19, src/manyBody.js:42-71)."""
import numpy as np
import pytest

import d3fgpu
import oracle as O
from helpers import make_nodes, make_nodes_fast

pytestmark = pytest.mark.gpu


def build_on_gpu(engine, pos, dims, strengths, theta=0.9):
    n = len(pos)
    engine.set_nodes(pos, ndim=dims)
    engine.set_forces([d3fgpu.many_body(strengths, theta=theta)])
    engine.apply_forces(1.0)
    engine.sync()
    return engine.debug_tree()


def check_tree(engine, pos, dims, strengths=None):
    n = len(pos)
    if strengths is None:
        strengths = np.full(n, -30.0)
    g = build_on_gpu(engine, pos, dims, strengths)
    o = O.gt_tree(pos.astype(np.float32), dims, strengths.astype(np.float32).astype(np.float64))
    print(f"n={n} dims={dims} cube={o['cube']} shift={o['shift']}", flush=True)
    np.testing.assert_array_equal(g["cube"], o["cube"])
    assert g["shift"] == o["shift"]
    np.testing.assert_array_equal(g["hi"], o["hi"])
    np.testing.assert_array_equal(g["lo"], o["lo"])
    np.testing.assert_array_equal(g["order"], o["order"])
    np.testing.assert_array_equal(g["leaf_depth"], o["leaf_depth"])
    if n > 1:
        np.testing.assert_array_equal(g["left"], o["left"])
        np.testing.assert_array_equal(g["right"], o["right"])
        np.testing.assert_array_equal(g["depth"], o["depth"])
        np.testing.assert_array_equal(g["kids"], o["kids"])
        vscale = float(np.max(np.abs(o["value"]))) if n > 1 else 1.0
        np.testing.assert_allclose(g["value"], o["value"].astype(np.float32), rtol=2e-7, atol=1e-9 * vscale)
        live = o["value"] != 0
        scale = np.max(np.abs(o["com"][live])) if live.any() else 1.0
        np.testing.assert_allclose(g["com"][live], o["com"][live], rtol=0, atol=1e-12 * max(1.0, scale))
    return g


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_500_nodes(engine_factory, dims):
    check_tree(engine_factory(), make_nodes(500, dims), dims)


@pytest.mark.parametrize("n", [1, 2, 3, 1023, 1024, 1025, 4097, 8192, 8193, 32769])
def test_sizes(engine_factory, n):
    """Either side of the one-workgroup build (1,024 points) and of the counting sort
    (8,192 points; 4,097 splits its codes over the most parts)."""
    check_tree(engine_factory(), make_nodes_fast(n, 3, 20 * np.cbrt(n)), 3)


@pytest.mark.parametrize("dims", [1, 2])
def test_one_workgroup_build_low_dims(engine_factory, dims):
    """1,024 points with 60 coincident: the one-workgroup build in 1D and 2D."""
    pos = make_nodes(1024, dims, 40)
    pos[100:160] = pos[7]
    check_tree(engine_factory(), pos, dims)


def test_mixed_signs(engine_factory):
    for dims in (2, 3):
        n = 1000
        s = np.array([-30.0 if i % 3 else 20.0 for i in range(n)])
        check_tree(engine_factory(), make_nodes(n, dims), dims, s)


def test_huge_cube(engine_factory):
    pos = make_nodes(1000, 3)
    pos[0, 0] = 1e5
    pos[1, 1] = -2e5
    check_tree(engine_factory(), pos, 3)


def test_far_from_origin(engine_factory):
    pos = make_nodes(1000, 3)
    pos[:, 0] = (pos[:, 0] + 5e4).astype(np.float32)
    pos[:, 1] = (pos[:, 1] - 3e4).astype(np.float32)
    pos[:, 2] = (pos[:, 2] + 1e4).astype(np.float32)
    check_tree(engine_factory(), pos, 3)


@pytest.mark.parametrize("dims,outlier", [(2, 120), (3, 120), (2, 2000), (3, 2000)])
def test_clumps_sharing_top_levels(engine_factory, dims, outlier):
    pos = make_nodes_fast(9000, dims, 2)
    pos[0, 0] = outlier
    check_tree(engine_factory(), pos, dims)


def test_coincident_points(engine_factory):
    pos = make_nodes(300, 3, 50)
    pos[:40] = [1.0, 2.0, 3.0]
    check_tree(engine_factory(), pos, 3)


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_long_runs_of_equal_high_halves(engine_factory, dims):
    """Clumps far smaller than a cell of the codes' high half (2,500, 150 and 40 points with
    distinct low halves, plus 40 coincident points): runs too long to be ordered in place,
    so their order comes from the long-run sort; order and topology stay bit-exact."""
    rng = np.random.default_rng(77 + dims)
    n = 6000
    pos = make_nodes_fast(n, dims)
    for start, count, centre, width in ((0, 2500, 10.0, 1e-3), (2500, 150, -40.0, 2e-3), (2650, 40, 60.0, 0.0),
                                        (2690, 40, -70.0, 5e-4)):
        blob = centre + width * rng.random((count, dims))
        pos[start:start + count, :dims] = blob.astype(np.float32).astype(np.float64)
    check_tree(engine_factory(), pos, dims)


@pytest.mark.parametrize("n", [200_000, 1_000_000])
def test_phyllotaxis_config_sizes(engine_factory, n):
    x, y, z = O.phyllotaxis(n, 3)
    pos = np.stack([x, y, z], 1).astype(np.float32).astype(np.float64)
    check_tree(engine_factory(), pos, 3)

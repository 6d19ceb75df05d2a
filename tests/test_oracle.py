"""The CPU oracle pinned against the reference's stored golden vectors and against what
the mathematics fixes (brute force, closed forms).  No GPU."""
import numpy as np
import pytest

import oracle as O
from helpers import make_nodes

GOLD = {
    # test/simulation-test.js:14-16 and test/gpu-simulation-test.js:41-43
    "phyllotaxis2d": [(7.0710678118654755, 0.0), (-9.03088751750192, 8.27303273571596),
                      (1.3823220809823638, -15.750847141167634)],
    # test/collide-test.js:34-36
    "collide": [(174.08616723117228, 66.51743051995625, 0.677346615710878, 0.26976816231064354),
                (-139.73606544743998, 95.69860503079263, -0.5300880593105067, 0.3545632444404687),
                (-34.9275994083864, -169.69384995620052, -0.1472585564003713, -0.6243314067511122)],
    # test/collide-test.js:67
    "jiggle": (-5.371433857229194, -2.6644608278592576),
}


def collide_desc(r, n, strength=1.0, it=1):
    radii = np.full(n, float(r))
    return O.ForceDesc(type=2, iterations=it, strength=strength, radii=O.P(radii)), radii


def test_phyllotaxis_golden():
    x, y, _ = O.phyllotaxis(3, 2)
    for i, (gx, gy) in enumerate(GOLD["phyllotaxis2d"]):
        assert abs(x[i] - gx) < 1e-12 and abs(y[i] - gy) < 1e-12


def test_collide_golden():
    x, y, _ = O.phyllotaxis(3, 2)
    s = O.Sim(np.stack([x, y, np.zeros(3)], 1), ndim=2)
    d, keep = collide_desc(1, 3)
    s.set_forces([d])
    s.tick(10)
    p, v = s.state()
    np.testing.assert_array_equal(p[:, :2], np.stack([x, y], 1))
    d, keep = collide_desc(100, 3)
    s.set_forces([d])
    s.tick(10)
    p, v = s.state()
    for i, g in enumerate(GOLD["collide"]):
        np.testing.assert_allclose([p[i, 0], p[i, 1], v[i, 0], v[i, 1]], g, rtol=1e-14, atol=1e-12)


def test_collide_respects_fixed_positions():
    """test/collide-test.js:40-50"""
    x, y, _ = O.phyllotaxis(3, 2)
    pos = np.stack([x, y, np.zeros(3)], 1)
    pos[0] = 0
    fix = np.full((3, 3), np.nan)
    fix[0, :2] = 0
    s = O.Sim(pos, fix=fix, ndim=2)
    d, keep = collide_desc(1, 3)
    s.set_forces([d])
    s.tick(10)
    d, keep = collide_desc(100, 3)
    s.set_forces([d])
    s.tick(10)
    p, v = s.state()
    assert p[0, 0] == 0 and p[0, 1] == 0 and v[0, 0] == 0 and v[0, 1] == 0


def test_collide_jiggle_golden():
    """Pins d3-quadtree's coincident chains, visit order and the LCG jiggle stream."""
    s = O.Sim(np.zeros((10, 3)), ndim=2)
    d, keep = collide_desc(1, 10)
    s.set_forces([d])
    s.tick(50)
    p, v = s.state()
    np.testing.assert_allclose(p[0, :2], GOLD["jiggle"], rtol=1e-14)
    assert abs(v[0, 0]) < 1e-6 and abs(v[0, 1]) < 1e-6


def test_center_golden():
    """test/center-test.js:4-13"""
    pos = np.array([[100.0, 0, 0], [200.0, 0, 0], [300.0, 0, 0]])
    s = O.Sim(pos, ndim=3)
    s.set_forces([O.ForceDesc(type=1, strength=1.0)])
    s.tick(1)
    p, _ = s.state()
    np.testing.assert_array_equal(p[:, 0], [-100.0, 0.0, 100.0])


def test_lcg_stream():
    """lcg.js: s0 = 1, s = (1664525 s + 1013904223) mod 2^32."""
    got = O.lcg_stream(3)
    s, want = 1, []
    for _ in range(3):
        s = (1664525 * s + 1013904223) % 2 ** 32
        want.append(s / 2 ** 32)
    np.testing.assert_array_equal(got, want)


def test_graph_generator():
    """examples/benchmark.html:86-93: n-1 tree links then ceil(n/2) random pairs."""
    src, dst = O.make_graph(1000, 42)
    assert len(src) == 1499
    assert np.all(dst[:999] < np.arange(1, 1000))
    s = 42
    s = (s * 1664525 + 1013904223) % 2 ** 32
    assert dst[0] == int(np.floor(s / 2 ** 32 * 1))


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_tree_with_tiny_theta_is_the_exact_sum(dims):
    """theta -> 0 opens every cell: d3's tree sum equals the brute-force sum."""
    pos = make_nodes(300, dims)
    s = np.where(np.arange(300) % 5 == 0, 20.0, -30.0)
    a, st = O.manybody(pos, np.zeros_like(pos), s, 0.7, theta2=1e-12, dmin2=4.0, ndim=dims)
    b = O.manybody_exact(pos, np.zeros_like(pos), s, 0.7, dmin2=4.0, ndim=dims)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)
    assert st[0] == 300 * 299


def test_tree_2d_is_an_unweighted_sum_of_cells():
    """In 2D c = 1: every cell's value is the plain sum of its strengths."""
    pos = make_nodes(400, 2)
    s = np.full(400, -30.0)
    g = O.gt_tree(pos.astype(np.float32), 2, s)
    # the root holds every point
    assert abs(g["value"][0] - (-30.0 * 400)) < 1e-9


def test_gpu_tree_restatement_invariants():
    """Karras topology: every point and every node but the root has exactly one parent;
    each node's d3 children are strictly deeper or points; leaf depth = parent depth + 1."""
    for dims in (1, 2, 3):
        pos = make_nodes(777, dims, 50)
        pos[:20] = pos[0]
        g = O.gt_tree(pos.astype(np.float32), dims)
        n = len(pos)
        refs = np.concatenate([g["left"], g["right"]])
        points = refs[refs >= 0x80000000] - 0x80000000
        nodes = refs[refs < 0x80000000]
        assert sorted(points.tolist()) == list(range(n))
        assert sorted(nodes.tolist()) == list(range(1, n - 1))
        for i in range(n - 1):
            for k in g["kids"][i]:
                if k == 0xffffffff:
                    continue
                if k >= 0x80000000:
                    assert g["leaf_depth"][k - 0x80000000] == g["depth"][i] + 1
                else:
                    assert g["depth"][k] >= g["depth"][i]
        order = g["order"]
        key = (g["hi"].astype(np.uint64) << np.uint64(32)) | g["lo"].astype(np.uint64)
        assert np.all(np.diff(key[order].astype(np.float64)) >= 0)


def test_gpu_tree_cells_match_d3_tree_sums():
    """FP64 values over the GPU tree equal d3's recursive accumulate on the root."""
    pos = make_nodes(500, 3)
    s = np.full(500, -30.0)
    g = O.gt_tree(pos.astype(np.float32), 3, s)
    lvl = g["leaf_depth"].astype(float)
    # value(root at its depth) = sum s c^(L - D)
    c = np.sqrt(0.5)
    want = np.sum(s[g["order"]] * c ** (lvl - g["depth"][0]))
    assert abs(g["value"][0] - want) < 1e-9 * abs(want)


def layout_rms(n, theta, dims, ticks=300):
    """The benchmark graph (examples/benchmark.html:86-93) laid out with link(30) +
    manyBody + center for 300 ticks from d3's start positions; RMS distance from 0."""
    x, y, z = O.phyllotaxis(n, dims)
    pos = np.stack([x, y, z], 1)
    src, dst = O.make_graph(n, 42)
    bias, st, dist = O.link_defaults(n, src, dst)
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, np.float64)
        keep.append(a)
        return O.P(a)

    links = np.ascontiguousarray(np.stack([src, dst], 1).reshape(-1), np.uint32)
    keep.append(links)
    d = [O.ForceDesc(type=3, iterations=1, count=len(src), links=O.P(links, O.C.c_uint32), strengths=arr(st),
                     distances=arr(dist), bias=arr(bias)),
         O.ForceDesc(type=4, strengths=arr(np.full(n, -30.0)), theta2=theta * theta, distance_min2=1.0,
                     distance_max2=np.inf),
         O.ForceDesc(type=1, strength=1.0)]
    s = O.Sim(pos, ndim=dims)
    s.set_forces(d)
    s.tick(ticks)
    p, _ = s.state()
    return float(np.sqrt(np.mean(np.sum(p ** 2, 1))))


@pytest.mark.parametrize("theta,want", [(0.9, 372), (0.5, 418)])
def test_layout_size_matches_published_3d(theta, want):
    """docs/differences.md:35-40: RMS radius of a 1,000-node random graph after 300 ticks,
    d3-force-3d CPU, 3D: 372 (theta 0.9), 418 (0.5).  Pins the octree Barnes-Hut with d3's
    sqrt(4/numChildren) damping, the sequential link force and the tick, in aggregate."""
    assert abs(layout_rms(1000, theta, 3) - want) < 1.0


def test_layout_size_matches_published_2d():
    """docs/differences.md:37: 2D CPU 552 (c = 1 in 2D, so theta does not shape it)."""
    assert abs(layout_rms(1000, 0.9, 2) - 552) < 1.0


@pytest.mark.parametrize("dims", [2, 3])
def test_collide_jacobi_cell_list_matches_brute_force(dims):
    """Above 4096 nodes the oracle's Jacobi collide finds partners through a cell list; it
    must give the all-pairs sum (here against numpy's brute force, summed in another order)."""
    n = 5000
    rng = np.random.default_rng(4)
    pos = np.zeros((n, 3))
    pos[:, :dims] = (rng.random((n, dims)) - 0.5) * 4 * n ** (1 / dims)
    vel = np.zeros((n, 3))
    vel[:, :dims] = rng.standard_normal((n, dims)) * 0.5
    radii = np.exp(np.log(2) + 0.5 * rng.standard_normal(n))
    x, y, z = (np.ascontiguousarray(pos[:, a]) for a in range(3))
    vx, vy, vz = (np.ascontiguousarray(vel[:, a]) for a in range(3))
    O.lib().or_collide_jacobi(n, dims, O.P(x), O.P(y), O.P(z), O.P(vx), O.P(vy), O.P(vz),
                              O.P(np.ascontiguousarray(radii)), 0.8, 1)
    got = np.stack([vx, vy, vz], 1) - vel
    snap = pos + vel
    want = np.zeros((n, 3))
    for lo in range(0, n, 500):
        d = snap[lo:lo + 500, None, :] - snap[None, :, :]
        l = (d * d).sum(-1)
        ri = radii[lo:lo + 500, None]
        r = ri + radii[None, :]
        hit = (l < r * r) & (l > 0)
        s = np.sqrt(np.where(hit, l, 1.0))
        f = np.where(hit, (r - s) / s * 0.8 * radii[None, :] ** 2 / (ri ** 2 + radii[None, :] ** 2), 0.0)
        want[lo:lo + 500, :dims] = (d[:, :, :dims] * f[:, :, None]).sum(1)
    assert np.count_nonzero(np.any(want != 0, 1)) > n // 4
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)

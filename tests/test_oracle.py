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


# ---- pins of the GPU-tree restatement (oracle/gputree.c) ------------------------------

LEAF, NONE = 0x80000000, 0xffffffff


def interleave(q, dims, bits):
    """Morton code by explicit bit interleaving, most significant level first and, within a
    level, z then y then x (d3's child index z << 2 | y << 1 | x): a method of its own,
    not the spread-by-masks of barnesHut.js:318-342."""
    m = 0
    for b in range(bits - 1, -1, -1):
        for a in range(dims - 1, -1, -1):
            m = (m << 1) | ((q[a] >> b) & 1)
    return m


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_cube_and_codes_known_answers(dims):
    """Hand-derived answers from barnesHut.js:169-181 (cover) and :343-366 (cellOf, keys).
    The points span [0, 1000] on every axis used, so the cube is origin 0, size 1024 and
    shift = LEVELS - 10 (cells of 2^-10 units in 3D, 2^-14 in 1D and 2D); the half of each
    cell coordinate above LEVELS/2 bits goes to the high code, the rest to the low one, each
    left-aligned in 32 bits."""
    levels = {1: 24, 2: 24, 3: 20}[dims]
    shift = levels - 10
    half = levels // 2
    cases = [(0.0, 0.0, 0.0), (1000.0, 1000.0, 1000.0)]
    # one unit on one axis sets bit `shift` of that axis' cell
    cases += [tuple(1.0 if a == k else 0.0 for a in range(3)) for k in range(dims)]
    # 512 units set its top bit; one cell of the deepest level its lowest
    cases += [tuple(512.0 if a == k else 0.0 for a in range(3)) for k in range(dims)]
    cases += [tuple(2.0 ** -shift if a == k else 0.0 for a in range(3)) for k in range(dims)]
    pos = np.zeros((len(cases), 3), np.float32)
    for i, c in enumerate(cases):
        pos[i, :dims] = c[:dims]
    g = O.gt_tree(pos, dims)
    np.testing.assert_array_equal(g["cube"], [0, 0, 0, 1024])
    assert g["shift"] == shift
    hb = half * dims  # bits of a half code
    for i, c in enumerate(cases):
        q = [min(int(np.floor(c[a] * 2.0 ** shift)), 2 ** levels - 1) for a in range(dims)]
        hi = interleave([v >> half for v in q], dims, half) << (32 - hb)
        lo = interleave([v & (2 ** half - 1) for v in q], dims, half) << (32 - hb)
        assert (g["hi"][i], g["lo"][i]) == (hi, lo), (dims, c)
    # the hand values themselves: axis a of the top level is bit 31 - (dims - 1 - a)
    for k in range(dims):
        assert g["hi"][2 + dims + k] == 1 << (31 - (dims - 1 - k))
        assert g["lo"][2 + 2 * dims + k] == 1 << (32 - hb + k)
        assert g["lo"][2 + k] == 0 and g["hi"][2 + k] == 1 << (32 - hb + (half - 10) * dims + k)


def test_cube_negative_shift_and_offset_origin():
    """A cube of 2^21 in 3D has shift -1: cells of 2 units counted from floor(x) - origin
    (barnesHut.js:343-351); an origin at floor(min) below 0 offsets every cell."""
    pos = np.array([[-3.2, 0.5, 7.0], [1.5e6, 5.0, 8.0], [3.7, 0.5, 7.0], [2.0 ** 20, 0.5, 7.0]], np.float32)
    g = O.gt_tree(pos, 3)
    np.testing.assert_array_equal(g["cube"], [-4, 0, 7, 2 ** 21])
    assert g["shift"] == -1
    # x = 3.7: floor 3, minus origin -4 -> 7 units -> cell 3 (0b11); y, z cell 0
    assert g["hi"][2] == 0 and g["lo"][2] == interleave([3, 0, 0], 3, 10) << 2
    # x = 2^20: 2^20 + 4 units -> cell 2^19 + 2: top bit of the high half, and 2 in the low half
    assert g["hi"][3] == (1 << 29) and g["lo"][3] == interleave([2, 0, 0], 3, 10) << 2


def gpu_tree_cells(pos, dims):
    """The d3 cells the restated GPU tree stands for: a binary node holds the cells from
    one below its parent's deepest cell down to its own; a point is its leaf cell.  A run of
    equal codes whose points coincide is d3's leaf one level below its parent's cell, and
    the nodes and points inside such a run stand for no cell of their own."""
    n = len(pos)
    g = O.gt_tree(pos.astype(np.float32), dims)
    levels = {1: 24, 2: 24, 3: 20}[dims]
    order = g["order"]
    if n == 1:
        return {(0, (int(order[0]),))}, g, {}
    left, right, depth = g["left"], g["right"], g["depth"]
    pts = {}

    def points(k):
        if k & LEAF:
            return [int(order[k & ~LEAF])]
        if k not in pts:
            pts[k] = points(int(left[k])) + points(int(right[k]))
        return pts[k]

    points(0)
    p32 = pos.astype(np.float32)
    out, rd = set(), {}

    def visit(k, top, inside):
        if k & LEAF:
            if not inside:
                s = k & ~LEAF
                out.add((int(g["leaf_depth"][s]), (int(order[s]),)))
            return
        d = int(depth[k])
        ps = pts[k]
        if d == levels and not inside:
            if all(np.array_equal(p32[ps[0], :dims], p32[j, :dims]) for j in ps):
                rd[k] = top
                out.add((top, tuple(sorted(ps))))
                inside = True
        if inside:
            visit(int(left[k]), 0, True)
            visit(int(right[k]), 0, True)
            return
        rd[k] = d
        for D in range(top, d + 1):
            out.add((D, tuple(sorted(ps))))
        visit(int(left[k]), d + 1, False)
        visit(int(right[k]), d + 1, False)

    visit(0, 0, False)
    return out, g, rd


def tiny_inputs(dims, seed):
    rng = np.random.default_rng(seed)
    from helpers import separate_cells
    for n in (1, 2, 5, 17, 40, 64):
        pos = np.zeros((n, 3))
        pos[:, :dims] = ((rng.random((n, dims)) - 0.5) * 50).astype(np.float32)
        if n >= 17:
            pos[: n // 3, :dims] = (pos[0, :dims] + rng.random((n // 3, dims)) * 0.01).astype(np.float32)
        pos = separate_cells(pos, dims)
        yield f"n={n}", pos
        if n >= 5:
            c = pos.copy()
            c[1: 1 + max(2, n // 4)] = c[n - 1]  # a coincident group (with node n - 1)
            yield f"n={n} coincident", c
    yield "all coincident", np.tile(np.array([[1.5, -2.0, 3.25]]), (7, 1))


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_gpu_tree_cells_are_d3_cells(dims):
    """On tiny inputs (clumps, coincident groups) the restated GPU tree -- cube, codes,
    order, Karras topology, leaf depths -- stands for exactly the cells of the d3 tree that
    tests/d3tree.py builds by d3-octree's own insertion, and each node's d3 children
    (collapse) are exactly its deepest cell's children there."""
    import d3tree
    for name, pos in tiny_inputs(dims, 10 + dims):
        want = set(d3tree.cells(pos.tolist(), dims))
        got, g, rd = gpu_tree_cells(pos, dims)
        assert got == want, (dims, name)
        kids_of = {}
        for D, ps in want:
            kids_of.setdefault(D, []).append(ps)
        n = len(pos)
        for k, d in rd.items():
            ps = set(gpu_point_set(g, k))
            if all(np.array_equal(pos[j, :dims], pos[next(iter(ps)), :dims]) for j in ps) and len(ps) > 1:
                continue  # a coincident group: d3's leaf, no children
            children = sorted(tuple(sorted(gpu_point_set(g, int(c)))) for c in g["kids"][k] if c != NONE)
            expect = sorted(c for c in kids_of.get(d + 1, []) if set(c) <= ps)
            assert children == expect, (dims, name, k)


def gpu_point_set(g, k):
    order = g["order"]
    if k & LEAF:
        return [int(order[k & ~LEAF])]
    return gpu_point_set(g, int(g["left"][k])) + gpu_point_set(g, int(g["right"][k]))


def mixed_input(n, dims, seed, coincident=True):
    from helpers import make_nodes_fast, separate_cells
    pos = make_nodes_fast(n, dims, 20 * np.cbrt(n), seed=seed)
    pos[: n // 8] = make_nodes_fast(n // 8, dims, 3.0, seed=seed + 1)
    pos = separate_cells(pos, dims)
    groups = []
    if coincident:
        for a, b, src in ((n // 8, n // 8 + 12, 0), (n // 2, n // 2 + 1, n // 2 + 3), (n - 3, n - 1, 5)):
            pos[a:b] = pos[src]
            groups += list(range(a, b)) + [src]
    return pos, np.array(sorted(set(groups)), int)


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("theta,signs,dmin,dmax", [(0.9, "neg", 1.0, np.inf), (0.5, "mixed", 1.0, np.inf),
                                                   (1.5, "neg", 1.0, np.inf), (0.9, "mixed", 3.0, 40.0)])
def test_oracle_manybody_matches_independent_d3(dims, theta, signs, dmin, dmax):
    """or_manybody (C, d3oracle.c) against tests/d3tree.py (Python, written separately from
    the d3 module algorithms and src/manyBody.js:42-115), per node, with clumps and
    coincident groups (their members meet the jiggle and are left out)."""
    import d3tree
    n = 400
    pos, members = mixed_input(n, dims, 3 + dims)
    s = -30.0 * (1 + 0.5 * np.sin(np.arange(n)))
    if signs == "mixed":
        s = np.where(np.arange(n) % 3 == 0, 25.0, s)
    s = s.astype(np.float32).astype(np.float64)
    ref, _ = O.manybody(pos, np.zeros_like(pos), s, 0.8, theta2=theta * theta, dmin2=dmin * dmin,
                        dmax2=dmax * dmax, ndim=dims)
    py = d3tree.manybody(pos.tolist(), s.tolist(), dims, alpha=0.8, theta=theta, dmin=dmin, dmax=dmax)
    keep = np.array([v is not None for v in py])
    keep[members] = False
    assert keep.sum() > n - 40
    py = np.array([v if v is not None else [0.0] * 3 for v in py])
    np.testing.assert_allclose(py[keep], ref[keep], rtol=1e-11, atol=1e-9)


def walk_gpu_tree(pos, dims, s, theta=0.9, dmin=1.0, dmax=np.inf):
    """The GPU walk's decisions over the restated GPU tree in FP64 (tree.hip cell_force and
    point_force): a node is tested at each cell depth from one below its parent's record
    depth down to its own record depth and taken at the first that passes, else opened; a
    point is exact.  Record depths follow the tree's rule for runs of equal codes."""
    g = O.gt_tree(pos.astype(np.float32), dims, s)
    _, _, rd = gpu_tree_cells(pos, dims)
    n = len(pos)
    order, kids, com, val = g["order"], g["kids"], g["com"], g["value"]
    levels = {1: 24, 2: 24, 3: 20}[dims]
    rdep = np.zeros(n - 1, int)
    for k in range(n - 1):
        rdep[k] = rd.get(k, -1)  # -1: inside a coincident run, always opened
        if k not in rd and g["depth"][k] == levels and not any(k == j for j in rd):
            pass
    size = float(g["cube"][3])
    c = np.sqrt(4.0 / 2 ** dims)
    pts = pos[order]
    th2, dmin2, dmax2 = theta * theta, dmin * dmin, dmax * dmax
    out = np.zeros((n, 3))
    for b in range(n):
        me = pts[b]
        acc = np.zeros(3)

        def cell(k, top):
            if top > rdep[k]:
                return False
            if val[k] == 0:
                return True
            d = com[k] - me
            d[dims:] = 0
            l = float(d @ d)
            w2, D = (size * 2.0 ** -top) ** 2, top
            while D <= rdep[k] and not (w2 < th2 * l):
                w2, D = w2 * 0.25, D + 1
            if D > rdep[k]:
                return False
            if l < dmax2:
                acc[:] += d * val[k] * c ** (rdep[k] - D) / (np.sqrt(dmin2 * l) if l < dmin2 else l)
            return True

        stack = [] if cell(0, 0) else [0]
        while stack:
            k = stack.pop()
            for q in kids[k]:
                if q == NONE:
                    continue
                if q & LEAF:
                    j = int(q) & 0x7fffffff
                    d = pts[j] - me
                    d[dims:] = 0
                    l = float(d @ d)
                    if j != b and 0 < l < dmax2:
                        acc[:] += d * s[order[j]] / (np.sqrt(dmin2 * l) if l < dmin2 else l)
                elif not cell(int(q), rdep[k] + 1):
                    stack.append(int(q))
        out[order[b]] = acc
    return out


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_gpu_tree_values_and_walk_rules_give_d3_forces(dims):
    """The restated GPU tree's FP64 cell values and centres (or_gt_cells) walked by the GPU's
    rules reproduce d3's forces node for node (tests/d3tree.py), coincident groups included:
    this pins the values the GPU's records are checked against, with their damping by
    depth, |value| weights and the chain depth of coincident points."""
    import d3tree
    n = 300
    pos, members = mixed_input(n, dims, 20 + dims)
    s = (-30.0 * (1 + 0.5 * np.sin(np.arange(n)))).astype(np.float32).astype(np.float64)
    s[::7] = 20.0
    got = walk_gpu_tree(pos, dims, s)
    py = d3tree.manybody(pos.tolist(), s.tolist(), dims)
    keep = np.array([v is not None for v in py])
    keep[members] = False
    py = np.array([v if v is not None else [0.0] * 3 for v in py])
    np.testing.assert_allclose(got[keep], py[keep], rtol=1e-9, atol=1e-9)


# ---- link: the GPU's Jacobi semantics --------------------------------------------------

def link_inputs(n, m, dims, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 3))
    pos[:, :dims] = (rng.random((n, dims)) - 0.5) * 100
    vel = np.zeros((n, 3))
    vel[:, :dims] = rng.standard_normal((n, dims))
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = ((src + 1 + rng.integers(0, n - 1, m)) % n).astype(np.uint32)
    st, di, bi = rng.random(m) + 0.1, 10 + 40 * rng.random(m), rng.random(m)
    return pos, vel, src, dst, st, di, bi


def run_link_jacobi(pos, vel, src, dst, st, di, bi, dims, it, alpha):
    x, y, z = (np.ascontiguousarray(pos[:, a]) for a in range(3))
    v = [np.ascontiguousarray(vel[:, a]).copy() for a in range(3)]
    O.lib().or_link_jacobi(len(pos), dims, O.P(x), O.P(y), O.P(z), O.P(v[0]), O.P(v[1]), O.P(v[2]), len(src),
                           O.P(src, O.C.c_uint32), O.P(dst, O.C.c_uint32), O.P(st), O.P(di), O.P(bi), it, alpha)
    return np.stack(v, 1)


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_link_jacobi_brute_force(dims):
    """or_link_jacobi against numpy (passes/link.js:25-46): every link sees one snapshot
    p + v; d *= (|d| - distance) / |d| * alpha * strength; target -= d * bias, source +=
    d * (1 - bias); `iterations` snapshots in turn."""
    n, m = 500, 1500
    pos, vel, src, dst, st, di, bi = link_inputs(n, m, dims, 8 + dims)
    got = run_link_jacobi(pos, vel, src, dst, st, di, bi, dims, 2, 0.7)
    v = vel.copy()
    for _ in range(2):
        snap = pos + v
        d = snap[dst] - snap[src]
        d[:, dims:] = 0
        l = np.sqrt((d * d).sum(1))
        f = (d * ((l - di) / l * 0.7 * st)[:, None])
        acc = np.zeros_like(v)
        np.add.at(acc, dst, -f * bi[:, None])
        np.add.at(acc, src, f * (1 - bi)[:, None])
        v = v + acc
    np.testing.assert_allclose(got, v, rtol=1e-12, atol=1e-12)


def test_link_single_spring_closed_form():
    """One link of length L along an axis: the target moves by -(L - distance) alpha k b,
    the source by +(L - distance) alpha k (1 - b) (src/link.js:34-54)."""
    for axis in range(3):
        pos = np.zeros((2, 3))
        pos[1, axis] = 50.0
        got = run_link_jacobi(pos, np.zeros((2, 3)), np.array([0], np.uint32), np.array([1], np.uint32),
                              np.array([0.4]), np.array([30.0]), np.array([0.25]), 3, 1, 0.5)
        want = np.zeros((2, 3))
        want[1, axis] = -(50 - 30) * 0.5 * 0.4 * 0.25
        want[0, axis] = (50 - 30) * 0.5 * 0.4 * 0.75
        np.testing.assert_allclose(got, want, rtol=1e-15, atol=1e-15)


def test_link_jacobi_equals_sequential_on_disjoint_links():
    """With no node shared by two links the Jacobi and d3's sequential order agree
    (test/gpu-forces-test.js:61-77)."""
    n = 400
    rng = np.random.default_rng(5)
    pos = (rng.random((n, 3)) - 0.5) * 100
    vel = rng.standard_normal((n, 3))
    perm = rng.permutation(n).astype(np.uint32)
    src, dst = perm[0::2].copy(), perm[1::2].copy()
    m = len(src)
    st, di, bi = rng.random(m) + 0.1, 10 + 40 * rng.random(m), rng.random(m)
    jac = run_link_jacobi(pos, vel, src, dst, st, di, bi, 3, 1, 0.9)
    x, y, z = (np.ascontiguousarray(pos[:, a]) for a in range(3))
    v = [np.ascontiguousarray(vel[:, a]).copy() for a in range(3)]
    r = O.C.c_uint64(1)
    O.lib().or_link_seq(n, 3, O.P(x), O.P(y), O.P(z), O.P(v[0]), O.P(v[1]), O.P(v[2]), m, O.P(src, O.C.c_uint32),
                        O.P(dst, O.C.c_uint32), O.P(st), O.P(di), O.P(bi), 1, 0.9, O.C.byref(r))
    np.testing.assert_allclose(jac, np.stack(v, 1), rtol=1e-13, atol=1e-13)


# ---- the GPU jiggle hash ----------------------------------------------------------------

def test_gpu_jiggle_hash():
    """or_gpu_jiggle is wgsl.js:28-39 (PCG-RXS-M-XS as a hash, chained over step * 3 + axis,
    seed, b, a): against a numpy uint32 statement of it, and the distribution d3's
    jiggle has, (random() - 0.5) * 1e-6: [-5e-7, 5e-7), mean 0, uniform."""
    def pcg(v):
        v = np.asarray(v, np.uint64)
        s = (v * np.uint64(747796405) + np.uint64(2891336453)) & np.uint64(0xffffffff)
        w = (((s >> ((s >> np.uint64(28)) + np.uint64(4))) ^ s) * np.uint64(277803737)) & np.uint64(0xffffffff)
        return (w >> np.uint64(22)) ^ w

    rng = np.random.default_rng(3)
    a, b = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64), rng.integers(0, 2 ** 32, 4000, dtype=np.uint64)
    axis, step, seed = rng.integers(0, 3, 4000), rng.integers(0, 10 ** 6, 4000), rng.integers(0, 2 ** 32, 4000)
    h = pcg(a ^ pcg(b ^ pcg(seed.astype(np.uint64) ^ pcg((step * 3 + axis) & 0xffffffff))))
    want = ((h.astype(np.float32) * np.float32(1 / 4294967296) - np.float32(0.5)) * np.float32(1e-6))
    got = np.array([O.lib().or_gpu_jiggle(int(a[i]), int(b[i]), int(axis[i]), int(step[i]), int(seed[i]))
                    for i in range(4000)], np.float32)
    np.testing.assert_array_equal(got, want)
    assert got.min() >= -5e-7 and got.max() < 5e-7
    assert abs(got.mean()) < 1e-8
    hist, _ = np.histogram(got, bins=10, range=(-5e-7, 5e-7))
    assert hist.min() > 300

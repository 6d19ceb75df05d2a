"""Every force on the GPU against the FP64 oracle on identical f32-exact input
(test/gpu-forces-test.js is the model).  One application of the force list (no
integrator) is compared as the velocity increment it produces:

  manyBody (theta > 0)   d3's own Barnes-Hut over d3-octree/quadtree/binarytree, FP64:
                         rel-L2 <= 1e-4 over all nodes, and the per-node rule of
                         gpu-helpers.js:63-84 with its outlier allowance
  manyBody (theta = 0)   exact all-pairs sum, FP64
  link, collide          the GPU's Jacobi semantics restated in FP64 (passes/link.js,
                         passes/collide.js); d3's sequential order differs by design
                         (docs/differences.md:54-65)
  center, radial, x/y/z  closed forms, over several ticks with the integrator
"""
import numpy as np
import pytest

import d3fgpu
import oracle as O
from helpers import PARITY, close_count, make_nodes, make_nodes_fast, rel_l2, separate_cells, tree_outliers

pytestmark = pytest.mark.gpu


def gpu_dv(engine, pos, dims, forces, alpha=1.0, vel=None):
    engine.set_nodes(pos, vel=vel, ndim=dims)
    engine.set_forces(forces)
    engine.apply_forces(alpha)
    p, v = engine.readback()
    return p, v - (0 if vel is None else vel)


def oracle_manybody(pos, dims, strengths, alpha=1.0, theta=0.9, dmin=1.0, dmax=np.inf):
    s32 = np.asarray(strengths, np.float32).astype(np.float64)
    dv, st = O.manybody(pos, np.zeros_like(pos), s32, alpha, theta2=theta * theta, dmin2=dmin * dmin,
                        dmax2=dmax * dmax, ndim=dims)
    return dv, st


def report(name, err, **kw):
    extra = " ".join(f"{k}={v}" for k, v in kw.items())
    PARITY.append({"case": name, "rel_l2": err, **kw})
    print(f"[parity] {name}: rel-L2 {err:.3e} {extra}", flush=True)


# ---- many-body, Barnes-Hut ----------------------------------------------------------

@pytest.mark.parametrize("dims", [1, 2, 3])
def test_manybody_tree_500(engine_factory, dims):
    """test/gpu-forces-test.js:141-147 (which gives 1D rtol 2e-2: there every level up
    scales a cell by 1.41, and points closer than the deepest cell, which d3 splits
    further, change every cell above them; these points have no such pair)."""
    pos = make_nodes(500, dims)
    s = np.full(500, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, dims, [d3fgpu.many_body(s)])
    ref, _ = oracle_manybody(pos, dims, s)
    err = rel_l2(dv, ref)
    report(f"manyBody tree {dims}D n=500", err)
    assert close_count(dv, ref, rtol=1e-3) <= tree_outliers(500)
    assert err < 1e-4


@pytest.mark.parametrize("theta", [0.5, 1.5])
def test_manybody_tree_theta(engine_factory, theta):
    pos = make_nodes(2000, 3)
    s = np.full(2000, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, 3, [d3fgpu.many_body(s, theta=theta)])
    ref, _ = oracle_manybody(pos, 3, s, theta=theta)
    err = rel_l2(dv, ref)
    report(f"manyBody tree theta={theta}", err)
    assert close_count(dv, ref, rtol=1e-3) <= tree_outliers(2000)
    assert err < 1e-4


def test_manybody_tree_per_node_strength_dmin_dmax(engine_factory):
    pos = make_nodes(2000, 3)
    s = np.array([-10.0 - i % 7 for i in range(2000)])
    f = d3fgpu.many_body(s, distance_min=3, distance_max=60)
    _, dv = gpu_dv(engine_factory(), pos, 3, [f])
    ref, _ = oracle_manybody(pos, 3, s, dmin=3, dmax=60)
    err = rel_l2(dv, ref)
    report("manyBody tree dmin/dmax", err)
    assert close_count(dv, ref, rtol=1e-3) <= tree_outliers(2000)
    assert err < 1e-4


@pytest.mark.parametrize("dims", [2, 3])
def test_manybody_tree_mixed_signs(engine_factory, dims):
    pos = make_nodes(1000, dims)
    s = np.array([-30.0 if i % 3 else 20.0 for i in range(1000)])
    _, dv = gpu_dv(engine_factory(), pos, dims, [d3fgpu.many_body(s)])
    ref, _ = oracle_manybody(pos, dims, s)
    err = rel_l2(dv, ref)
    report(f"manyBody tree mixed signs {dims}D", err)
    assert close_count(dv, ref, rtol=1e-3) <= tree_outliers(1000)
    assert err < 1e-4


def test_manybody_tree_huge_cube(engine_factory):
    pos = make_nodes(1000, 3)
    pos[0, 0] = 1e5
    pos[1, 1] = -2e5
    s = np.full(1000, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, 3, [d3fgpu.many_body(s)])
    ref, _ = oracle_manybody(pos, 3, s)
    report("manyBody tree huge cube", rel_l2(dv, ref))
    assert close_count(dv, ref, rtol=1e-3) <= tree_outliers(1000)


def test_manybody_tree_far_from_origin(engine_factory):
    pos = make_nodes(1000, 3)
    pos[:, 0] = (pos[:, 0] + 5e4).astype(np.float32)
    pos[:, 1] = (pos[:, 1] - 3e4).astype(np.float32)
    pos[:, 2] = (pos[:, 2] + 1e4).astype(np.float32)
    s = np.full(1000, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, 3, [d3fgpu.many_body(s)])
    ref, _ = oracle_manybody(pos, 3, s)
    report("manyBody tree far from origin", rel_l2(dv, ref))
    assert close_count(dv, ref, rtol=1e-2) <= tree_outliers(1000)


@pytest.mark.parametrize("n", [2, 3, 1025, 8192, 8193, 32769])
def test_manybody_tree_sizes(engine_factory, n):
    pos = make_nodes_fast(n, 3, 20 * np.cbrt(n))
    s = np.full(n, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, 3, [d3fgpu.many_body(s)])
    ref, _ = oracle_manybody(pos, 3, s)
    err = rel_l2(dv, ref)
    report(f"manyBody tree n={n}", err)
    assert close_count(dv, ref, rtol=1e-3, atol=1e-3) <= tree_outliers(n)
    assert err < 1e-4


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("coincident", [False, True])
def test_manybody_tree_forces_do_not_depend_on_tiling(monkeypatch, dims, coincident):
    """Trees of up to 163840 points are walked in the exact mode: every point's force comes
    out bit for bit the same whatever the tile size (points per wave) and whichever mode
    (one lane per point, or packed (point, child) items) ran each opening -- packed limit 0
    runs every opening one lane per point, 64 every opening packed -- and in the walk that
    gives each point a group of lanes of its own (solo, the default for small trees).  A
    clump makes deep cell chains; 40 copies of one point make a deep run of equal codes,
    which the tree gives d3's depth for such a group (one leaf where it parts from the
    other points).  Against the FP64 oracle over every node but the copies, whose push
    from each other is the hash jiggle (d3 draws from its random source instead)."""
    n = 6000
    pos = make_nodes_fast(n, dims, 20 * np.cbrt(n))
    pos[: n // 10] = make_nodes_fast(n // 10, dims, 3.0)
    pos = separate_cells(pos, dims)
    members = np.zeros(n, bool)
    if coincident:
        pos[n // 10: n // 10 + 40] = pos[0]
        members[n // 10: n // 10 + 40] = members[0] = True
    s = np.full(n, -30.0) * (1 + 0.5 * np.sin(np.arange(n)))
    out = []
    monkeypatch.setenv("D3FGPU_WALK_WIDE", "0")
    monkeypatch.setenv("D3FGPU_WALK_GROUP", "1")
    for tile, packed, solo in [(64, 24, 0), (64, 0, 0), (64, 64, 0), (32, 0, 0), (16, 24, 0), (7, 64, 0),
                               (1, 24, 0), (0, 24, n)]:
        monkeypatch.setenv("D3FGPU_WALK_TILE", str(tile))
        monkeypatch.setenv("D3FGPU_WALK_PACKED_MAX", str(packed))
        monkeypatch.setenv("D3FGPU_WALK_SOLO", str(solo))
        out.append(((tile, packed, solo), walk_dv(pos, dims, s)))
    for key, dv in out[1:]:
        assert np.array_equal(dv, out[0][1]), f"tile/packed/solo {key} differs from {out[0][0]}"
    # the wave-per-point walk (the default at this size) and the group walks (2, 4, 8 points
    # per wave) sum the same terms in other orders
    for name in ("D3FGPU_WALK_WIDE", "D3FGPU_WALK_TILE", "D3FGPU_WALK_PACKED_MAX", "D3FGPU_WALK_SOLO",
                 "D3FGPU_WALK_GROUP"):
        monkeypatch.delenv(name, raising=False)
    others = [("wave per point", walk_dv(pos, dims, s))]
    for g in (2, 4, 8):
        monkeypatch.setenv("D3FGPU_WALK_GROUP", str(g))
        others.append((f"{g} points per wave", walk_dv(pos, dims, s)))
    monkeypatch.delenv("D3FGPU_WALK_GROUP")
    for label, dv in others:
        assert rel_l2(dv, out[0][1]) < 1e-6, label
    ref, _ = oracle_manybody(pos, dims, s)
    keep = ~members
    for label, dv in [("any tiling", out[0][1])] + others:
        err = rel_l2(dv[keep], ref[keep])
        report(f"manyBody tree {dims}D n={n} clump{' + 40 coincident' if coincident else ''}, {label}", err)
        assert err < 1e-4
        assert close_count(dv[keep], ref[keep], rtol=1e-3) <= tree_outliers(n)


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("groups", ["pair", "forty", "one percent"])
def test_manybody_tree_coincident_groups(engine_factory, dims, groups):
    """Coincident points are one d3 leaf (a chain) in the cell where they part from the
    other points, their strength the chain's undamped sum (src/manyBody.js:61-68): the
    tree gives such a group that depth and value, so every other node's force matches the
    FP64 oracle in 1D (x1.41 per level), 2D and 3D (x0.71).  The group's own members
    push each other along the hash jiggle, which d3 draws from its random source instead:
    they are left out of the comparison and checked to separate."""
    n = 6000
    pos = make_nodes_fast(n, dims, 20 * np.cbrt(n), seed=31 + dims)
    members = np.zeros(n, bool)
    rng = np.random.default_rng(dims)
    if groups == "pair":
        pos[17] = pos[4000]
        members[[17, 4000]] = True
    elif groups == "forty":
        pos[1000:1040] = pos[2500]
        members[1000:1040] = members[2500] = True
    else:
        src = rng.choice(n, n // 100, replace=False)
        dst = rng.choice(np.setdiff1d(np.arange(n), src), n // 100, replace=False)
        pos[dst] = pos[src]
        members[src] = members[dst] = True
    s = -30.0 * (1 + 0.5 * np.cos(np.arange(n)))
    e = engine_factory()
    _, dv = gpu_dv(e, pos, dims, [d3fgpu.many_body(s)])
    assert e.stats().walk_errors == 0
    ref, _ = oracle_manybody(pos, dims, s)
    keep = ~members
    err = rel_l2(dv[keep], ref[keep])
    report(f"manyBody tree {dims}D n={n} coincident {groups}", err, members=int(members.sum()))
    assert err < 1e-4
    assert close_count(dv[keep], ref[keep], rtol=1e-3) <= tree_outliers(n)
    assert np.all(np.isfinite(dv)) and np.all(np.linalg.norm(dv[members], axis=1) > 0)


@pytest.mark.parametrize("theta", [0.9, 0.0])
def test_gpu_jiggle_known_answer(theta, engine_factory):
    """Two coincident points: each pushes the other with strength * alpha / distanceMin along
    the hash jiggle displacement of (lo, hi) = (0, 1) at step 0 under the engine's seed
    (wgsl.js:28-52, restated in the oracle as or_gpu_jiggle), opposite on the two nodes."""
    for seed in (0, 12345, 0xdeadbeef):
        for dims in (1, 2, 3):
            e = d3fgpu.Engine(0)
            try:
                e.set_seed(seed)
                pos = np.zeros((2, 3))
                pos[:, :dims] = [3.25, -7.5, 11.0][:dims]
                s = np.array([-30.0, -12.0])
                _, dv = gpu_dv(e, pos, dims, [d3fgpu.many_body(s, theta=theta)], alpha=0.5)
            finally:
                e.close()
            d = np.array([O.lib().or_gpu_jiggle(0, 1, a, 0, seed) for a in range(3)], np.float64)
            d[dims:] = 0
            u = d / np.linalg.norm(d)
            # |d| < distanceMin: l' = sqrt(1 * l) = |d|, so each push is strength * alpha * d / |d|
            np.testing.assert_allclose(dv[0], u * s[1] * 0.5, rtol=2e-6, atol=1e-12)
            np.testing.assert_allclose(dv[1], -u * s[0] * 0.5, rtol=2e-6, atol=1e-12)


def walk_dv(pos, dims, s):
    """One many-body force on a fresh engine (walk knobs from the environment), with no
    walk errors."""
    e = d3fgpu.Engine(0)
    try:
        _, dv = gpu_dv(e, pos, dims, [d3fgpu.many_body(s)])
        assert e.stats().walk_errors == 0
    finally:
        e.close()
    return dv


@pytest.mark.parametrize("dims,n", [(3, 12288), (3, 12289), (2, 24576), (2, 24577), (1, 20000), (1, 32768),
                                    (1, 32769)])
def test_manybody_lane_group_walk_switch_points(monkeypatch, dims, n):
    """With the wave-per-point walk off, either side of the sizes up to which a tree is
    walked with a group of 2^D lanes per point (12,288 points in 3D, 24,576 in 2D, 32,768 in
    1D, with 1D's largest workgroups in LDS at 20,000): that walk equals the tiled
    exact-mode walk bit for bit, has no walk errors, and matches the FP64 oracle."""
    pos = make_nodes_fast(n, dims, 20 * n ** (1 / dims), seed=n)
    s = np.full(n, -30.0) * (1 + 0.25 * np.sin(np.arange(n)))
    out = []
    monkeypatch.setenv("D3FGPU_WALK_WIDE", "0")
    monkeypatch.setenv("D3FGPU_WALK_GROUP", "1")
    for solo in (None, "0"):
        if solo is None:
            monkeypatch.delenv("D3FGPU_WALK_SOLO", raising=False)
        else:
            monkeypatch.setenv("D3FGPU_WALK_SOLO", solo)
        out.append(walk_dv(pos, dims, s))
    assert np.array_equal(out[0], out[1])
    ref, _ = oracle_manybody(pos, dims, s)
    err = rel_l2(out[0], ref)
    report(f"manyBody tree {dims}D n={n} (lane-group walk switch)", err)
    assert err < 1e-4
    assert close_count(out[0], ref, rtol=1e-3) <= tree_outliers(n)


@pytest.mark.parametrize("dims,n", [(3, 16384), (3, 16385), (2, 16385), (1, 16384), (3, 262144), (3, 262145),
                                    (2, 262145)])
def test_manybody_walk_switch_points(monkeypatch, dims, n):
    """Either side of the sizes where the walk changes: one point per wave up to 16,384
    points, two per wave up to 262,144, 64-point tiles beyond.  The default walk has no
    walk errors, is within rounding of the tiled walk on the same tree where it is another
    walk (the same terms, summed in another order) and that walk itself beyond 262,144,
    and matches the FP64 oracle."""
    pos = make_nodes_fast(n, dims, 20 * n ** (1 / dims), seed=n + dims)
    s = np.full(n, -30.0) * (1 + 0.25 * np.cos(np.arange(n)))
    for name in ("D3FGPU_WALK_WIDE", "D3FGPU_WALK_GROUP"):
        monkeypatch.delenv(name, raising=False)
    dflt = walk_dv(pos, dims, s)
    monkeypatch.setenv("D3FGPU_WALK_WIDE", "0")
    monkeypatch.setenv("D3FGPU_WALK_GROUP", "1")
    tiled = walk_dv(pos, dims, s)
    if n <= 262144:
        assert 0 < rel_l2(dflt, tiled) < 1e-6
    else:
        assert np.array_equal(dflt, tiled)
    ref, _ = oracle_manybody(pos, dims, s)
    err = rel_l2(dflt, ref)
    report(f"manyBody tree {dims}D n={n} (walk switch)", err)
    assert err < 1e-4
    assert close_count(dflt, ref, rtol=1e-3) <= tree_outliers(n)


@pytest.mark.parametrize("dims", [2, 3])
def test_manybody_pool_of_single_point_openings(monkeypatch, dims):
    """The tiled walk's cheaper mode (past 262,144 points) sends openings needed by few of a
    tile's points to a pool of single-point openings.  No pool, the default, every opening
    below the whole tile pooled, and a pool of 8 entries (always full: its entries go to the
    stack as one-point openings) sum the same terms in other orders: within rounding of each
    other, no walk errors, and within the FP64 oracle's bar."""
    n = 300000
    pos = make_nodes_fast(n, dims, 20 * n ** (1 / dims), seed=7 + dims)
    s = np.full(n, -30.0) * (1 + 0.25 * np.sin(np.arange(n)))
    settings = [{"D3FGPU_WALK_POOL_K": "0"}, {}, {"D3FGPU_WALK_POOL_K": "63"},
                {"D3FGPU_WALK_POOL_K": "63", "D3FGPU_WALK_POOL_CAP": "8"}]
    out = []
    for env in settings:
        for name in ("D3FGPU_WALK_POOL_K", "D3FGPU_WALK_POOL_CAP"):
            monkeypatch.delenv(name, raising=False)
        for name, value in env.items():
            monkeypatch.setenv(name, value)
        out.append(walk_dv(pos, dims, s))
    for dv in out[1:]:
        assert rel_l2(dv, out[0]) < 1e-6
    ref, _ = oracle_manybody(pos, dims, s)
    for env, dv in zip(settings, out):
        err = rel_l2(dv, ref)
        report(f"manyBody tree {dims}D n={n} pool {env or 'default'}", err)
        assert err < 1e-4
        assert close_count(dv, ref, rtol=1e-3) <= tree_outliers(n)


def test_manybody_tree_above_exact_mode(engine_factory):
    """Past 262,144 points the walk takes 64-point tiles and the cheaper accumulation."""
    n = 300000
    pos = make_nodes_fast(n, 3, 20 * np.cbrt(n))
    s = np.full(n, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, 3, [d3fgpu.many_body(s)])
    ref, _ = oracle_manybody(pos, 3, s)
    err = rel_l2(dv, ref)
    report(f"manyBody tree n={n}", err)
    assert close_count(dv, ref, rtol=1e-3, atol=1e-3) <= tree_outliers(n)
    assert err < 1e-4


@pytest.mark.parametrize("dims,outlier", [(2, 120), (3, 120), (2, 2000), (3, 2000)])
def test_manybody_tree_clumps(engine_factory, dims, outlier):
    pos = make_nodes_fast(9000, dims, 2)
    pos[0, 0] = outlier
    s = np.full(9000, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, dims, [d3fgpu.many_body(s)])
    ref, _ = oracle_manybody(pos, dims, s)
    report(f"manyBody tree clump {dims}D outlier {outlier}", rel_l2(dv, ref))
    assert close_count(dv, ref, rtol=1e-3) <= tree_outliers(9000)


def test_manybody_tree_stats_match_oracle(engine_factory):
    n = 20000
    pos = make_nodes_fast(n, 3, 20 * np.cbrt(n))
    s = np.full(n, -30.0)
    e = engine_factory()
    gpu_dv(e, pos, 3, [d3fgpu.many_body(s)])
    st = e.stats()
    _, ost = oracle_manybody(pos, 3, s)
    print(f"[stats] interactions gpu={st.interactions} oracle={ost[0]} opened(wave)={st.opened}", flush=True)
    assert st.walk_errors == 0
    assert abs(int(st.interactions) - int(ost[0])) <= 0.01 * ost[0]


# ---- many-body, exact ---------------------------------------------------------------

@pytest.mark.parametrize("dims", [1, 2, 3])
def test_manybody_exact(engine_factory, dims):
    pos = make_nodes(300, dims)
    s = np.full(300, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, dims, [d3fgpu.many_body(s, theta=0)])
    ref = O.manybody_exact(pos, np.zeros_like(pos), s, 1.0, ndim=dims)
    report(f"manyBody exact {dims}D", rel_l2(dv, ref))
    assert close_count(dv, ref) == 0
    assert rel_l2(dv, ref) < 1e-5


def test_manybody_exact_strength_dmin_dmax(engine_factory):
    pos = make_nodes(200, 3)
    s = np.array([15.0 if i % 3 == 0 else -20.0 - i % 7 for i in range(200)])
    _, dv = gpu_dv(engine_factory(), pos, 3, [d3fgpu.many_body(s, theta=0, distance_min=5, distance_max=150)])
    ref = O.manybody_exact(pos, np.zeros_like(pos), s, 1.0, dmin2=25, dmax2=150 ** 2)
    assert close_count(dv, ref, rtol=1e-3) == 0


@pytest.mark.parametrize("n", [1, 2, 3, 7, 8, 9, 127, 128, 129, 255, 257])
def test_manybody_exact_small(engine_factory, n):
    pos = make_nodes(n, 3)
    s = np.full(n, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, 3, [d3fgpu.many_body(s, theta=0)])
    ref = O.manybody_exact(pos, np.zeros_like(pos), s, 1.0)
    assert close_count(dv, ref, rtol=1e-3, atol=1e-3) == 0


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_manybody_exact_many_blocks(engine_factory, dims):
    """Several blocks of the all-pairs kernel (bodies outside the source tile's block take
    the fast path), with distanceMin/distanceMax in force."""
    n = 5000
    pos = make_nodes_fast(n, dims)
    s = np.where(np.arange(n) % 5 == 0, 12.0, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, dims, [d3fgpu.many_body(s, theta=0, distance_min=2, distance_max=120)],
                   alpha=0.7)
    ref = O.manybody_exact(pos, np.zeros_like(pos), s, 0.7, dmin2=4, dmax2=120 ** 2, ndim=dims)
    report(f"manyBody exact {n} nodes {dims}D", rel_l2(dv, ref))
    assert rel_l2(dv, ref) < 1e-5
    # distanceMax is a step: a pair within float rounding of it may fall on either side in
    # f32 and in f64; nodes with such a pair are left out of the per-node bar
    edge = np.zeros(n, bool)
    for lo in range(0, n, 500):
        l = ((pos[lo:lo + 500, None, :] - pos[None, :, :]) ** 2).sum(-1)
        edge[lo:lo + 500] = (np.abs(l - 120.0 ** 2) < 2e-6 * 120.0 ** 2).any(1)
    assert edge.sum() < n // 50
    assert close_count(dv[~edge], ref[~edge], rtol=1e-4, atol=1e-4) == 0


def test_manybody_exact_coincident_across_blocks(engine_factory):
    """Coincident pairs in different blocks and in one block: each pair adds a push of
    |strength| * alpha / distanceMin along the hash jiggle, equal and opposite on its two
    nodes (the FP64 oracle leaves coincident pairs out, so the push is the residual)."""
    n, alpha = 3000, 0.5
    pos = make_nodes_fast(n, 3)
    pairs = [(5, 1705), (100, 2900), (600, 601), (1201, 2410)]
    for i, j in pairs:
        pos[j] = pos[i]
    s = np.full(n, -30.0)
    _, dv = gpu_dv(engine_factory(), pos, 3, [d3fgpu.many_body(s, theta=0)], alpha=alpha)
    ref = O.manybody_exact(pos, np.zeros_like(pos), s, alpha)
    r = dv - ref
    hit = np.zeros(n, bool)
    for i, j in pairs:
        hit[i] = hit[j] = True
        np.testing.assert_allclose(np.linalg.norm(r[i]), 30.0 * alpha, rtol=1e-3)
        np.testing.assert_allclose(r[i], -r[j], rtol=1e-3, atol=1e-3)
    assert close_count(dv[~hit], ref[~hit]) == 0


# ---- link ---------------------------------------------------------------------------

def oracle_link(pos, vel, dims, src, dst, strengths, distances, bias, iterations, alpha):
    x, y, z = (O.dv(pos[:, a]) for a in range(3))
    vx, vy, vz = (O.dv(vel[:, a]) for a in range(3))
    u = lambda a: O.P(np.ascontiguousarray(a, np.uint32), O.C.c_uint32)
    f32 = lambda a: O.dv(np.asarray(a, np.float32))
    O.lib().or_link_jacobi(len(pos), dims, O.P(x), O.P(y), O.P(z), O.P(vx), O.P(vy), O.P(vz), len(src), u(src), u(dst),
                           O.P(f32(strengths)), O.P(f32(distances)), O.P(f32(bias)), iterations, alpha)
    return np.stack([vx, vy, vz], 1)


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_link_graph(engine_factory, dims):
    n = 3000
    pos = make_nodes_fast(n, dims, 400)
    vel = make_nodes_fast(n, dims, 2, seed=9)
    src, dst = O.make_graph(n, 42)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    bias, strengths, distances = O.link_defaults(n, src, dst)
    f = d3fgpu.link(src, dst, strengths, distances, bias, iterations=2)
    _, dv = gpu_dv(engine_factory(), pos, dims, [f], alpha=0.7, vel=vel)
    ref = oracle_link(pos, vel, dims, src, dst, strengths, distances, bias, 2, 0.7) - vel
    err = rel_l2(dv, ref)
    report(f"link {dims}D", err)
    assert err < 1e-4
    assert close_count(dv, ref, rtol=1e-3) == 0


def test_link_defaults_built_on_the_device(engine_factory):
    """Links given without strengths, distances or bias take d3's defaults (src/link.js:
    56-97) -- counted on the device from the CSR -- including repeated links; the CSR keeps
    each node's links in link order."""
    n, dims = 2500, 3
    pos = make_nodes_fast(n, dims, 400, seed=3)
    vel = make_nodes_fast(n, dims, 2, seed=4)
    src, dst = O.make_graph(n, 7)
    keep = src != dst
    src, dst = np.concatenate([src[keep], src[keep][:300]]), np.concatenate([dst[keep], dst[keep][:300]])
    bias, strengths, distances = O.link_defaults(n, src, dst)
    links = np.stack([src, dst], 1).astype(np.uint32)
    f = d3fgpu.Force(d3fgpu.LINK, links=links, iterations=1)
    _, dv = gpu_dv(engine_factory(), pos, dims, [f], alpha=0.6, vel=vel)
    ref = oracle_link(pos, vel, dims, src, dst, strengths, np.full(len(src), 30.0), bias, 1, 0.6) - vel
    err = rel_l2(dv, ref)
    report("link defaults on the device", err)
    assert err < 1e-4
    assert close_count(dv, ref, rtol=1e-3) == 0


def test_link_rejects_bad_indices(engine_factory):
    e = engine_factory()
    e.set_nodes(make_nodes_fast(10, 2, 10), ndim=2)
    with pytest.raises(d3fgpu.FgError):
        e.set_forces([d3fgpu.Force(d3fgpu.LINK, links=np.array([[0, 1], [2, 10]], np.uint32))])


# ---- collide ------------------------------------------------------------------------

def oracle_collide(pos, vel, dims, radii, strength, iterations):
    x, y, z = (O.dv(pos[:, a]) for a in range(3))
    vx, vy, vz = (O.dv(vel[:, a]) for a in range(3))
    r = O.dv(np.asarray(radii, np.float32))
    O.lib().or_collide_jacobi(len(pos), dims, O.P(x), O.P(y), O.P(z), O.P(vx), O.P(vy), O.P(vz), O.P(r),
                              strength, iterations)
    return np.stack([vx, vy, vz], 1)


@pytest.mark.parametrize("n,dims", [(2000, 2), (2000, 3), (10000, 1), (10000, 2), (10000, 3)])
def test_collide(engine_factory, n, dims):
    pos = make_nodes_fast(n, dims, 12 * n ** (1 / dims))
    radii = np.array([6 + i % 3 for i in range(n)], float)
    f = d3fgpu.collide(radii, strength=0.9, iterations=1)
    _, dv = gpu_dv(engine_factory(), pos, dims, [f])
    ref = oracle_collide(pos, np.zeros_like(pos), dims, radii, 0.9, 1)
    err = rel_l2(dv, ref)
    report(f"collide n={n} {dims}D", err, touched=int(np.sum(np.any(ref != 0, 1))))
    assert err < 1e-4
    assert close_count(dv, ref, rtol=1e-3) == 0


def test_collide_grid_two_iterations_lognormal(engine_factory):
    n, dims = 20000, 3
    rng = np.random.default_rng(5)
    pos = make_nodes_fast(n, dims, 10 * np.cbrt(n))
    radii = np.exp(np.log(2) + 0.4 * rng.standard_normal(n))
    f = d3fgpu.collide(radii, strength=1.0, iterations=2)
    _, dv = gpu_dv(engine_factory(), pos, dims, [f])
    ref = oracle_collide(pos, np.zeros_like(pos), dims, radii, 1.0, 2)
    err = rel_l2(dv, ref)
    report("collide grid lognormal 2 iterations", err)
    assert err < 1e-4


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_collide_grid_later_iterations_reuse_the_grid(engine_factory, dims):
    """Four iterations where the points move a little: the later ones query the first
    one's grid (points filed where they were then, pairs taken at their positions now,
    cells within the largest move visited too).  Against four collide forces of one
    iteration each (every one builds its own grid; the same Jacobi steps), and in 2D/3D
    against the oracle's four iterations.  (On a line points pass each other: after the
    first iteration nodes 9885 and 17247 stand 0.001 apart at x = -48050 in FP64, the same
    f32 value there (its spacing is 0.004), so the f32 path takes them as coincident (the
    hash jiggle) where the FP64 oracle pushes them apart along their 0.001: rel-L2 1.7e-2
    from the second iteration on, with one grid per iteration as well; the oracle is left
    out in 1D.)"""
    n = 20000
    rng = np.random.default_rng(11)
    pos = make_nodes_fast(n, dims, 10 * n ** (1 / dims), seed=3)
    radii = np.exp(np.log(2) + 0.4 * rng.standard_normal(n))
    _, dv = gpu_dv(engine_factory(), pos, dims, [d3fgpu.collide(radii, strength=0.7, iterations=4)])
    _, dv1 = gpu_dv(engine_factory(), pos, dims, [d3fgpu.collide(radii, strength=0.7, iterations=1) for _ in range(4)])
    err1 = rel_l2(dv, dv1)
    report(f"collide grid 4 iterations {dims}D against 4 grids", err1)
    assert err1 < 1e-5
    if dims > 1:
        ref = oracle_collide(pos, np.zeros_like(pos), dims, radii, 0.7, 4)
        err = rel_l2(dv, ref)
        report(f"collide grid 4 iterations {dims}D", err, touched=int(np.sum(np.any(ref != 0, 1))))
        assert err < 1e-4
        assert close_count(dv, ref, rtol=1e-3) == 0


def test_collide_grid_rebuilds_after_large_moves(engine_factory):
    """Points crowded deep inside each other's radii (about 50 overlaps each) move several
    units -- more than a quarter of a 6-unit cell -- in one iteration, so the next ones build
    grids of their own.  Against the oracle's three iterations."""
    n, dims = 12000, 3
    pos = make_nodes_fast(n, dims, 60.0, seed=17)
    radii = np.full(n, 3.0)
    f = d3fgpu.collide(radii, strength=1.0, iterations=3)
    _, dv = gpu_dv(engine_factory(), pos, dims, [f])
    ref = oracle_collide(pos, np.zeros_like(pos), dims, radii, 1.0, 3)
    moved = float(np.max(np.linalg.norm(ref, axis=1)))
    err = rel_l2(dv, ref)
    report("collide grid rebuilt after large moves", err, largest_move=round(moved, 2))
    assert moved > 1.5  # the first iteration's moves exceed a quarter cell
    assert err < 1e-4


@pytest.mark.parametrize("dims", [2, 3])
def test_collide_grid_heavy_tail(engine_factory, dims):
    """1% of the nodes ten times larger than the rest: the two-class grid, where a large node
    searches several rings of the fine cells."""
    n = 12000
    rng = np.random.default_rng(9)
    pos = make_nodes_fast(n, dims, 6 * n ** (1 / dims), seed=21)
    radii = 2.0 + rng.random(n)
    radii[rng.choice(n, n // 100, replace=False)] = 40.0
    f = d3fgpu.collide(radii, strength=0.7, iterations=1)
    _, dv = gpu_dv(engine_factory(), pos, dims, [f])
    ref = oracle_collide(pos, np.zeros_like(pos), dims, radii, 0.7, 1)
    err = rel_l2(dv, ref)
    report(f"collide grid heavy tail {dims}D", err, touched=int(np.sum(np.any(ref != 0, 1))))
    assert err < 1e-4
    assert close_count(dv, ref, rtol=1e-3) == 0


# ---- center / radial / position / integrate, several ticks --------------------------

def run_both(engine, pos, dims, forces_gpu, forces_oracle, ticks, fix=None, fixmask=None):
    engine.set_nodes(pos, fix=fix, fixmask=fixmask, ndim=dims)
    engine.set_forces(forces_gpu)
    alpha, decay_a = 1.0, 1 - 0.001 ** (1 / 300)
    alphas = []
    for _ in range(ticks):
        alpha += (0 - alpha) * decay_a
        alphas.append(alpha)
    engine.enqueue(alphas, 0.6)
    gp, gv = engine.readback()
    fixarr = None
    if fix is not None:
        fixarr = np.where(np.stack([(fixmask >> a) & 1 for a in range(3)], 1) == 1, fix, np.nan)
    sim = O.Sim(pos, fix=fixarr, ndim=dims)
    keep = [f for f in forces_oracle]
    sim.set_forces([f[0] for f in keep], jacobi=True)
    sim.tick(ticks)
    op, ov = sim.state()
    return gp, gv, op, ov


def odesc(type, **kw):
    keep = []
    d = O.ForceDesc(type=type, iterations=kw.pop("iterations", 1))
    for k, v in kw.items():
        if isinstance(v, np.ndarray):
            a = O.dv(np.asarray(v, np.float32))
            keep.append(a)
            setattr(d, k, O.P(a))
        else:
            setattr(d, k, v)
    return d, keep


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_center(engine_factory, dims):
    pos = make_nodes(1000, dims, 100)
    pos[:, 0] += 40
    if dims > 1:
        pos[:, 1] -= 25
    if dims > 2:
        pos[:, 2] += 10
    pos = pos.astype(np.float32).astype(np.float64)
    c = (3.0, 4.0 if dims > 1 else 0.0, 5.0 if dims > 2 else 0.0)
    gp, gv, op, ov = run_both(engine_factory(), pos, dims, [d3fgpu.center(*c, strength=0.7)],
                              [odesc(1, x=c[0], y=c[1], z=c[2], strength=0.7)], 2)
    assert close_count(gp, op, rtol=1e-3) == 0


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_radial(engine_factory, dims):
    n = 200
    pos = make_nodes(n, dims)
    r = np.array([30.0 + i % 5 * 10 for i in range(n)])
    w = np.full(n, 0.3)
    c = (5.0, -5.0 if dims > 1 else 0.0, 2.0 if dims > 2 else 0.0)
    gp, gv, op, ov = run_both(engine_factory(), pos, dims, [d3fgpu.radial(r, w, *c)],
                              [odesc(6, radii=r, strengths=w, x=c[0], y=c[1], z=c[2])], 3)
    assert close_count(np.hstack([gp, gv]), np.hstack([op, ov]), rtol=1e-3) == 0


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_position(engine_factory, dims):
    n = 200
    pos = make_nodes(n, dims)
    gf = [d3fgpu.position(0, np.array([i % 10 for i in range(n)], float),
                          np.array([0.2 if i % 4 else 0.05 for i in range(n)]))]
    of = [odesc(5, axis=0, targets=np.array([i % 10 for i in range(n)], float),
                strengths=np.array([0.2 if i % 4 else 0.05 for i in range(n)]))]
    if dims > 1:
        gf.append(d3fgpu.position(1, np.full(n, -20.0), np.full(n, 0.3)))
        of.append(odesc(5, axis=1, targets=np.full(n, -20.0), strengths=np.full(n, 0.3)))
    if dims > 2:
        gf.append(d3fgpu.position(2, np.full(n, 15.0), np.full(n, 0.1)))
        of.append(odesc(5, axis=2, targets=np.full(n, 15.0), strengths=np.full(n, 0.1)))
    gp, gv, op, ov = run_both(engine_factory(), pos, dims, gf, of, 3)
    assert close_count(np.hstack([gp, gv]), np.hstack([op, ov]), rtol=1e-3) == 0


def test_pins_and_force_order(engine_factory):
    n = 250
    pos = make_nodes(n, 3, 120)
    pos[:, 0] = (pos[:, 0] + 30).astype(np.float32)
    fix = np.zeros((n, 3))
    mask = np.zeros(n, np.uint8)
    fix[0] = [5, -6, 7]
    mask[0] = 7
    fix[1, 0] = 100
    mask[1] = 1
    s = np.full(n, -15.0)
    gf = [d3fgpu.many_body(s, theta=0), d3fgpu.position(0, np.full(n, 10.0), np.full(n, 0.05)),
          d3fgpu.center(), d3fgpu.radial(np.full(n, 60.0), np.full(n, 0.02))]
    of = [odesc(4, strengths=s, theta2=0.0, distance_min2=1.0, distance_max2=np.inf),
          odesc(5, axis=0, targets=np.full(n, 10.0), strengths=np.full(n, 0.05)),
          odesc(1, strength=1.0), odesc(6, radii=np.full(n, 60.0), strengths=np.full(n, 0.02))]
    gp, gv, op, ov = run_both(engine_factory(), pos, 3, gf, of, 5, fix=fix, fixmask=mask)
    assert gp[0].tolist() == [5.0, -6.0, 7.0]
    assert gv[0].tolist() == [0.0, 0.0, 0.0]
    assert gp[1, 0] == 100.0
    assert close_count(np.hstack([gp, gv]), np.hstack([op, ov]), rtol=2e-3, atol=1e-3) == 0


def test_coincident_nodes_separate(engine_factory):
    e = engine_factory()
    pos = np.tile([1.0, 2.0, 3.0], (10, 1))
    e.set_nodes(pos, ndim=3)
    e.set_forces([d3fgpu.many_body(np.full(10, -30.0)), d3fgpu.collide(np.full(10, 2.0))])
    alpha, a = 1.0, []
    for _ in range(30):
        alpha += -alpha * (1 - 0.001 ** (1 / 300))
        a.append(alpha)
    e.enqueue(a, 0.6)
    p, v = e.readback()
    assert np.all(np.isfinite(p))
    d = np.linalg.norm(p[:, None] - p[None], axis=2) + np.eye(10) * 1e9
    assert d.min() > 0.5


@pytest.mark.parametrize("dims,n", [(1, 700), (2, 1024), (3, 1), (3, 513), (3, 1025)])
def test_center_last_in_one_launch(dims, n):
    """A tick whose last force is center over at most 1,024 nodes computes the shift and
    integrates in one launch; it must leave the state center followed by a force that adds
    nothing leaves (center then runs on its own kernels, its shift applied before the
    integration), bit for bit, with velocities and pinned axes.  1,025 nodes: the usual path
    both ways."""
    rng = np.random.default_rng(dims * 10007 + n)
    pos = np.zeros((n, 3))
    pos[:, :dims] = rng.normal(0.0, 50.0, (n, dims)).astype(np.float32)
    vel = np.zeros((n, 3))
    vel[:, :dims] = rng.normal(0.0, 2.0, (n, dims)).astype(np.float32)
    fix = np.full((n, 3), np.nan)
    fix[::37, :dims] = 7.0
    out = []
    for forces in ([d3fgpu.center(1.0, 2.0, 3.0)],
                   [d3fgpu.center(1.0, 2.0, 3.0), d3fgpu.radial(np.full(n, 10.0), np.zeros(n))]):
        e = d3fgpu.Engine(0)
        try:
            e.set_nodes(pos, vel=vel, fix=fix, ndim=dims)
            e.set_forces(forces)
            e.enqueue([0.5, 0.49, 0.48], 0.6)
            out.append(e.readback())
        finally:
            e.close()
    assert np.all(np.isfinite(out[0][0]))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_center_in_fresh_engines_over_reused_memory():
    """Engines made and freed in turn get memory back that held other data: a center pass's
    counters must be cleared before its first launch reads them (set-up clears are waited
    for), so every fresh engine moves every node by the FP64 shift."""
    rng = np.random.default_rng(8)
    for k in range(12):
        n = int(rng.integers(300, 5000))
        pos = make_nodes_fast(n, 3, 300, seed=k)
        e = d3fgpu.Engine(0)
        try:  # dirty the allocator with a tree, links and velocities
            src = rng.integers(0, n, n).astype(np.uint32)
            dst = ((src + 1) % n).astype(np.uint32)
            b, st, di = O.link_defaults(n, src, dst)
            gpu_dv(e, pos, 3, [d3fgpu.link(src, dst, st, di, b), d3fgpu.many_body(np.full(n, -30.0))],
                   vel=rng.standard_normal((n, 3)))
        finally:
            e.close()
        shifted = pos + np.array([1e-3, -2.5e-4, 4e-5]) * (k + 1)
        e = d3fgpu.Engine(0)
        try:
            p, _ = gpu_dv(e, shifted, 3, [d3fgpu.center()])
        finally:
            e.close()
        p32 = shifted.astype(np.float32)
        mean = p32.astype(np.float64).mean(0)
        want = (p32 - mean.astype(np.float32)).astype(np.float64)
        assert np.array_equal(p, want), f"engine {k}: center did not move the nodes by the mean"


def test_gpu_buffers_hold_the_state(engine_factory):
    """fg_buffers: a float4 per node of the device state, kept current after every step
    (engine.js:418-420, test/gpu-simulation-test.js:421-443), read through the pointers
    bit-equal to the readback."""
    n = 777
    e = engine_factory()
    pos = make_nodes_fast(n, 3, 300)
    e.set_nodes(pos, ndim=3)
    e.set_forces([d3fgpu.many_body(np.full(n, -30.0)), d3fgpu.center()])
    b = e.buffers()
    assert b.count == n and b.stride == 16
    for ticks in (1, 4):
        e.enqueue(np.full(ticks, 0.5), 0.6)
        p, v = e.readback()
        dp = e.copy_from_device(b.positions, 4 * n).reshape(n, 4)
        dv = e.copy_from_device(b.velocities, 4 * n).reshape(n, 4)
        np.testing.assert_array_equal(dp[:, :3].astype(np.float64), p)
        np.testing.assert_array_equal(dv[:, :3].astype(np.float64), v)

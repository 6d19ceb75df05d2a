"""The accuracy protocol at the benchmark sizes: from identical f32 states (d3's
phyllotaxis start, and the state after GPU ticks), one application of each force on the GPU
against the FP64 oracle.

  config A: 1,000 nodes, 3D, link + manyBody(theta 0.9) + center, at ticks 0, 1, 150 and
            299 of the GPU's own run; and the layout's size after 300 ticks
  config B: 200,000 nodes, 3D, link + manyBody (theta 0.9, and theta 0 on a sample of
            2,000 nodes against the exact all-pairs sum) + center
  config C: 1,000,000 nodes, 3D, link + manyBody(theta 0.9) + center + collide (2 iterations,
            lognormal radii) + radial, 1% pinned

Bars: manyBody rel-L2 <= 1e-4 with at most one node per 1000 outside the per-node rule
(gpu-helpers.js:63-84); collide (two Jacobi iterations, the second fed by the f32 result of
the first) rel-L2 <= 1e-4 against the FP64 Jacobi restatement; link and radial rel-L2 <=
1e-5; center: positions after the move within two ulps.  Every figure goes to the parity
report.
"""
import numpy as np
import pytest

import d3fgpu
import oracle as O
from bench import alphas, link_defaults, make_graph, phyllotaxis3
from helpers import PARITY, close_count, rel_l2, tree_outliers

pytestmark = pytest.mark.gpu

ALPHA = 0.5


def record(config, force, tick, err, **kw):
    PARITY.append({"case": f"{config} {force} tick {tick}", "rel_l2": err, **kw})
    print(f"[parity] {config} {force} tick {tick}: rel-L2 {err:.3e} {kw}", flush=True)


def gpu_increment(pos, vel, forces):
    """One application of `forces` on the GPU: (dpos, dvel)."""
    e = d3fgpu.Engine(0)
    e.set_nodes(pos, vel=vel, ndim=3)
    e.set_forces(forces)
    e.apply_forces(ALPHA)
    p, v = e.readback()
    e.close()
    return p - pos, v - vel


def oracle_increment(pos, vel, descs):
    sim = O.Sim(pos, vel=vel, ndim=3)
    sim.set_forces(descs, jacobi=True)
    sim.apply_forces(ALPHA)
    p, v = sim.state()
    return p - pos, v - vel


def keep(a, store):
    a = np.ascontiguousarray(a, np.float64)
    store.append(a)
    return O.P(a)


def gpu_state_after(pos, forces, ticks, fix=None, fixmask=None):
    e = d3fgpu.Engine(0)
    e.set_nodes(pos, fix=fix, fixmask=fixmask, ndim=3)
    e.set_forces(forces)
    a, _ = alphas(1.0, ticks)
    e.enqueue(a, 0.6)
    p, v = e.readback()
    e.close()
    return p, v


def check_center(config, tick, pos, vel):
    """center moves every position by one shift (positions, not velocities): the positions
    after the move against the FP64 ones, each within two ulps of the larger operand of its
    f32 subtraction (position or shift)."""
    dp, _ = gpu_increment(pos, vel, [d3fgpu.center()])
    rp, _ = oracle_increment(pos, vel, [O.ForceDesc(type=1, strength=1.0)])
    err = rel_l2(pos + dp, pos + rp)
    scale = np.maximum(np.abs(pos), np.abs(rp)) * 2.0 ** -23
    ulps = float(np.max(np.abs(dp - rp) / np.maximum(scale, 1e-30)))
    record(config, "center (positions after the move)", tick, err, max_ulps=round(ulps, 2))
    assert err < 1e-6
    assert ulps <= 2.0


def check_manybody(config, tick, pos, vel, s, theta=0.9):
    _, dv = gpu_increment(pos, vel, [d3fgpu.many_body(s, theta=theta)])
    after, _ = O.manybody(pos, vel, s, ALPHA, theta2=theta * theta, dmin2=1.0, dmax2=np.inf, ndim=3)
    ref = after - vel  # the oracle returns the velocities with the force added
    err = rel_l2(dv, ref)
    out = close_count(dv, ref, rtol=1e-3)
    record(config, f"manyBody theta {theta}", tick, err, outliers=out)
    assert err < 1e-4
    assert out <= tree_outliers(len(pos))


@pytest.fixture(scope="module")
def config_b():
    n = 200_000
    src, dst = make_graph(n, 42)
    bias, strengths, distances = link_defaults(n, src, dst)
    return n, src, dst, bias, strengths, distances


@pytest.mark.parametrize("tick", [0, 150])
def test_config_b_forces(config_b, tick):
    n, src, dst, bias, strengths, distances = config_b
    s = np.full(n, -30.0)
    link = d3fgpu.link(src, dst, strengths, distances, bias)
    pos = phyllotaxis3(n).astype(np.float32).astype(np.float64)
    vel = np.zeros_like(pos)
    if tick:
        pos, vel = gpu_state_after(pos, [link, d3fgpu.many_body(s), d3fgpu.center()], tick)
    check_manybody("B 200k", tick, pos, vel, s)
    # link, Jacobi
    store = []
    links = np.ascontiguousarray(np.stack([src, dst], 1).reshape(-1), np.uint32)
    store.append(links)
    ld = O.ForceDesc(type=3, iterations=1, count=len(src), links=O.P(links, O.C.c_uint32),
                     strengths=keep(strengths, store), distances=keep(distances, store), bias=keep(bias, store))
    _, dv = gpu_increment(pos, vel, [link])
    _, ref = oracle_increment(pos, vel, [ld])
    err = rel_l2(dv, ref)
    record("B 200k", "link", tick, err)
    assert err < 1e-5
    check_center("B 200k", tick, pos, vel)


def test_config_b_exact_sample(config_b):
    """theta 0 (the all-pairs path): 2,000 sampled nodes against the exact FP64 sum."""
    n = config_b[0]
    s = np.full(n, -30.0)
    pos = phyllotaxis3(n).astype(np.float32).astype(np.float64)
    _, dv = gpu_increment(pos, np.zeros_like(pos), [d3fgpu.many_body(s, theta=0.0)])
    idx = np.random.default_rng(1).choice(n, 2000, replace=False)
    ref = np.zeros((len(idx), 3))
    for k0 in range(0, len(idx), 250):
        rows = idx[k0:k0 + 250]
        d = pos[None, :, :] - pos[rows, None, :]
        l = (d * d).sum(-1)
        l[np.arange(len(rows)), rows] = np.inf
        l = np.where(l < 1.0, np.sqrt(l), l)  # distanceMin 1 (manyBody.js:95)
        ref[k0:k0 + 250] = (d * (s[None, :, None] * ALPHA / l[:, :, None])).sum(1)
    err = rel_l2(dv[idx], ref)
    record("B 200k", "manyBody theta 0 (2000-node sample)", 0, err)
    assert err < 1e-4


@pytest.mark.parametrize("tick", [0, 20])
def test_config_c_forces(tick):
    n = 1_000_000
    src, dst = make_graph(n, 42)
    bias, strengths, distances = link_defaults(n, src, dst)
    s = np.full(n, -30.0)
    rng = np.random.default_rng(7)
    radii = np.exp(np.log(2.0) + 0.4 * rng.standard_normal(n))
    rr, rw = np.full(n, 200.0), np.full(n, 0.05)
    link = d3fgpu.link(src, dst, strengths, distances, bias)
    coll = d3fgpu.collide(radii, strength=1.0, iterations=2)
    rad = d3fgpu.radial(rr, rw)
    pos = phyllotaxis3(n).astype(np.float32).astype(np.float64)
    vel = np.zeros_like(pos)
    if tick:
        mask = np.zeros(n, np.uint8)
        mask[::100] = 7
        fix = np.zeros((n, 3))
        fix[::100] = np.random.default_rng(11).uniform(-500, 500, (len(mask[::100]), 3))
        pos[::100] = fix[::100].astype(np.float32)
        pos, vel = gpu_state_after(pos, [link, d3fgpu.many_body(s), d3fgpu.center(), coll, rad], tick,
                                   fix=fix, fixmask=mask)
    check_manybody("C 1M", tick, pos, vel, s)
    store = []
    links = np.ascontiguousarray(np.stack([src, dst], 1).reshape(-1), np.uint32)
    store.append(links)
    cases = [
        ("link", [link], O.ForceDesc(type=3, iterations=1, count=len(src), links=O.P(links, O.C.c_uint32),
                                     strengths=keep(strengths, store), distances=keep(distances, store),
                                     bias=keep(bias, store))),
        ("collide 2 iterations", [coll], O.ForceDesc(type=2, iterations=2, strength=1.0, radii=keep(radii, store))),
        ("radial", [rad], O.ForceDesc(type=6, strengths=keep(rw, store), radii=keep(rr, store))),
    ]
    for name, gf, od in cases:
        _, dv = gpu_increment(pos, vel, gf)
        _, ref = oracle_increment(pos, vel, [od])
        err = rel_l2(dv, ref)
        record("C 1M", name, tick, err, touched=int(np.count_nonzero(np.any(ref != 0, 1))))
        assert err < (1e-4 if name.startswith("collide") else 1e-5)
    check_center("C 1M", tick, pos, vel)


# ---- config A: the reference's 1,000-node benchmark graph ------------------------------

def config_a():
    n = 1000
    src, dst = make_graph(n, 42)
    bias, strengths, distances = link_defaults(n, src, dst)
    s = np.full(n, -30.0)
    forces = [d3fgpu.link(src, dst, strengths, distances, bias), d3fgpu.many_body(s, theta=0.9), d3fgpu.center()]
    return n, src, dst, bias, strengths, distances, s, forces


@pytest.mark.parametrize("tick", [0, 1, 150, 299])
def test_config_a_forces(tick):
    """Config A (examples/benchmark.html:86-93: 1,000 nodes, 1,499 links, link + manyBody +
    center from d3's start): the state the GPU reaches after `tick` ticks, then one
    application of each force on the GPU and in the FP64 oracle."""
    n, src, dst, bias, strengths, distances, s, forces = config_a()
    pos = phyllotaxis3(n).astype(np.float32).astype(np.float64)
    vel = np.zeros_like(pos)
    if tick:
        pos, vel = gpu_state_after(pos, forces, tick)
    check_manybody("A 1k", tick, pos, vel, s)
    store = []
    links = np.ascontiguousarray(np.stack([src, dst], 1).reshape(-1), np.uint32)
    store.append(links)
    ld = O.ForceDesc(type=3, iterations=1, count=len(src), links=O.P(links, O.C.c_uint32),
                     strengths=keep(strengths, store), distances=keep(distances, store), bias=keep(bias, store))
    _, dv = gpu_increment(pos, vel, [forces[0]])
    _, ref = oracle_increment(pos, vel, [ld])
    err = rel_l2(dv, ref)
    record("A 1k", "link", tick, err)
    assert err < 1e-5
    check_center("A 1k", tick, pos, vel)


def test_config_a_layout_size():
    """300 ticks of config A on the GPU: the layout's RMS radius against the FP64 oracle's
    with the GPU's Jacobi link (374.19) and the reference's published GPU figure, 374
    (docs/differences.md:35-40; d3's sequential CPU layout is 372)."""
    n, src, dst, bias, strengths, distances, s, forces = config_a()
    pos = phyllotaxis3(n)
    p, _ = gpu_state_after(pos, forces, 300)
    rms = float(np.sqrt(np.mean(np.sum(p ** 2, 1))))
    store = []
    links = np.ascontiguousarray(np.stack([src, dst], 1).reshape(-1), np.uint32)
    store.append(links)
    descs = [O.ForceDesc(type=3, iterations=1, count=len(src), links=O.P(links, O.C.c_uint32),
                         strengths=keep(strengths, store), distances=keep(distances, store), bias=keep(bias, store)),
             O.ForceDesc(type=4, strengths=keep(s, store), theta2=0.81, distance_min2=1.0, distance_max2=np.inf),
             O.ForceDesc(type=1, strength=1.0)]
    sim = O.Sim(pos, ndim=3)
    sim.set_forces(descs, jacobi=True)
    sim.tick(300)
    op, _ = sim.state()
    want = float(np.sqrt(np.mean(np.sum(op ** 2, 1))))
    PARITY.append({"case": "A 1k RMS radius after 300 ticks", "gpu": rms, "oracle_jacobi": want, "published_gpu": 374})
    print(f"[parity] A 1k RMS radius after 300 ticks: GPU {rms:.2f}, FP64 Jacobi {want:.2f}, published 374", flush=True)
    assert abs(rms - want) < 1.0
    assert abs(rms - 374) < 1.5

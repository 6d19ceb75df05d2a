"""The batch engine (fg_batch_*): many layouts ticked together.

  * forces of every job in a mixed batch (1D/2D/3D, sizes 300..6000) against the FP64
    oracle from identical f32 input, to the single-simulation bars (rel-L2 <= 1e-4 for
    the tree force);
  * a job advances bit for bit the same alone and in a mix (the serving guarantee),
    with and without graph capture and under any tunables;
  * the batch and the single-simulation engine agree job by job over many ticks
    (bit for bit for link + manyBody; within rounding where center and collide sum in
    another order);
  * jobs finish after exactly the tick d3's schedule says (alpha < alphaMin), and the
    HBM budget refuses what does not fit.
"""
import numpy as np
import pytest

import d3fgpu
import oracle as O
from bench import ba_graph, link_defaults, phyllotaxis
from helpers import PARITY, close_count, make_nodes_fast, rel_l2, tree_outliers

pytestmark = pytest.mark.gpu

DECAY = d3fgpu.ALPHA_DECAY


def layout_forces(n, ndim, seed, collide=True, center=True, manybody=True, link=True, radius=3.0):
    """Link (BA m=2, d3's defaults) + manyBody(-30) + collide(radius) + center; radius None:
    lognormal radii (mu ln 2, sigma 0.4), the heavy tail that takes the two-class grid."""
    fs = []
    if link and n > 1:
        src, dst = ba_graph(n, 2, seed)
        bias, strengths, distances = link_defaults(n, src, dst)
        fs.append(d3fgpu.link(src, dst, strengths, distances, bias))
    if manybody:
        fs.append(d3fgpu.many_body(np.full(n, -30.0)))
    if collide:
        if radius is None:
            r = np.exp(np.log(2.0) + 0.4 * np.random.default_rng(seed).standard_normal(n))
        else:
            r = np.full(n, radius)
        fs.append(d3fgpu.collide(r))
    if center:
        fs.append(d3fgpu.center())
    return fs


def make_job(n, ndim, seed, **kw):
    return d3fgpu.Job(phyllotaxis(n, ndim)[:, :ndim], layout_forces(n, ndim, seed, **kw), ndim=ndim,
                      seed=seed)


def run_alone(job, ticks, **tune):
    b = d3fgpu.Batch(0)
    if tune:
        b.set_tune(**tune)
    jid = b.submit(job)
    b.step(ticks)
    out = b.read(jid)
    b.close()
    return out


def test_batch_forces_vs_oracle():
    """One tick from zero velocity with no velocity decay: v = the force increment."""
    sizes = [(300, 3), (1500, 2), (6000, 3), (700, 1), (4000, 2)]
    b = d3fgpu.Batch(0)
    jobs = {}
    for k, (n, nd) in enumerate(sizes):
        pos = make_nodes_fast(n, nd, seed=100 + k)
        s = np.full(n, -30.0) * (1 + 0.5 * np.sin(np.arange(n)))
        job = d3fgpu.Job(pos[:, :nd], [d3fgpu.many_body(s)], ndim=nd, velocity_decay=0.0, seed=k)
        jobs[b.submit(job)] = (pos, s, nd)
    b.step(1)
    alpha = 1.0 + (0.0 - 1.0) * DECAY
    for jid, (pos, s, nd) in jobs.items():
        _, v = b.read(jid)
        ref, _ = O.manybody(pos, np.zeros_like(pos), s.astype(np.float32).astype(np.float64), alpha,
                            theta2=0.81, dmin2=1.0, dmax2=np.inf, ndim=nd)
        err = rel_l2(v[:, :nd], ref[:, :nd])
        PARITY.append({"case": f"batch manyBody {nd}D n={len(pos)}", "rel_l2": err})
        print(f"[parity] batch manyBody {nd}D n={len(pos)}: rel-L2 {err:.3e}", flush=True)
        assert close_count(v, ref, rtol=1e-3) <= tree_outliers(len(pos))
        assert err < 1e-4
    st = b.stats()
    assert st["walk_errors"] == 0
    b.close()


def oracle_tick(pos, ndim, descs, alpha_decay=DECAY):
    """One d3 tick (alpha, forces in order, integrate with no velocity decay) in the FP64
    oracle, link and collide with the GPU's Jacobi semantics: (positions, velocities)."""
    sim = O.Sim(pos, ndim=ndim)
    sim.set_forces(descs, jacobi=True)
    sim.set_params(alpha=1.0, alpha_decay=alpha_decay, decay=1.0)
    sim.tick(1)
    return sim.state()


def config_d_job(n, nd, seed, kinds):
    """A config-D layout from d3's start: BA m=2 links with d3's defaults, manyBody(-30,
    theta 0.9), collide(radius 3), center -- or the subset `kinds` -- as a GPU job and as
    oracle descriptors (same order)."""
    pos = phyllotaxis(n, nd)
    if kinds == ("collide",):  # d3's start spaces points wider than radius-3 circles reach
        pos = pos / 4
    pos = pos.astype(np.float32).astype(np.float64)
    gpu, descs, store = [], [], []
    if "link" in kinds:
        src, dst = ba_graph(n, 2, seed)
        bias, strengths, distances = link_defaults(n, src, dst)
        gpu.append(d3fgpu.link(src, dst, strengths, distances, bias))
        links = np.ascontiguousarray(np.stack([src, dst], 1).reshape(-1), np.uint32)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (strengths, distances, bias)]
        store += [links] + arrs
        descs.append(O.ForceDesc(type=3, iterations=1, count=len(src), links=O.P(links, O.C.c_uint32),
                                 strengths=O.P(arrs[0]), distances=O.P(arrs[1]), bias=O.P(arrs[2])))
    if "manyBody" in kinds:
        s = np.full(n, -30.0)
        store.append(s)
        gpu.append(d3fgpu.many_body(s))
        descs.append(O.ForceDesc(type=4, strengths=O.P(s), theta2=0.81, distance_min2=1.0, distance_max2=np.inf))
    if "collide" in kinds:
        r = np.full(n, 3.0)
        store.append(r)
        gpu.append(d3fgpu.collide(r))
        descs.append(O.ForceDesc(type=2, iterations=1, strength=1.0, radii=O.P(r)))
    if "center" in kinds:
        gpu.append(d3fgpu.center())
        descs.append(O.ForceDesc(type=1, strength=1.0))
    job = d3fgpu.Job(pos[:, :nd], gpu, ndim=nd, velocity_decay=0.0, seed=seed)
    return job, pos, descs, store


def test_config_d_batch_vs_oracle():
    """Config D's batch (BA m=2 layouts from d3's start, link + manyBody(theta 0.9) +
    collide(3) + center, 2D and 3D): a 50,000-node 3D job and a 20,000-node 2D job among
    smaller ones (1D too), and 5,000-node jobs with one force each (collide alone from d3's
    start scaled by 1/4, so that the circles overlap).  One tick with no
    velocity decay, so each job's velocities are its force increments: every job against
    the FP64 oracle's tick, velocities to rel-L2 <= 1e-4 and positions (moved by center and
    the increments, then rounded to f32) to rel-L2 <= 1e-6."""
    ALL = ("link", "manyBody", "collide", "center")
    specs = [(50000, 3, ALL), (20000, 2, ALL), (5000, 3, ("link",)), (5000, 3, ("collide",)),
             (5000, 3, ("center",)), (5000, 3, ("manyBody",)), (5000, 2, ("collide",)), (3000, 2, ALL),
             (800, 1, ALL), (500, 3, ALL)]
    b = d3fgpu.Batch(0)
    jobs = []
    for k, (n, nd, kinds) in enumerate(specs):
        job, pos, descs, store = config_d_job(n, nd, 40 + k, kinds)
        jobs.append((b.submit(job), n, nd, kinds, pos, descs, store))
    b.step(1)
    assert b.stats()["walk_errors"] == 0
    for jid, n, nd, kinds, pos, descs, store in jobs:
        p, v = b.read(jid)
        rp, rv = oracle_tick(pos, nd, descs)
        ev = rel_l2(v[:, :nd], rv[:, :nd])
        ep = rel_l2(p[:, :nd], rp[:, :nd])
        moved = rel_l2(pos[:, :nd], rp[:, :nd])  # how far the oracle's tick moved the layout
        name = f"config D batch {nd}D n={n} {'+'.join(kinds)}"
        PARITY.append({"case": name, "rel_l2": ev, "rel_l2_pos": ep, "moved": moved})
        print(f"[parity] {name}: velocities rel-L2 {ev:.3e}, positions {ep:.3e} (moved {moved:.2e})", flush=True)
        assert ev < 1e-4 and ep < 1e-6 and ep <= 1e-2 * moved
        if "manyBody" in kinds and len(kinds) == 1:
            assert close_count(v, rv, rtol=1e-3) <= tree_outliers(n)
    b.close()


def test_batch_alone_equals_mix():
    """Each job of a mixed batch ends bit for bit where it ends when run alone."""
    specs = [(2500, 3, 1), (800, 2, 2), (6000, 3, 3), (1200, 2, 4), (400, 3, 5), (3000, 1, 6)]
    radius = [3.0, None, None, 3.0, None, None]  # None: lognormal radii (two-class grid when n >= 1024)
    ticks = 25
    mix = d3fgpu.Batch(0)
    ids = [mix.submit(make_job(n, nd, sd, radius=r)) for (n, nd, sd), r in zip(specs, radius)]
    mix.step(ticks)
    for jid, (n, nd, sd), r in zip(ids, specs, radius):
        p_mix, v_mix = mix.read(jid)
        p_one, v_one = run_alone(make_job(n, nd, sd, radius=r), ticks)
        assert np.all(np.isfinite(p_mix))
        assert np.array_equal(p_mix, p_one), f"job n={n} {nd}D differs alone vs mix"
        assert np.array_equal(v_mix, v_one)
    mix.close()


def test_batch_graphs_and_tunables_change_nothing():
    specs = [(3000, 3, 11), (900, 2, 12), (5000, 3, 13)]
    ticks = 12
    results = []
    for graphs, tune in [(True, {}), (False, {}),
                         (True, dict(radix_bits=11, walk_waves=8, walk_stack=96)),
                         (True, dict(radix_bits=8, walk_waves=2, walk_stack=32))]:
        b = d3fgpu.Batch(0)
        b.use_graphs(graphs)
        if tune:
            b.set_tune(**tune)
        ids = [b.submit(make_job(n, nd, sd)) for n, nd, sd in specs]
        b.step(ticks)
        results.append([b.read(j)[0] for j in ids])
        assert b.stats()["walk_errors"] == 0
        b.close()
    for other in results[1:]:
        for a, c in zip(results[0], other):
            assert np.array_equal(a, c)


def plain_run(n, nd, seed, ticks, **kw):
    e = d3fgpu.Engine(0)
    e.set_nodes(phyllotaxis(n, nd), ndim=nd)
    e.set_forces(layout_forces(n, nd, seed, **kw))
    e.set_seed(seed)
    alphas, a = [], 1.0
    for _ in range(ticks):
        a += (0.0 - a) * DECAY
        alphas.append(a)
    e.enqueue(alphas, 0.6)
    p, v = e.readback()
    e.close()
    return p, v


@pytest.mark.parametrize("n,nd,ticks", [(3000, 3, 30), (2000, 2, 30), (170000, 3, 3)])
def test_batch_equals_single_simulation_link_manybody(monkeypatch, n, nd, ticks):
    """The batch walks every job with 64-point tiles in the cheaper mode (the pool of
    single-point openings): bit for bit the single simulation with that walk (tiles of 64,
    no exact mode), and within rounding of the single simulation's default walk for such
    sizes (one or two points per wave, the same terms summed in another order) after one
    tick: Barnes-Hut's opening decisions are not continuous, so over many ticks such
    rounding differences grow to the approximation's own size."""
    p2, v2 = run_alone(make_job(n, nd, 21, collide=False, center=False), ticks)
    _, v0 = plain_run(n, nd, 21, 1, collide=False, center=False)
    _, v2_1 = run_alone(make_job(n, nd, 21, collide=False, center=False), 1)
    assert rel_l2(v0, v2_1) < 1e-6
    monkeypatch.setenv("D3FGPU_WALK_WIDE", "0")
    monkeypatch.setenv("D3FGPU_WALK_GROUP", "1")
    monkeypatch.setenv("D3FGPU_WALK_EXACT", "0")
    p1, v1 = plain_run(n, nd, 21, ticks, collide=False, center=False)
    assert np.array_equal(p1, p2)
    assert np.array_equal(v1, v2)


@pytest.mark.parametrize("n,nd,radius,ticks,tol", [(9000, 3, 3.0, 10, 1e-4), (3000, 2, 3.0, 5, 1e-5),
                                                   (12000, 3, None, 10, 1e-4)])
def test_batch_close_to_single_simulation_full_mix(monkeypatch, n, nd, radius, ticks, tol):
    """center sums and collide visits in another order: agreement to rounding (the single
    simulation with the batch's walk, 64-point tiles in the cheaper mode).

    Below 8192 nodes the single simulation collides all pairs and the batch through its
    grid, so the sums round differently from the first tick (rel-L2 ~5e-9 after one tick
    at 3000 nodes in 2D); the overlap-resolving collide dynamics then multiply such a
    difference several times a tick (to ~1e-4 after 10 ticks, a value set by chance
    rounding), so that case is held to 1e-5 after 5 ticks."""
    monkeypatch.setenv("D3FGPU_WALK_WIDE", "0")
    monkeypatch.setenv("D3FGPU_WALK_GROUP", "1")
    monkeypatch.setenv("D3FGPU_WALK_EXACT", "0")
    p1, _ = plain_run(n, nd, 31, ticks, radius=radius)
    p2, _ = run_alone(make_job(n, nd, 31, radius=radius), ticks)
    err = rel_l2(p2 - phyllotaxis(n, nd), p1 - phyllotaxis(n, nd))
    print(f"[parity] batch vs single, {nd}D n={n}, displacement after {ticks} ticks: rel-L2 {err:.2e}")
    assert err < tol


def test_batch_finish_at_alpha_min():
    """Jobs stop after the tick whose alpha falls below alphaMin (simulation.js:54-73)."""
    cases = [(1.0, 0.001, DECAY), (0.3, 0.01, 0.05), (1.0, 0.2, 0.1), (0.5, 0.001, 0.0228)]
    b = d3fgpu.Batch(0)
    ids = []
    want = []
    for k, (a0, amin, dec) in enumerate(cases):
        job = make_job(200 + 50 * k, 2 + k % 2, 40 + k)
        job.alpha, job.alpha_min, job.alpha_decay = a0, amin, dec
        ids.append(b.submit(job))
        a, t = a0, 0
        while True:
            a += (0.0 - a) * dec
            t += 1
            if a < amin:
                break
        want.append((t, a))
    b.step(max(t for t, _ in want) + 5)
    info = {j["id"]: j for j in b.jobs()}
    for jid, (t, a) in zip(ids, want):
        assert info[jid]["status"] == d3fgpu.JOB_DONE
        assert info[jid]["ticks"] == t
        assert info[jid]["alpha"] == a
    b.close()


def test_batch_budget_and_release():
    job = make_job(20000, 3, 51)
    need = job.bytes()
    b = d3fgpu.Batch(0, budget=int(need * 1.5))
    j1 = b.submit(job)
    assert b.try_submit(make_job(20000, 3, 52)) is None
    b.step(2)
    b.release(j1)
    j2 = b.try_submit(make_job(20000, 3, 52))
    assert j2 is not None
    b.step(2)
    p, _ = b.read(j2)
    assert np.all(np.isfinite(p))
    st = b.stats()
    assert st["bytes"] <= st["budget"]
    b.close()


def test_batch_autotune_keeps_results(tmp_path):
    specs = [(4000, 3, 61), (1500, 2, 62)]
    ticks = 8
    b = d3fgpu.Batch(0)
    ids = [b.submit(make_job(n, nd, sd)) for n, nd, sd in specs]
    cache = str(tmp_path / "tune.txt")
    chosen = b.autotune(cache)
    assert chosen["radix_bits"] in (8, 11) and chosen["walk_waves"] in (2, 4, 8)
    b.step(ticks)
    got = [b.read(j)[0] for j in ids]
    b.close()
    for (n, nd, sd), p in zip(specs, got):
        q, _ = run_alone(make_job(n, nd, sd), ticks)
        assert np.array_equal(p, q)
    # the cache answers the same shape without timing
    b = d3fgpu.Batch(0)
    for n, nd, sd in specs:
        b.submit(make_job(n, nd, sd))
    assert b.autotune(cache) == chosen
    b.close()


def test_batch_walk_spill():
    """LDS stack of 16 entries: walks deeper than that keep going in the global spill area
    and give the same forces as a 48-entry LDS stack."""
    n, nd = 60000, 3
    out = []
    for stack in (48, 16):
        b = d3fgpu.Batch(0)
        b.set_tune(walk_stack=stack)
        jid = b.submit(make_job(n, nd, 71, collide=False))
        b.step(3)
        st = b.stats()
        assert st["walk_errors"] == 0
        if stack == 16:
            assert st["walk_max_stack"] > 16
        out.append(b.read(jid))
        b.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])

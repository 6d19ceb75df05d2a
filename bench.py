"""Benchmark: ms per tick of one 1,000,000-node 3D simulation on one MI355X.

Workload (the 1M-node configuration of the benchmark set, synthetic):
  * 1,000,000 nodes, 3D, d3's phyllotaxis start positions (src/simulation.js:79-107)
  * links from examples/benchmark.html:86-93 (LCG seed 42): node i > 0 to a random
    earlier node, then n/2 random pairs -> 1,499,999 links, d3's default strength,
    distance 30 and bias (src/link.js:56-97)
  * forces in this order: link, manyBody(-30, theta 0.9), center, collide (radii
    exp(ln 2 + 0.4 N(0,1)), numpy PCG64 seed 7, 2 iterations), radial(r 200, strength 0.05)
  * 1% of nodes pinned (every 100th index) at uniform positions in [-500, 500]^3 (seed 11)
  * default alpha schedule (alpha 1, alphaDecay 1 - 0.001^(1/300)), velocityDecay 0.4

`value` is wall-clock ms per tick over `--steps` ticks (inputs resident in HBM, no
readback inside the timed region).  A second, event-timed run gives the per-phase split,
the walk kernel's time for its VALU roofline and each HBM-bound phase's GB/s.  The CPU
baseline is forceSimulationGPU's CPU path -- d3-force-3d's physics in FP64, JavaScript on one
thread -- timed in Node on this host (a bounded sample of the same workload, and the
1,000-node benchmark graph in full); the C FP64 restatement of the same tick is timed beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "d3-force-3d-webgpu_amd", "py"))

METRIC = "ms/tick @1M nodes 3D; batched layouts/s + p50/p99 latency vs plain, 1 GPU"
FLOP_PER_INTERACTION = 20          # n-body convention for one body-source evaluation
FP32_PEAK_TFLOPS = 157.3           # MI355X FP32 (vector and dense matrix)
HBM_PEAK_GBS = 8000.0              # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def collide_radii(n):
    """The workload's collide radii: exp(ln 2 + 0.4 N(0, 1)), numpy PCG64 seed 7."""
    return np.exp(np.log(2.0) + 0.4 * np.random.default_rng(7).standard_normal(n))


def pinned(n):
    """The workload's pins: every 100th node at a uniform position in [-500, 500]^3 (seed 11)."""
    idx = np.arange(0, n, 100)
    return idx, np.random.default_rng(11).uniform(-500, 500, (len(idx), 3))


def make_workload(n, with_collide=True, with_radial=True, with_pins=True):
    """Node positions, the force list and pins of the 1M-node configuration."""
    import d3fgpu
    pos = phyllotaxis3(n)
    src, dst = make_graph(n, 42)
    bias, strengths, distances = link_defaults(n, src, dst)
    forces = [d3fgpu.link(src, dst, strengths, distances, bias, iterations=1),
              d3fgpu.many_body(np.full(n, -30.0), theta=0.9),
              d3fgpu.center()]
    if with_collide:
        forces.append(d3fgpu.collide(collide_radii(n), strength=1.0, iterations=2))
    if with_radial:
        forces.append(d3fgpu.radial(np.full(n, 200.0), np.full(n, 0.05)))
    fix = mask = None
    if with_pins:
        idx, at = pinned(n)
        mask = np.zeros(n, np.uint8)
        mask[idx] = 7
        fix = np.zeros((n, 3))
        fix[idx] = at
        pos[idx] = at
    return pos, forces, fix, mask, len(src)


def phyllotaxis3(n):
    """src/simulation.js:85-100, 3D branch (vectorised)."""
    i = np.arange(n, dtype=np.float64)
    radius = 10 * np.cbrt(0.5 + i)
    roll = i * (np.pi * (3 - np.sqrt(5)))
    yaw = i * (np.pi * 20 / (9 + np.sqrt(221)))
    return np.stack([radius * np.sin(roll) * np.cos(yaw), radius * np.cos(roll),
                     radius * np.sin(roll) * np.sin(yaw)], 1)


def make_graph(n, seed):
    """examples/benchmark.html:86-93 with the LCG of src/lcg.js."""
    s = seed
    m = n - 1 + (n + 1) // 2
    src = np.zeros(m, np.uint32)
    dst = np.zeros(m, np.uint32)
    k = 0
    a, c, M = 1664525, 1013904223, 4294967296
    for i in range(1, n):
        s = (a * s + c) % M
        src[k], dst[k] = i, int(s / M * i)
        k += 1
    for _ in range((n + 1) // 2):
        s = (a * s + c) % M
        u = int(s / M * n)
        s = (a * s + c) % M
        src[k], dst[k] = u, int(s / M * n)
        k += 1
    return src, dst


def ba_graph(n, m, seed):
    """Barabasi-Albert graph: node i >= 1 attaches m links, each to an endpoint drawn uniformly
    from the endpoints of all earlier links (linear preferential attachment; node 1 attaches to
    node 0).  Vectorised: a link whose draw lands on an earlier link's target takes that link's
    target, resolved by pointer jumping."""
    if n < 2:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    rng = np.random.default_rng(seed)
    M = m * (n - 1)
    e = np.arange(M, dtype=np.int64)
    src = 1 + e // m
    slots = 2 * m * (src - 1)                      # endpoints of the earlier nodes' links
    q = np.floor(rng.random(M) * slots).astype(np.int64)
    tgt = np.zeros(M, np.int64)
    done = slots == 0
    ptr = q.copy()
    while not done.all():
        todo = ~done
        even = todo & (ptr % 2 == 0)
        tgt[even] = src[ptr[even] // 2]
        done |= even
        todo = ~done
        par = ptr[todo] // 2
        ready = done[par]
        idx = np.flatnonzero(todo)
        tgt[idx[ready]] = tgt[par[ready]]
        done[idx[ready]] = True
        ptr[idx[~ready]] = q[par[~ready]]
        first = idx[~ready][slots[par[~ready]] == 0]   # the earlier link was node 1's: target 0
        tgt[first] = 0
        done[first] = True
    return src.astype(np.uint32), tgt.astype(np.uint32)


def phyllotaxis(n, ndim):
    """d3's initial positions in 1, 2 or 3 dimensions (src/simulation.js:85-100)."""
    i = np.arange(n, dtype=np.float64)
    out = np.zeros((n, 3))
    if ndim == 3:
        out[:] = phyllotaxis3(n)
    elif ndim == 2:
        r = 10 * np.sqrt(0.5 + i)
        a = i * (np.pi * (3 - np.sqrt(5)))
        out[:, 0], out[:, 1] = r * np.cos(a), r * np.sin(a)
    else:
        out[:, 0] = 10 * i
    return out


def link_defaults(n, src, dst):
    """src/link.js:65-97: bias, default strength 1 / min(count), distance 30."""
    count = np.bincount(src, minlength=n).astype(np.float64) + np.bincount(dst, minlength=n)
    cs, ct = count[src], count[dst]
    return cs / (cs + ct), 1.0 / np.minimum(cs, ct), np.full(len(src), 30.0)


def alphas(start, count, decay=1 - 0.001 ** (1 / 300)):
    out, a = [], start
    for _ in range(count):
        a += (0 - a) * decay
        out.append(a)
    return out, a


def run_gpu(n, steps, warmup, light=False):
    import d3fgpu
    t0 = time.time()
    pos, forces, fix, mask, m = make_workload(n, with_collide=not light, with_radial=not light,
                                              with_pins=not light)
    log(f"workload built in {time.time() - t0:.1f}s: n={n} links={m} forces={len(forces)}")
    e = d3fgpu.Engine(0)
    e.set_nodes(pos, fix=fix, fixmask=mask, ndim=3)
    e.set_forces(forces)
    e.set_seed(0x2545F491)
    alpha = 1.0
    a, alpha = alphas(alpha, warmup)
    e.enqueue(a, 0.6)
    e.sync()
    log(f"warmup {warmup} ticks done")
    a, alpha = alphas(alpha, steps)
    e.sync()
    t0 = time.perf_counter()
    e.enqueue(a, 0.6)
    e.sync()
    wall = time.perf_counter() - t0
    ms = 1e3 * wall / steps
    log(f"timed {steps} ticks: {ms:.3f} ms/tick")
    # event-timed run for the per-phase split and the walk kernel's roofline
    k = max(5, min(steps, 20))
    a, alpha = alphas(alpha, k)
    e.timing_enable(True)
    e.enqueue(a, 0.6)
    e.sync()
    t = e.timing()
    st = e.stats()
    e.timing_enable(False)
    # device -> pinned host -> FP64 arrays, PCIe included: the first readback (it allocates the
    # pinned staging buffer, and fresh arrays take their page faults), then one into those arrays
    t0 = time.perf_counter()
    p, v = e.readback()
    readback_first_ms = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    p, v = e.readback(out=(p, v))
    readback_ms = 1e3 * (time.perf_counter() - t0)
    finite = bool(np.all(np.isfinite(p)))
    e.close()
    phases = {name: round(v[0] / k, 4) for name, v in t.items() if v[1]}
    walk_ms = t["walk"][0] / max(1, t["walk"][1])
    return dict(ms=ms, phases=phases, walk_ms=walk_ms, walk_launches=t["walk"][1], interactions=int(st.interactions),
                opened=int(st.opened), lane_visits=int(st.lane_visits), walk_errors=int(st.walk_errors), links=m,
                finite=finite, timed_ticks=k, readback_ms=readback_ms, readback_first_ms=readback_first_ms, max_stack=int(st.walk_max_stack))


def batch_specs(jobs, seed=2024):
    """The batched workload: `jobs` layouts, sizes floor(exp(U(ln 500, ln 50000))), a quarter
    2D and the rest 3D, Barabasi-Albert graphs with m = 2 seeded per job."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(500), np.log(50000), jobs))).astype(int)
    dims = np.array([2] * (jobs // 4) + [3] * (jobs - jobs // 4))
    rng.shuffle(dims)
    return [(int(n), int(d), seed + 1 + k) for k, (n, d) in enumerate(zip(sizes, dims))]


def batch_forces(n, seed):
    """link (d3 defaults) + manyBody(-30, theta 0.9) + collide(radius 3) + center."""
    import d3fgpu
    src, dst = ba_graph(n, 2, seed)
    bias, strengths, distances = link_defaults(n, src, dst)
    return [d3fgpu.link(src, dst, strengths, distances, bias), d3fgpu.many_body(np.full(n, -30.0), theta=0.9),
            d3fgpu.collide(np.full(n, 3.0)), d3fgpu.center()]


def run_batched(specs, cache=None):
    """All layouts resident in one batch, ticked together until every one is done."""
    import d3fgpu
    jobs = [d3fgpu.Job(phyllotaxis(n, d)[:, :d], batch_forces(n, sd), ndim=d, seed=sd) for n, d, sd in specs]
    t0 = time.perf_counter()
    b = d3fgpu.Batch(0)
    ids = [b.submit(j) for j in jobs]
    t_submit = time.perf_counter() - t0
    tune = b.autotune(cache)
    t_tune = time.perf_counter() - t0 - t_submit
    ticks = 0
    while True:
        b.step(50)
        ticks += 50
        if all(j["status"] == d3fgpu.JOB_DONE for j in b.jobs()):
            break
    out = {jid: b.read(jid)[0] for jid in ids}
    wall = time.perf_counter() - t0
    st = b.stats()
    b.close()
    return dict(wall=wall, submit_s=t_submit, tune_s=t_tune, ticks=ticks, tune=tune, stats=st,
                layouts=[out[j] for j in ids])


# The single simulation's walk overrides that leave every job on the tiled walk in its cheaper
# mode (the walk limits are read when an engine is created).
TILED_WALK_ENV = {"D3FGPU_WALK_WIDE": "0", "D3FGPU_WALK_GROUP": "1", "D3FGPU_WALK_EXACT": "0"}


def run_plain_jobs(specs, env=None):
    """The single-simulation engine, job after job, each to alpha < alphaMin (`env`: variables
    set while the engine is created)."""
    import d3fgpu
    saved = {k: os.environ.get(k) for k in (env or {})}
    os.environ.update(env or {})
    try:
        e = d3fgpu.Engine(0)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    outs = []
    t0 = time.perf_counter()
    for n, d, sd in specs:
        e.set_nodes(phyllotaxis(n, d), ndim=d)
        e.set_forces(batch_forces(n, sd))
        e.set_seed(sd)
        a, alphas = 1.0, []
        while True:
            a += (0.0 - a) * (1 - 0.001 ** (1 / 300))
            alphas.append(a)
            if a < 0.001:
                break
        e.enqueue(alphas, 0.6)
        outs.append(e.readback()[0])
    wall = time.perf_counter() - t0
    e.close()
    return dict(wall=wall, layouts=outs)


def bench_batch(jobs, with_plain=True):
    specs = batch_specs(jobs)
    nodes = sum(n for n, _, _ in specs)
    log(f"batched: {jobs} layouts, {nodes} nodes")
    bt = run_batched(specs, cache=os.path.join(ROOT, "gpurun_out", "batch_tune.txt"))
    log(f"batched: {bt['wall']:.2f}s for {jobs} layouts ({bt['ticks']} ticks stepped, tune {bt['tune']})")
    out = {"jobs": jobs, "nodes": nodes, "dims": "1:3 2D:3D", "graph": "Barabasi-Albert m=2",
           "forces": "link + manyBody(theta 0.9) + collide(radius 3) + center",
           "layouts_per_s": round(jobs / bt["wall"], 3), "wall_s": round(bt["wall"], 3),
           "submit_s": round(bt["submit_s"], 3), "autotune_s": round(bt["tune_s"], 3),
           "ticks_stepped": bt["ticks"], "tune": bt["tune"],
           "launches_per_tick": bt["stats"]["launches_per_tick"], "walk_errors": bt["stats"]["walk_errors"]}
    if with_plain:
        pl = run_plain_jobs(specs)
        log(f"plain job by job: {pl['wall']:.2f}s")
        # control: the same engine again on the first jobs with only its walk changed (the sum
        # of each point's forces taken in another order); over the ticks to alpha < alphaMin
        # the layouts move apart by as much as the batch's do from the plain ones
        few = min(64, len(specs))
        other = run_plain_jobs(specs[:few], TILED_WALK_ENV)
        log(f"plain job by job, tiled walk, first {few}: {other['wall']:.2f}s")

        def spread(xs, ys, sp):
            errs = []
            for (n, d, sd), a, c in zip(sp, xs, ys):
                den = np.linalg.norm(c - phyllotaxis(n, d))
                errs.append(float(np.linalg.norm(a - c) / (den if den > 0 else 1.0)))
            return {"max": max(errs), "median": float(np.median(errs))}
        out.update({"plain_layouts_per_s": round(jobs / pl["wall"], 3), "plain_wall_s": round(pl["wall"], 3),
                    "speedup_vs_plain": round(pl["wall"] / bt["wall"], 2),
                    "layout_rel_l2_vs_plain": spread(bt["layouts"], pl["layouts"], specs),
                    "layout_rel_l2_plain_tiled_walk_vs_plain": dict(
                        jobs=few, **spread(other["layouts"], pl["layouts"][:few], specs[:few]))})
    return out


def serve_specs(jobs, seed, lo=1000, hi=500_000):
    """The serving stream's jobs: sizes log-uniform in [lo, hi], 70% 3D, Barabasi-Albert with
    m in {1, 2, 3}, each force of {link, manyBody, collide, center, radial} present with
    probability 0.6 (at least one), in that order."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(jobs):
        n = int(np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        d = 3 if rng.random() < 0.7 else 2
        m = int(rng.integers(1, 4))
        mix = rng.random(5) < 0.6
        if not mix.any():
            mix[int(rng.integers(0, 5))] = True
        out.append(dict(n=n, ndim=d, m=m, mix=[bool(x) for x in mix], seed=seed * 1000 + k))
    return out


def serve_job(spec):
    """A d3fgpu.Job for one serving spec."""
    import d3fgpu
    n, d, sd = spec["n"], spec["ndim"], spec["seed"]
    fs = []
    if spec["mix"][0] and n > 1:
        src, dst = ba_graph(n, spec["m"], sd)
        bias, strengths, distances = link_defaults(n, src, dst)
        fs.append(d3fgpu.link(src, dst, strengths, distances, bias))
    if spec["mix"][1]:
        fs.append(d3fgpu.many_body(np.full(n, -30.0), theta=0.9))
    if spec["mix"][2]:
        fs.append(d3fgpu.collide(np.full(n, 3.0)))
    if spec["mix"][3]:
        fs.append(d3fgpu.center())
    if spec["mix"][4]:
        fs.append(d3fgpu.radial(np.full(n, 10.0 * n ** (1.0 / d)), np.full(n, 0.05)))
    return d3fgpu.Job(phyllotaxis(n, d)[:, :d], fs, ndim=d, seed=sd)


def serve_run(jobs, rate, budget, seed=7):
    """Poisson arrivals at `rate` jobs/s (inf: all at once) into a server; every job runs to
    alpha < alphaMin.  Returns the server's latency and throughput figures."""
    import threading
    import d3fgpu
    srv = d3fgpu.Server(0, budget=budget, ticks_per_round=10)
    rng = np.random.default_rng(seed)
    gaps = np.zeros(len(jobs)) if not np.isfinite(rate) else rng.exponential(1.0 / rate, len(jobs))
    ids = []
    t0 = time.perf_counter()
    due = 0.0
    for job, gap in zip(jobs, gaps):
        due += gap
        delay = due - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(job))
    for j in ids:
        srv.wait(j, timeout=900.0)
    st = srv.stats()
    srv.close()
    return st


def bench_serve(jobs, budget_gb=1.0):
    specs = serve_specs(jobs, seed=5)
    log(f"serving: building {jobs} jobs ({sum(s['n'] for s in specs)} nodes)")
    built = [serve_job(s) for s in specs]
    budget = int(budget_gb * 2 ** 30)
    runs = []
    capacity = None
    for frac in (None, 0.8, 0.5):
        # all jobs at once first (the capacity), then Poisson arrivals at fractions of it
        rate = float("inf") if frac is None else frac * capacity
        st = serve_run(built, rate, budget)
        if frac is None:
            capacity = st["layouts_per_s"]
        log(f"serving at {rate} jobs/s: p50 {st['p50_latency_s']:.3f}s p99 {st['p99_latency_s']:.3f}s "
            f"{st['layouts_per_s']:.2f} layouts/s, peak running {st['peak_running']}")
        runs.append({"arrival_rate_per_s": None if not np.isfinite(rate) else round(rate, 3),
                     "load": "all at once" if frac is None else f"{frac:.0%} of capacity",
                     "p50_latency_s": round(st["p50_latency_s"], 4), "p99_latency_s": round(st["p99_latency_s"], 4),
                     "mean_latency_s": round(st["mean_latency_s"], 4),
                     "layouts_per_s": round(st["layouts_per_s"], 3), "finished": st["finished"],
                     "failed": st["failed"], "peak_running": st["peak_running"],
                     "peak_bytes_gb": round(st["peak_bytes"] / 2 ** 30, 3)})
    return {"jobs": jobs, "nodes": sum(s["n"] for s in specs), "sizes": "log-uniform [1e3, 5e5]",
            "dims": "70% 3D", "graph": "Barabasi-Albert m in {1,2,3}",
            "forces": "each of link/manyBody/collide/center/radial with p=0.6",
            "hbm_budget_gb": budget_gb, "ticks_per_round": 10, "runs": runs}


def bench_config_b(n=200_000, steps=20, warmup=5):
    """200,000 nodes, 3D, link + manyBody + center at theta 0.9 and theta 0 (exact all-pairs):
    ms/tick, the tree sort's throughput, interactions/s of the walk and of the all-pairs sum."""
    import d3fgpu
    src, dst = make_graph(n, 42)
    bias, strengths, distances = link_defaults(n, src, dst)
    out = {"nodes": n, "links": len(src)}
    for theta in (0.9, 0.0):
        e = d3fgpu.Engine(0)
        e.set_nodes(phyllotaxis3(n), ndim=3)
        e.set_forces([d3fgpu.link(src, dst, strengths, distances, bias), d3fgpu.many_body(np.full(n, -30.0), theta=theta),
                      d3fgpu.center()])
        a, alpha = alphas(1.0, warmup)
        e.enqueue(a, 0.6)
        e.sync()
        a, alpha = alphas(alpha, steps)
        t0 = time.perf_counter()
        e.enqueue(a, 0.6)
        e.sync()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        k = 5
        a, alpha = alphas(alpha, k)
        e.timing_enable(True)
        e.enqueue(a, 0.6)
        e.sync()
        t = e.timing()
        st = e.stats()
        e.timing_enable(False)
        e.close()
        key = "theta_0.9" if theta else "theta_0"
        leg = {"ms_per_tick": round(ms, 4), "phases_ms_per_tick": {nm: round(v[0] / k, 4) for nm, v in t.items() if v[1]}}
        if theta:
            sort_ms = t["tree_sort"][0] / max(1, t["tree_sort"][1])
            walk_ms = t["walk"][0] / max(1, t["walk"][1])
            passes = 4  # the 30-bit high halves of the 60-bit codes, 8-bit digits (then the runs are ordered)
            leg.update({"sort_passes": passes, "sort_ms": round(sort_ms, 4),
                        "sort_gb_per_s": round(24 * n * passes / (sort_ms * 1e-3) / 1e9, 1),
                        "sort_note": "u64 code + u32 index per element, read and written each pass: 24 B/element/pass; "
                                     "sort_ms also holds the run ordering",
                        "walk_ms": round(walk_ms, 4), "interactions": int(st.interactions),
                        "walk_interactions_per_s": round(st.interactions / (walk_ms * 1e-3), 1)})
        else:
            pairs_ms = t["pairs"][0] / max(1, t["pairs"][1])
            inter = n * (n - 1)
            tf = inter * FLOP_PER_INTERACTION / (pairs_ms * 1e-3) / 1e12
            leg.update({"pairs_ms": round(pairs_ms, 4), "interactions": inter,
                        "pairs_interactions_per_s": round(inter / (pairs_ms * 1e-3), 1),
                        "pairs_tflops": round(tf, 2), "pairs_roofline_frac": round(tf / FP32_PEAK_TFLOPS, 4)})
        out[key] = leg
        log(f"config B theta {theta}: {ms:.3f} ms/tick")
    return out


# The reference's own size table (README.md:86-92): link + charge + center, 3D, ms/tick on an
# RTX 4090 (WebGPU).
REF_SIZES_MS = {1000: 0.067, 5000: 0.10, 20000: 0.18, 50000: 0.27, 100000: 0.37, 200000: 0.71, 1000000: 2.6}


def bench_ref_sizes(sizes=(1000, 5000, 20000, 50000, 100000, 200000)):
    """The reference's benchmark sizes: link + manyBody(theta 0.9) + center on its benchmark
    graph (examples/benchmark.html:86-93) from d3's phyllotaxis start, ms/tick after 10
    warmup ticks (no readback in the timed region)."""
    import d3fgpu
    out = {}
    for n in sizes:
        src, dst = make_graph(n, 42)
        bias, strengths, distances = link_defaults(n, src, dst)
        e = d3fgpu.Engine(0)
        e.set_nodes(phyllotaxis3(n), ndim=3)
        e.set_forces([d3fgpu.link(src, dst, strengths, distances, bias), d3fgpu.many_body(np.full(n, -30.0)),
                      d3fgpu.center()])
        a, alpha = alphas(1.0, 10)
        e.enqueue(a, 0.6)
        e.sync()
        steps = 300 if n <= 50_000 else (150 if n <= 100_000 else 100)
        a, alpha = alphas(alpha, steps)
        t0 = time.perf_counter()
        e.enqueue(a, 0.6)
        e.sync()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        e.close()
        out[str(n)] = {"ms_per_tick": round(ms, 4), "reference_ms": REF_SIZES_MS[n],
                       "vs_reference": round(REF_SIZES_MS[n] / ms, 3)}
        log(f"size {n}: {ms:.4f} ms/tick (reference {REF_SIZES_MS[n]})")
    return out


def cpu_baseline(n, ticks=1):
    """The FP64 restatement of d3-force-3d's tick (d3 semantics: sequential link and
    collide, d3-octree Barnes-Hut), one thread, on `ticks` ticks of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    pos = phyllotaxis3(n)
    src, dst = make_graph(n, 42)
    bias, strengths, distances = link_defaults(n, src, dst)
    rng = np.random.default_rng(7)
    radii = np.exp(np.log(2.0) + 0.4 * rng.standard_normal(n))
    rng = np.random.default_rng(11)
    fix = np.full((n, 3), np.nan)
    fix[::100] = rng.uniform(-500, 500, (len(fix[::100]), 3))
    pos[::100] = fix[::100]
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, np.float64)
        keep.append(a)
        return O.P(a)

    links = np.ascontiguousarray(np.stack([src, dst], 1).reshape(-1), np.uint32)
    keep.append(links)
    descs = [O.ForceDesc(type=3, iterations=1, count=len(src), links=O.P(links, O.C.c_uint32),
                         strengths=arr(strengths), distances=arr(distances), bias=arr(bias)),
             O.ForceDesc(type=4, strengths=arr(np.full(n, -30.0)), theta2=0.81, distance_min2=1.0,
                         distance_max2=np.inf),
             O.ForceDesc(type=1, strength=1.0),
             O.ForceDesc(type=2, iterations=2, strength=1.0, radii=arr(radii)),
             O.ForceDesc(type=6, strengths=arr(np.full(n, 0.05)), radii=arr(np.full(n, 200.0)))]
    sim = O.Sim(pos, fix=fix, ndim=3)
    sim.set_forces(descs)
    t0 = time.perf_counter()
    sim.tick(ticks)
    dt = time.perf_counter() - t0
    return 1e3 * dt / ticks


def node_cpu_baseline(n, light=False, ticks=None, min_ticks=3, budget=1.5):
    """forceSimulationGPU's CPU path -- d3-force-3d's physics in FP64 on one thread, the
    JavaScript the simulation runs without a GPU -- timed in Node on this host with the
    reference benchmark page's method (examples/benchmark.html:105-117: one untimed tick,
    then tick(1) calls, at least `min_ticks`, for `budget` seconds; or exactly `ticks`).
    light: link + manyBody + center only; otherwise the whole 1M-node mix with its pins."""
    import shutil
    import subprocess
    import tempfile
    node = shutil.which("node")
    if node is None:
        return None
    d = tempfile.mkdtemp(prefix="d3f-cpu-")
    try:
        src, dst = make_graph(n, 42)
        np.ascontiguousarray(np.stack([src, dst], 1), np.uint32).tofile(os.path.join(d, "links.u32"))
        case = {"n": n, "dims": 3, "forces": ["link", "manyBody", "center"], "min": min_ticks, "budget": budget}
        if ticks:
            case["ticks"] = ticks
        if not light:
            collide_radii(n).tofile(os.path.join(d, "radii.f64"))
            idx, at = pinned(n)
            np.ascontiguousarray(np.column_stack([idx, at]), np.float64).tofile(os.path.join(d, "pins.f64"))
            case["forces"] += ["collide", "radial"]
            case.update(collideIterations=2, radialRadius=200.0, radialStrength=0.05)
        with open(os.path.join(d, "case.json"), "w") as f:
            json.dump(case, f)
        script = os.path.join(ROOT, "d3-force-3d-webgpu_amd", "js", "cpu-benchmark.mjs")
        r = subprocess.run([node, script, d], capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            log("node CPU baseline failed: " + r.stderr[-2000:])
            return None
        return json.loads(r.stdout.strip().splitlines()[-1])
    finally:
        shutil.rmtree(d, ignore_errors=True)


def node_tick_async(sizes=(1000, 5000, 20000, 50000, 100000, 200000, 1000000)):
    """The reference's own GPU timing, through this backend's Node host: forceSimulationGPU
    with link (distance 30) + manyBody + center on the benchmark page's graph, three warm-up
    ticks, then await tickAsync(k) for k = 10, 20, ... until one call takes 0.75 s or k
    reaches 640 (examples/benchmark.html:119-132); each call holds its ticks, the readback
    and the update of the nodes in JavaScript."""
    import shutil
    import subprocess
    node = shutil.which("node")
    if node is None:
        return None
    script = os.path.join(ROOT, "d3-force-3d-webgpu_amd", "js", "gpu-benchmark.mjs")
    r = subprocess.run([node, script] + [str(n) for n in sizes], capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        log("node tickAsync leg failed: " + r.stderr[-2000:])
        return None
    out = {}
    for line in r.stdout.strip().splitlines():
        row = json.loads(line)
        ms = row["ms_per_tick"]
        out[str(row["n"])] = {"ms_per_tick": round(ms, 4), "ticks": row["ticks"], "links": row["links"],
                              "reference_ms": REF_SIZES_MS.get(row["n"]),
                              "vs_reference": round(REF_SIZES_MS[row["n"]] / ms, 3) if row["n"] in REF_SIZES_MS else None}
        log(f"node tickAsync {row['n']}: {ms:.4f} ms/tick over {row['ticks']} ticks")
    out["node"] = json.loads(r.stdout.strip().splitlines()[-1])["node"] if r.stdout.strip() else None
    return out


# Algorithmic HBM bytes per node of each phase of a 1M-node tick (each input element read
# once, each output written once), as this backend's kernels move them:
PASS_BYTES = {
    # bbox 12; keys 12 + 8 + 4 + 16; leaves + Karras nodes 16 + 16 + 4 + 8 + 16 + 4;
    # reduction 16 + 16 + 64 record; d3 children 16 + 32
    "tree_build": 12 + 40 + 64 + 96 + 48,
    # histogram 8; four 8-bit passes over (u64 code, u32 node) read and written, 24 each;
    # the run ordering 24 + 12
    "tree_sort": 8 + 4 * 24 + 36,
    # snapshot p + v -> float4 (24 + 16); CSR row 4, 2m/n = 3 entries of 20 B and the other
    # end's float4 snapshot 16, velocities read and written 24
    "link": 40 + 4 + 3 * 36 + 24,
    # two iterations of keys, counting build, scan, scatter, run order and query (~180 B)
    "collide": 2 * 180,
    # the mean is taken by the tree's box kernel (its read counted there); the shift applied
    "center": 24,
    "radial": 12 + 8 + 24,
    "integrate": 48 + 4,
}


def pass_rooflines(n, phases):
    """GB/s and fraction of 8 TB/s of each HBM-bound phase from its event-timed ms/tick."""
    out = {}
    for name, per_node in PASS_BYTES.items():
        ms = phases.get(name)
        if not ms:
            continue
        gbs = per_node * n / (ms * 1e-3) / 1e9
        out[name] = {"bytes_per_node": per_node, "ms": ms, "gb_per_s": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def copy_calibration():
    """A device copy of 256 MiB, three dispatches: what FETCH_SIZE and WRITE_SIZE read for a
    known number of bytes (the scale factor of the walk's counter readings)."""
    import torch
    a = torch.ones(64 * 1024 * 1024, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copied_bytes_per_dispatch": a.numel() * 4, "dispatches": 3}))


def walk_traffic(n):
    """HBM bytes per walk launch from the committed counter summary (1M nodes only)."""
    path = os.path.join(ROOT, "profiles", "tree_walk_1M_3d_hbm_counters.txt")
    if n != 1_000_000 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nodes", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ref-mix", action="store_true", help="skip the link+charge+center run")
    ap.add_argument("--batch-jobs", type=int, default=256, help="layouts in the batched leg (0: skip)")
    ap.add_argument("--no-plain-jobs", action="store_true", help="batched leg without the job-by-job run")
    ap.add_argument("--only-batch", action="store_true", help="run only the batched leg")
    ap.add_argument("--serve-jobs", type=int, default=48, help="jobs per serving run (0: skip)")
    ap.add_argument("--no-config-b", action="store_true", help="skip the 200k-node theta 0.9 / theta 0 leg")
    ap.add_argument("--no-ref-sizes", action="store_true", help="skip the reference's size table (1k..200k)")
    ap.add_argument("--no-node-gpu", action="store_true",
                    help="skip the reference's tickAsync timing through the Node host")
    ap.add_argument("--only-serve", action="store_true", help="run only the serving leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (counter calibration) and exit")
    ap.add_argument("--profile-ticks", type=int, default=0,
                    help="only run this many ticks of the --nodes workload (after 2 untimed) and exit, "
                         "for profiler runs")
    ap.add_argument("--light", action="store_true",
                    help="with --profile-ticks: link + manyBody + center only (the reference's size table)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.profile_ticks:
        r = run_gpu(args.nodes, args.profile_ticks, 2, light=args.light)
        print(json.dumps({"profile_ticks": args.profile_ticks, "ms_per_tick": round(r["ms"], 4)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs one simulation per GPU on one GPU; "
                                   "--gpus above 1 is not supported (see the sharding notes)"}))
        sys.exit(2)
    n = args.nodes
    if args.only_serve:
        print(json.dumps({"serving": bench_serve(args.serve_jobs)}))
        return
    if args.only_batch:
        print(json.dumps({"batched": bench_batch(args.batch_jobs, not args.no_plain_jobs)}))
        return
    r = run_gpu(n, args.steps, args.warmup)
    ref = None
    if not args.no_ref_mix:
        ref = run_gpu(n, args.steps, args.warmup, light=True)
    achieved = r["interactions"] * FLOP_PER_INTERACTION / (r["walk_ms"] * 1e-3) / 1e12 if r["walk_ms"] > 0 else 0.0
    cpu = None
    if not args.no_cpu:
        log("CPU baseline: forceSimulationGPU's CPU path in Node, a sample of the same workload ...")
        js = node_cpu_baseline(n, min_ticks=2, budget=1.5)
        log("CPU baseline: the 1,000-node benchmark graph, 300 ticks in Node ...")
        js_a = node_cpu_baseline(1000, light=True, ticks=300)
        log("CPU baseline: 1 tick of the same workload in the C FP64 restatement ...")
        c_ms = cpu_baseline(n, 1)
        c_port = {"value": round(c_ms, 1), "unit": "ms/tick", "cores": 1,
                  "sample": f"1 tick of the same {n}-node workload, C FP64 restatement of d3-force-3d's tick "
                            "(the parity oracle), one thread"}
        if js:
            cpu = {"value": round(js["ms_per_tick"], 1), "unit": "ms/tick", "cores": 1, "kind": "port",
                   "sample": f"forceSimulationGPU's CPU path (d3-force-3d's physics, FP64 JavaScript, one thread) in "
                             f"Node {js['node']}: one untimed tick then {js['ticks_timed']} timed ticks of the same "
                             f"{n}-node workload (examples/benchmark.html:105-117)",
                   "host_cpu": js["cpu_model"], "host_os_cpus": js["os_cpus"], "node": js["node"],
                   "config_a_ms_per_tick": round(js_a["ms_per_tick"], 3) if js_a else None,
                   "config_a_sample": "1,000-node benchmark graph, link + manyBody + center, 300 timed ticks after "
                                      "one untimed (the reference publishes 3.5 ms/tick, Ryzen 9 9950X3D, README.md:86)",
                   "c_port": c_port}
        else:
            cpu = dict(c_port, kind="port")
    out = {
        "metric": METRIC,
        "value": round(r["ms"], 4),
        "unit": "ms/tick",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "1M-node 3D layout: link + manyBody(theta 0.9) + center + collide(2 iterations, "
                               "lognormal radii) + radial(200, 0.05), 1% pinned",
                   "nodes": n, "links": r["links"], "dims": 3},
        "roofline": {"bound": "valu", "kernel": "tree_walk", "achieved": round(achieved, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 5), "traffic": walk_traffic(n),
                     "walk_ms": round(r["walk_ms"], 4), "interactions": r["interactions"],
                     "flop_per_interaction": FLOP_PER_INTERACTION,
                     "bound_note": "no dense contraction on this path, so no MFMA: the walk is bound by VALU issue; "
                                   "peak is MI355X's FP32 vector rate",
                     "passes": pass_rooflines(n, r["phases"])},
        "cpu_baseline": cpu,
        "phases_ms_per_tick": r["phases"],
        "walk_errors": r["walk_errors"],
        "walk_max_stack": r["max_stack"],
        "readback_ms": round(r["readback_ms"], 3),
        "readback_first_ms": round(r["readback_first_ms"], 3),
        "walk_waves_openings": r["opened"],
        "walk_lane_utilisation": round(r["lane_visits"] / max(1, 64 * r["opened"]), 4),
        "finite": r["finite"],
    }
    if not args.no_config_b:
        out["config_b"] = bench_config_b()
    if not args.no_ref_sizes:
        out["reference_sizes"] = bench_ref_sizes()
    if not args.no_node_gpu:
        out["node_tick_async"] = node_tick_async()
    if args.batch_jobs > 0:
        out["batched"] = bench_batch(args.batch_jobs, not args.no_plain_jobs)
    if args.serve_jobs > 0:
        out["serving"] = bench_serve(args.serve_jobs)
    if ref is not None:
        out["ref_mix_ms_per_tick"] = round(ref["ms"], 4)
        out["ref_mix_note"] = ("link + manyBody + center only, the mix of the published 2.6 ms/tick "
                               "at 1M nodes (RTX 4090, README.md:82,92)")
        out["ref_mix_vs_published"] = round(2.6 / ref["ms"], 3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

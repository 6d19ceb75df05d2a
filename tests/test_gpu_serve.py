"""The request loop (fg_serve_*): jobs that arrive while others run are admitted as memory
frees up, finish at alphaMin, and each result is bit for bit the one the job gives alone."""
import threading
import time

import numpy as np
import pytest

import d3fgpu
from bench import serve_job, serve_specs

pytestmark = pytest.mark.gpu


def alone(job):
    b = d3fgpu.Batch(0)
    jid = b.submit(job)
    while b.jobs()[0]["status"] != d3fgpu.JOB_DONE:
        b.step(25)
    out = b.read(jid)
    ticks = b.jobs()[0]["ticks"]
    b.close()
    return out, ticks


def test_serve_same_alone_and_in_mix():
    specs = serve_specs(10, seed=3, lo=300, hi=6000)
    jobs = [serve_job(s) for s in specs]
    budget = int(2.5 * max(j.bytes() for j in jobs))  # a few jobs at a time: admission waits
    srv = d3fgpu.Server(0, budget=budget, ticks_per_round=7)
    ids = []
    for k, j in enumerate(jobs):  # staggered arrivals: later jobs join a running batch
        ids.append(srv.submit(j))
        time.sleep(0.01 * (k % 3))
    results = [srv.wait(i, timeout=300.0) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["finished"] == len(jobs) and st["failed"] == 0
    assert st["peak_running"] >= 2
    assert st["p99_latency_s"] >= st["p50_latency_s"] > 0
    for job, (pos, vel, r) in zip(jobs, results):
        assert r["status"] == d3fgpu.SERVE_FINISHED
        (p1, v1), ticks = alone(job)
        assert r["ticks"] == ticks
        assert np.array_equal(pos, p1) and np.array_equal(vel, v1)


def test_serve_refuses_what_never_fits():
    big = serve_job(dict(n=5000, ndim=3, m=2, mix=[True, True, True, True, False], seed=9))
    small = serve_job(dict(n=200, ndim=2, m=1, mix=[True, True, False, True, False], seed=10))
    srv = d3fgpu.Server(0, budget=int(1.5 * small.bytes()))
    a = srv.submit(big)
    b = srv.submit(small)
    with pytest.raises(d3fgpu.FgError):
        srv.wait(a, timeout=120.0)
    pos, _, r = srv.wait(b, timeout=120.0)
    assert r["status"] == d3fgpu.SERVE_FINISHED and np.all(np.isfinite(pos))
    st = srv.stats()
    assert st["failed"] == 1 and st["finished"] == 1
    srv.close()


def test_serve_concurrent_submitters():
    specs = serve_specs(12, seed=4, lo=200, hi=3000)
    jobs = [serve_job(s) for s in specs]
    srv = d3fgpu.Server(0)
    ids = [None] * len(jobs)

    def push(k0):
        for k in range(k0, len(jobs), 3):
            ids[k] = srv.submit(jobs[k])

    ts = [threading.Thread(target=push, args=(k,)) for k in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    srv.drain(timeout=300.0)
    for i in ids:
        pos, _, r = srv.wait(i, timeout=10.0)
        assert r["status"] == d3fgpu.SERVE_FINISHED and np.all(np.isfinite(pos))
    srv.close()


def test_serve_config_e_large_job_alone_and_in_mix():
    """Config E's size range reaches 500,000 nodes: a 300,000-node 3D job (BA m=2, link +
    manyBody + collide + center + radial) served among smaller jobs arriving while it runs
    ends bit for bit where it ends alone, after the same number of ticks."""
    big = dict(n=300_000, ndim=3, m=2, mix=[True, True, True, True, True], seed=77)
    specs = [big] + serve_specs(6, seed=5, lo=1000, hi=20000)
    jobs = [serve_job(s) for s in specs]
    srv = d3fgpu.Server(0, ticks_per_round=10)
    ids = [srv.submit(jobs[0])]
    time.sleep(0.05)
    for j in jobs[1:]:
        ids.append(srv.submit(j))
        time.sleep(0.02)
    results = [srv.wait(i, timeout=300.0) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["finished"] == len(jobs) and st["failed"] == 0 and st["peak_running"] >= 2
    pos, vel, r = results[0]
    assert r["status"] == d3fgpu.SERVE_FINISHED and np.all(np.isfinite(pos))
    (p1, v1), ticks = alone(jobs[0])
    assert r["ticks"] == ticks
    assert np.array_equal(pos, p1) and np.array_equal(vel, v1)

"""Runs the Node host's tests: the API without a GPU here, the whole host on a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "d3-force-3d-webgpu_amd", "lib", "d3fgpu.node")


def run_node(script, timeout):
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", script)], capture_output=True, text=True,
                       timeout=timeout, cwd=ROOT)
    print(r.stdout)
    print(r.stderr[-4000:])
    assert r.returncode == 0, r.stdout[-4000:]


@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_js_api_without_gpu():
    assert os.path.exists(ADDON), "the N-API addon is not built"
    run_node("api_test.mjs", 120)


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_js_host_on_gpu():
    run_node("gpu_test.mjs", 240)


@pytest.mark.skipif(NODE is None, reason="node is not installed")
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_js_cpu_path_matches_oracle(tmp_path, dims):
    """The Node host's CPU path (js/cpu.js: d3's physics in FP64, taken for cpuTicks and when
    no GPU can run) against the C oracle's d3 tick, sequential link and collide, over 20
    ticks: link + manyBody + collide + center + radial + x with a pinned node."""
    import json
    import numpy as np
    import oracle as O
    from bench import make_graph, link_defaults
    n, ticks = 1000, 20
    rng = np.random.default_rng(dims)
    pos = np.zeros((n, 3))
    pos[:, :dims] = (rng.random((n, dims)) - 0.5) * 300
    src, dst = make_graph(n, 42)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    pins = [[3, 5.0, -5.0, 0.0]]
    case = dict(n=n, dims=dims, ticks=ticks, positions=pos.tolist(), pins=pins,
                links=np.stack([src, dst], 1).tolist())
    inp, out = tmp_path / "case.json", tmp_path / "out.json"
    inp.write_text(json.dumps(case))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "cpu_tick.mjs"), str(inp), str(out)],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-3000:]
    got = json.loads(out.read_text())
    bias, st, di = link_defaults(n, src, dst)
    radii = 1.0 + (np.arange(n) % 5)
    store = []

    def arr(a):
        a = np.ascontiguousarray(a, np.float64)
        store.append(a)
        return O.P(a)

    links = np.ascontiguousarray(np.stack([src, dst], 1).reshape(-1), np.uint32)
    store.append(links)
    descs = [O.ForceDesc(type=3, iterations=1, count=len(src), links=O.P(links, O.C.c_uint32), strengths=arr(st),
                         distances=arr(di), bias=arr(bias)),
             O.ForceDesc(type=4, strengths=arr(np.full(n, -30.0)), theta2=0.81, distance_min2=1.0,
                         distance_max2=np.inf),
             O.ForceDesc(type=2, iterations=1, strength=1.0, radii=arr(radii)),
             O.ForceDesc(type=1, strength=1.0),
             O.ForceDesc(type=6, radii=arr(np.full(n, 80.0)), strengths=arr(np.full(n, 0.05))),
             O.ForceDesc(type=5, axis=0, targets=arr(np.full(n, 10.0)), strengths=arr(np.full(n, 0.02)))]
    fix = np.full((n, 3), np.nan)
    fix[3, :dims] = pins[0][1:1 + dims]
    p0 = pos.copy()
    p0[3, :dims] = fix[3, :dims]
    sim = O.Sim(p0, fix=fix, ndim=dims)
    sim.set_forces(descs)
    sim.tick(ticks)
    op, ov = sim.state()
    gp, gv = np.array(got["pos"]), np.array(got["vel"])
    assert abs(got["alpha"] - sim.alpha) == 0
    np.testing.assert_allclose(gp[:, :dims], op[:, :dims], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(gv[:, :dims], ov[:, :dims], rtol=1e-9, atol=1e-9)

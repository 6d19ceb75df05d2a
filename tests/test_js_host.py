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

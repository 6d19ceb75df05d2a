import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "d3-force-3d-webgpu_amd", "py"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")


@pytest.fixture(scope="session")
def engine_factory():
    import d3fgpu
    made = []

    def make():
        e = d3fgpu.Engine(0)
        made.append(e)
        return e

    yield make
    for e in made:
        e.close()


def pytest_sessionfinish(session, exitstatus):
    """Write the rel-L2 figures the GPU parity tests measured to parity.json in the run's
    output folder."""
    import json
    try:
        from helpers import PARITY
    except Exception:
        return
    if PARITY:
        out = os.path.join(ROOT, "gpurun_out")
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, "parity.json"), "w") as f:
            json.dump(PARITY, f, indent=1)

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "d3-force-3d-webgpu_amd", "py"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


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

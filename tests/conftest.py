import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libaph.so on the GPU)")
    # The oracle is cheap to build; build it when absent (e.g. a fresh checkout).
    if not os.path.exists(os.path.join(ROOT, "oracle", "_build", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def teapot_path():
    return os.path.join(GOLDEN, "teapot.npz")


@pytest.fixture(scope="session")
def renderer():
    from aurora_amd.session import Renderer

    r = Renderer(0)
    yield r
    r.close()

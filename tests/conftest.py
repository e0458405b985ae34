import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtopopt_hip.so)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """Build the oracle (and the HIP library) only if they are not built yet."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "topopt_amd", "libtopopt_hip.so")):
        from topopt_amd import build
        build.build()
    yield


@pytest.fixture(autouse=True)
def _fresh_mg_latch(request):
    """Each GPU test starts as a new run: the multigrid stagnation latch is
    per thread and sticky (the reference's driver resets it at run start), and
    every test runs on the same thread."""
    if request.node.get_closest_marker("gpu") is not None:
        import topopt_amd
        topopt_amd.reset_mg_stagnation_latch()
    yield

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def engine():
    import fastlinkingnumbers_amd as flk
    eng = flk.Engine(0)
    eng.default_tune = eng.get_tune()
    yield eng
    eng.close()


@pytest.fixture(autouse=True)
def _restore_tune(request):
    """A test that changes the shared engine's launch shapes does not leak
    them into later tests: the default shapes are restored after each one."""
    yield
    if "engine" in request.fixturenames:
        eng = request.getfixturevalue("engine")
        if getattr(eng, "h", None):
            eng.set_tune(**eng.default_tune)

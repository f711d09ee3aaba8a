import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def titanic():
    from sinksim_amd.data_paths import titanic as load
    return load("titanic")


@pytest.fixture(scope="session")
def goldens():
    """The reference's three golden traces of titanic64/titanic, keyed by numerics."""
    from sinksim_amd.trace import load_trace
    out = {}
    for numerics, name in (("portable", "titanic.portable"), ("portable32", "titanic.portable32"), ("oracle", "titanic")):
        out[numerics] = load_trace(os.path.join(GOLDEN, name + ".trace.json.gz"))
    return out

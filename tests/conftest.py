import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import gputimeseriescompressionoptimizer_amd as g

    c = g.Context(0)
    yield c
    c.close()

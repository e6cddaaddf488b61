import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def ctx():
    from engine_quant_amd import eqhip
    if not eqhip.is_gpu_backend_available():
        pytest.fail("GPU test run without a gfx950 device: the HIP path must be the one that runs")
    c = eqhip.Context(0)
    yield c
    c.close()

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer statistical runs")


@pytest.fixture(scope="session")
def pyro():
    import pyro_cmc_amd

    pyro_cmc_amd.lib()  # raises if the HIP library has not been built
    return pyro_cmc_amd


@pytest.fixture(scope="session")
def gpu(pyro):
    if pyro.device_count() < 1:
        raise RuntimeError("GPU test started without a visible HIP device")
    return pyro


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    pyoracle.lib()
    return pyoracle

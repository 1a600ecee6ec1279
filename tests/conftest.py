import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for sub in ("gsplat-rs_amd", "oracle", "tests"):
    p = str(ROOT / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def abi():
    import gsplat_abi

    gsplat_abi.lib()
    return gsplat_abi


@pytest.fixture(scope="session")
def oracle():
    import oracle_py

    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def gpu(abi):
    if not abi.device_available():
        pytest.fail("GPU test on a host without a visible gfx950 device")
    return abi

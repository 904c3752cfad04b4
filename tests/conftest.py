import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    if not os.path.exists(os.path.join(oracle.HERE, "liboracle.so")):
        oracle.build(ref=False)
    oracle.set_threads(1)
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail (not skip) when it is missing."""
    import rochpcg_amd
    if rochpcg_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return rochpcg_amd

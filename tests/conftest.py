import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libespec_hip.so)")
    config.addinivalue_line("markers", "slow: long CPU-only check")


@pytest.fixture(scope="session")
def orc():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import Reference
    try:
        return Reference()
    except OSError as e:
        pytest.skip(f"reference build not present: {e}")


@pytest.fixture(scope="session")
def gpu():
    import espec_amd
    n = espec_amd.device_count()
    if n < 1:
        pytest.fail("no HIP device visible to libespec_hip.so")
    return espec_amd

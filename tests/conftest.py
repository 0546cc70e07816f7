import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "dplus-dev_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def gpu_available():
    try:
        import dplus_mi355x as D
        return bool(D.device_info().get("ok"))
    except Exception:
        return False


@pytest.fixture(scope="session")
def dplus():
    import dplus_mi355x as D
    if not D.device_info().get("ok"):
        pytest.fail("the HIP device is not usable: GPU tests must run on the GPU")
    return D


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build(with_ref=False)
    return O

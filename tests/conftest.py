import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libndipc.so")


def gpu_available():
    try:
        import torch  # noqa: F401  (device count only; no GPU initialisation here)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def engine_factory():
    if not gpu_available():
        pytest.skip("no HIP device")
    from nexdynipc_amd._lib import Engine

    engines = []

    def make(sc, params=None):
        e = Engine(0)
        e.upload(sc, params)
        engines.append(e)
        return e

    yield make
    for e in engines:
        e.close()

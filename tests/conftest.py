import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI of libiqm_hip.so)")


@pytest.fixture(scope="session")
def ctx():
    import iqm_amd

    c = iqm_amd.Context(0)
    print(f"\n[device] {c.device_name}", flush=True)
    yield c
    c.close()

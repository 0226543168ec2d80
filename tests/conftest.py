import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP path)")


@pytest.fixture(scope="session")
def ctx():
    from visor_amd import render, vk
    icd = vk.load()
    if not icd.vh_device_available():
        pytest.fail("no HIP device: GPU tests need the MI355X")
    return render.Context()

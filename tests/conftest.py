import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libccnn.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle.bind import Oracle
    return Oracle()


@pytest.fixture(scope="session", autouse=True)
def _torch_device_first(request):
    # torch's HIP runtime must initialise before libccnn's in a process (the other order
    # leaves torch without a device), so a run that holds GPU tests sets torch up first
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch
        if torch.cuda.is_available():
            torch.zeros(1, device="cuda:0")
    yield

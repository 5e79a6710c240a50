import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long-running test (full-size configurations)")


@pytest.fixture(scope="session")
def rm_ctx():
    """A library context on cuda:0 (GPU tests only). Fails loudly if the HIP library or the
    device is missing: there is no CPU fallback to pass on."""
    import torch
    from renderer_ray_marching_amd import api
    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    return api.Context(0)

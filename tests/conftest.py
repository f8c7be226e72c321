import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    """Device context for GPU tests: (torch, queue, allocation cache)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import opensycl_amd as acpp
    cache = acpp.AllocationCache()
    yield torch, cache
    torch.cuda.synchronize()
    cache.close()


@pytest.fixture()
def q(gpu):
    import opensycl_amd as acpp
    return acpp.Queue()


@pytest.fixture()
def scratch(gpu):
    import opensycl_amd as acpp
    torch, cache = gpu
    g = acpp.AllocationGroup(cache, 0)
    yield g
    torch.cuda.synchronize()
    g.close()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """PyTorch bundles its own HIP and HSA runtimes; in a process that also loads
    libsvo_amd.so (linked to the system ROCm), PyTorch finds the GPU only if it
    initializes first. GPU sessions therefore start PyTorch's device before any test."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    yield


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    """One device context for the GPU tests (fails loudly without the HIP library)."""
    from ooc_svo_builder_amd import SvoContext
    ctx = SvoContext(0)
    yield ctx
    ctx.close()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test collected without a GPU"
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def vctx(torch_dev):
    """One library context on torch's current stream, for the whole session."""
    import torch
    import velodb_amd
    ctx = velodb_amd.Context(stream=torch.cuda.current_stream().cuda_stream)
    yield ctx
    ctx.close()

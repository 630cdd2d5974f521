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


def pytest_collection_modifyitems(config, items):
    """The C++ operator suite (a child process) starts resident queues, which
    are cooperative launches; on this runtime a process cannot make one on a
    GPU once another process has.  Its tests therefore run before every
    other selected test, so that this process has made none yet."""
    first = [it for it in items if os.path.basename(str(it.fspath)) == "test_gpu_cpp_operators.py"]
    if first:
        items[:] = first + [it for it in items if it not in first]

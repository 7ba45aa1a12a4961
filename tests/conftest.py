import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libmsamd.so")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mountainsort_amd import _capi
    _capi.lib()  # raises MsamdLibraryMissing when the HIP library was not built
    return torch.device("cuda:0")


def say(*args):
    """Print and flush a line naming the case about to run on the GPU."""
    print(*args, flush=True)

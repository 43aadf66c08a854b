import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def cuda():
    """The GPU device for gpu-marked tests; a missing GPU or library is an error, not a skip."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("gpu test requires a GPU")
    import tandem_amd._lib as L

    L.lib()  # raises if libtandem_dg.so has not been built
    return torch.device("cuda:0")

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "jblas.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libjblas_hip.so on cuda:0)")


@pytest.fixture(scope="session")
def gpu():
    """The GPU session: torch on cuda:0 plus the library initialised on device 0.
    A GPU test run without a GPU fails here (never skips)."""
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    import jblas_amd

    jblas_amd.init(0)
    return torch


def to_dev(torch, a, dtype=None):
    """numpy column-major matrix -> column-major cuda tensor (stride(0) == 1)."""
    import numpy as np

    a = np.asfortranarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda")
    if dtype is not None:
        t = t.to(dtype)
    return t.T


def to_host(t):
    import numpy as np

    return np.asfortranarray(t.T.contiguous().cpu().numpy().T)

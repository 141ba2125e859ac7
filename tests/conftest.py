import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


@pytest.fixture(scope="session")
def gd():
    """The product package (libgdict.so binding)."""
    return entry.load_package()


@pytest.fixture(scope="session")
def orc():
    """The oracle (checker)."""
    return entry.load_oracle()


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")


class DeviceIO:
    """Host<->device helpers for 32/64-bit unsigned arrays."""

    def __init__(self, device):
        import torch
        self.torch = torch
        self.device = device

    def up(self, a, dtype=np.uint32):
        a = np.ascontiguousarray(a, dtype=dtype)
        view = np.int32 if a.dtype.itemsize == 4 else np.int64
        if a.dtype == np.uint8:
            view = np.uint8
        return self.torch.from_numpy(a.view(view)).to(self.device)

    def empty(self, n, itemsize=4):
        dt = {1: self.torch.uint8, 4: self.torch.int32,
              8: self.torch.int64}[itemsize]
        # poison so that unwritten outputs are visible
        return self.torch.full((max(n, 1),), -7 if itemsize > 1 else 0xAB,
                               dtype=dt, device=self.device)

    def down(self, t, n=None, dtype=np.uint32):
        self.torch.cuda.synchronize()
        a = t.cpu().numpy()
        if n is not None:
            a = a[:n]
        return a.view(dtype)


@pytest.fixture(scope="session")
def io(dev):
    return DeviceIO(dev)

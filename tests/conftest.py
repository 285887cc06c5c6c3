import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import __graft_entry__  # noqa: E402
import oracle as _oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def orc():
    _oracle.lib()
    return _oracle


@pytest.fixture(scope="session")
def blas():
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def dev():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def to_dev(view: np.ndarray, dev):
    """Copy a (possibly padded) numpy storage view to the GPU, keeping its row
    stride, padding included, and return the device view of the same shape."""
    import torch

    base = view
    while base.base is not None and isinstance(base.base, np.ndarray):
        base = base.base
    rows, cols = view.shape
    if base.ndim == 2 and base.shape[1] == view.strides[0] // view.itemsize:
        t = torch.from_numpy(np.ascontiguousarray(base)).to(dev)
        return t[:rows, :cols]
    t = torch.from_numpy(np.ascontiguousarray(view)).to(dev)
    return t

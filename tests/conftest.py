import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


def load_basis(name):
    with np.load(os.path.join(GOLDEN, "basis_%s.npz" % name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def load_npz(fname):
    with np.load(os.path.join(GOLDEN, fname), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")


def make_cpab(name):
    """Cpab on the GPU with the committed basis of tessellation `name`."""
    from libcpab_amd import Cpab
    b = load_basis(name)
    nc = [int(v) for v in b["nc"]]
    return Cpab(nc, backend="pytorch", device="gpu", zero_boundary=bool(b["zero_boundary"]), basis=b["B"]), b

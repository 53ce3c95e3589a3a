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


def _chunks(n, k):
    step = max(1, (n + k - 1) // k)
    return [(i, min(i + step, n)) for i in range(0, n, step)]


def par_cf_forward(ndim, nc, A, planes, nbr, points, workers=8):
    """oracle.cf_forward over point slices on a thread pool (the C oracle runs
    without the GIL); points [ndim, nP] shared by every theta of A."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle as O
    points = np.ascontiguousarray(points, dtype=np.float64)
    parts = _chunks(points.shape[-1], workers)
    with ThreadPoolExecutor(len(parts)) as ex:
        outs = list(ex.map(lambda r: O.cf_forward(ndim, nc, A, planes, nbr,
                                                  np.ascontiguousarray(points[:, r[0]:r[1]])), parts))
    return np.concatenate(outs, axis=-1)


def par_cf_backward(ndim, nc, A, planes, nbr, points, grad_out, workers=8):
    """oracle.cf_backward over point slices on a thread pool: dL/dA summed
    over the slices, dL/dpoints concatenated."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle as O
    points = np.ascontiguousarray(points, dtype=np.float64)
    grad_out = np.asarray(grad_out, dtype=np.float64)
    parts = _chunks(points.shape[-1], workers)

    def one(r):
        return O.cf_backward(ndim, nc, A, planes, nbr, np.ascontiguousarray(points[:, r[0]:r[1]]),
                             np.ascontiguousarray(grad_out[..., r[0]:r[1]]))

    with ThreadPoolExecutor(len(parts)) as ex:
        res = list(ex.map(one, parts))
    return sum(r[0] for r in res), np.concatenate([r[1] for r in res], axis=-1)

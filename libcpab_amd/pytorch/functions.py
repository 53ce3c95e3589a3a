"""The 'pytorch' backend plugin module: the same module-level function set the
reference's Cpab calls through ``self.backend.<fn>``
(libcpab/pytorch/functions.py:8-129), with the integration, interpolation,
cell lookup and velocity field running as HIP kernels on gfx950.
"""
import ctypes

import torch

from .. import _native as N
from . import hip_ops
from .hip_ops import Interpolate, TransformData, TransformGrid


def assert_version():
    major, minor = torch.__version__.split(".")[:2]
    assert float(major + "." + minor) >= 1.0, "pytorch >= 1.0 is required"


def _device(device):
    if isinstance(device, str):
        return torch.device("cuda") if device == "gpu" else torch.device("cpu")
    return device


def to(x, dtype=torch.float32, device=None):
    device = _device(device)
    if isinstance(x, torch.Tensor):
        return x.to(dtype=dtype, device=device if device is not None else x.device)
    return torch.tensor(x, dtype=dtype, device=device)


def tonumpy(x):
    return x.detach().cpu().numpy()


def check_device(x, device_name):
    return x.is_cuda == (device_name == "gpu")


def backend_type():
    return torch.Tensor


def pdist(mat):
    norm = torch.sum(mat * mat, 1).reshape(-1, 1)
    return norm - 2 * mat.mm(mat.t()) + norm.t()


def norm(x):
    return torch.norm(x)


def matmul(x, y):
    return torch.matmul(x, y)


def transpose(x):
    return x.t()


def exp(x):
    return torch.exp(x)


def zeros(*s, device=None):
    return torch.zeros(*s, device=_device(device))


def ones(*s, device=None):
    return torch.ones(*s, device=_device(device))


def arange(x):
    return torch.arange(x)


def repeat(x, reps):
    return x.repeat(reps)


def batch_repeat(x, reps):
    return x.repeat(reps, *(x.dim() * [1]))


def maximum(x):
    return x.max()


def sample_transformation(d, n_sample=1, mean=None, cov=None, device="cpu"):
    device = _device(device)
    mean = torch.zeros(d, dtype=torch.float32, device=device) if mean is None else mean
    cov = torch.eye(d, dtype=torch.float32, device=device) if cov is None else cov
    dist = torch.distributions.MultivariateNormal(mean, cov)
    return dist.sample((n_sample,)).to(device)


def identity(d, n_sample=1, epsilon=0, device="cpu"):
    assert epsilon >= 0, "epsilon need to be larger than 0"
    return torch.zeros(n_sample, d, dtype=torch.float32, device=_device(device)) + epsilon


def uniform_meshgrid(ndim, domain_min, domain_max, n_points, device="cpu"):
    """[ndim, prod(n_points)], x varying fastest (libcpab/pytorch/functions.py:102-108)."""
    device = _device(device)
    lin = [torch.linspace(domain_min[i], domain_max[i], n_points[i], device=device) for i in range(ndim)]
    mesh = torch.meshgrid(lin[::-1], indexing="ij")
    grid = torch.cat([g.reshape(1, -1) for g in mesh[::-1]], dim=0)
    if ndim == 2:
        # traversal hint for the grid transform (2D tiles); it changes no result
        grid._cpab_grid_shape = (int(n_points[0]), int(n_points[1]))
    return grid


def transformer(points, theta, params):
    """Integrate points under theta (replaces CPAB_transformer,
    libcpab/pytorch/transformer.py:72-87).  Returns [n_theta, ndim, nP].
    use_slow selects the reference's pure-torch stepping path (its autograd
    gradient, w.r.t. theta and the points); otherwise method='closed_form'
    (exact integration) or 'numeric' (the reference's compiled fast path)."""
    if getattr(params, "use_slow", False):
        from .slow import transformer_slow
        return transformer_slow(points, theta, params)
    method = getattr(params, "method", "closed_form")
    if method != "closed_form":
        from .numeric import transformer_numeric
        return transformer_numeric(points, theta, params)
    shape = getattr(points, "_cpab_grid_shape", None)
    out = TransformGrid.apply(points, theta, params, 1.0, None, shape)
    if shape is not None:
        out._cpab_grid_shape = shape  # warped meshgrid points keep the meshgrid's order
    return out


def interpolate(ndim, data, grid, outsize):
    """Linear interpolation, output [n_batch, C, *outsize]
    (libcpab/pytorch/interpolation.py:11-171)."""
    assert grid.shape[0] == data.shape[0], "grid and data must have the same batch size"
    return Interpolate.apply(data, grid, ndim, list(outsize))


def transform_data(data, theta, params, outsize):
    """Fused meshgrid -> integration -> interpolation (Cpab.transform_data)."""
    assert data.shape[0] == theta.shape[0], "data and theta must have the same batch size"
    return TransformData.apply(data, theta, params, list(outsize), 1.0)


def findcellidx(ndim, p, nc):
    """Cell index of each point [ndim, nP] (int64), numpy-backend semantics."""
    hip_ops._require_gpu_f32("points", p)
    p = p.contiguous()
    idx = torch.empty(p.shape[-1], dtype=torch.int32, device=p.device)
    N.check(N.lib().cpab_findcell(ndim, hip_ops._int_arr(nc), hip_ops._ptr(p), p.shape[-1], hip_ops._ptr(idx),
                                  hip_ops._stream(p.device)))
    return idx.to(torch.int64)


def calc_vectorfield(grid, theta, params):
    """Velocity [ndim, nP] of a single theta at grid points."""
    hip_ops._require_gpu_f32("grid", grid)
    hip_ops._require_gpu_f32("theta", theta)
    dev = hip_ops.device_tess(params, theta.device)
    grid = grid.contiguous()
    th = theta.reshape(1, -1).contiguous()
    nP = grid.shape[-1]
    v = torch.empty(dev.ndim, nP, device=grid.device, dtype=torch.float32)
    ws, nb = dev.workspace(1, nP)
    N.check(N.lib().cpab_vectorfield(dev.ref, hip_ops._ptr(th), hip_ops._ptr(grid), nP, hip_ops._ptr(v),
                                     hip_ops._ptr(ws), nb, hip_ops._stream(grid.device)))
    return v

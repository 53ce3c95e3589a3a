"""The fused transform_data (one HIP kernel forward, one backward) against the
reference's own composition in plain PyTorch fp32: the grid from
transform_grid, then libcpab's interpolation written in torch ops
(libcpab/pytorch/interpolation.py:18-171: scale by size-1, floor, clamp both
corners, lerp axis 0 first, reshape / permute to [N, C, out_0, out_1, ...]),
differentiated by torch autograd (index backward for dL/ddata, the xd / yd / zd
path for dL/dgrid, then transform_grid's backward for dL/dtheta).

Tolerances.  Both paths integrate in float32; their warped grids differ by
at most DELTA = 4e-7, twice the largest error against the FP64 flow measured
on these configurations (<= 2e-7; the two grids differ by about one ulp in
practice).  A grid difference of DELTA moves a sample by DELTA * (size - 1)
pixels per axis, so with amp = DELTA * sum(size - 1):
  forward      |out difference| <= amp * range(data) + 1e-6;
  dL/ddata     per cell, <= amp * (sum of |grad_out| over the points reading
               the cell) + 1e-6 (each bilinear weight moves by <= amp);
  dL/dtheta    1e-3 relative, the path's gradient tolerance: a grid
               difference across a pixel edge switches a point's corner pair,
               which changes its d out / d grid by O(1)."""
import pytest
import torch

from conftest import make_cpab
from torch_interp_ref import torch_interpolate, touching_mass

pytestmark = pytest.mark.gpu

DELTA = 4e-7


CASES = [
    # the image workload's tessellation and image size, 8 of its 256 thetas
    ("2d_4x4", 8, 0.3, 3, [256, 256], [256, 256]),
    # upsampling, downsampling (the LDS box overflows: global atomics), ragged
    ("2d_4x4", 3, 0.5, 3, [40, 50], [100, 90]),
    ("2d_4x4", 2, 0.3, 3, [512, 512], [64, 64]),
    ("2d_3x3", 4, 0.8, 2, [33, 7], [50, 5]),
    # zero_boundary=False: grid points leave [0, 1] and both corners clamp
    ("2d_2x2_nzb", 3, 0.7, 3, [64, 48], [64, 48]),
    ("1d_50", 64, 0.5, 1, [1024], [1024]),
    ("1d_10", 5, 1.0, 2, [300], [257]),
    ("3d_2x2x2", 2, 0.5, 1, [32, 32, 32], [32, 32, 32]),
    ("3d_2x2x2", 2, 0.4, 2, [20, 16, 12], [24, 16, 9]),
]


@pytest.mark.parametrize("name,n_theta,scale,C,dsize,outsize", CASES)
def test_fused_transform_data_matches_torch_composition(cuda, name, n_theta, scale, C, dsize, outsize):
    print("case", name, n_theta, scale, C, dsize, outsize, flush=True)
    T, _ = make_cpab(name)
    g = torch.Generator().manual_seed(101)
    theta = (torch.randn(n_theta, T.get_theta_dim(), generator=g) * scale).float()
    data = torch.rand([n_theta, C] + dsize, generator=g)
    gout = torch.randn([n_theta, C] + outsize, generator=g).to(cuda)

    th1 = theta.to(cuda).requires_grad_(True)
    d1 = data.to(cuda).requires_grad_(True)
    out1 = T.transform_data(d1, th1, outsize)
    (out1 * gout).sum().backward()

    th2 = theta.to(cuda).requires_grad_(True)
    d2 = data.to(cuda).requires_grad_(True)
    grid = T.transform_grid(T.uniform_meshgrid(outsize), th2)
    out2 = torch_interpolate(d2, grid, outsize)
    (out2 * gout).sum().backward()

    assert out1.shape == out2.shape == (n_theta, C, *outsize)
    amp = DELTA * sum(n - 1 for n in dsize)
    tol_out = amp * (data.max() - data.min()).item() + 1e-6
    assert (out1 - out2).abs().max().item() <= tol_out
    mass = touching_mass(grid.detach(), dsize, outsize, gout.abs())
    assert bool(((d1.grad - d2.grad).abs() <= amp * mass + 1e-6).all())
    gt1, gt2 = th1.grad.double(), th2.grad.double()
    assert ((gt1 - gt2).abs().max() / gt2.abs().max()).item() <= 1e-3

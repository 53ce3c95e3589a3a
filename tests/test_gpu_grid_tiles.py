"""The grid transform walks meshgrid points in 2D tiles when told the grid's
shape (CpabSequential.transform_data does): outputs bit-identical to the
run order, gradients equal up to the order of the per-tile sums."""
import pytest
import torch

from conftest import make_cpab
from libcpab_amd.pytorch.hip_ops import TransformGrid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [[64, 64], [100, 37], [37, 100], [256, 9], [33, 7]])
@pytest.mark.parametrize("per_theta,store", [(False, True), (False, False), (True, True)])
def test_tiled_grid_transform_equals_run_order(cuda, shape, per_theta, store):
    T, b = make_cpab("2d_4x4")
    g = torch.Generator().manual_seed(sum(shape))
    n_theta = 3
    theta = (torch.randn(n_theta, T.get_theta_dim(), generator=g) * 0.4).to(cuda)
    grid = T.uniform_meshgrid(shape)
    if per_theta:
        grid = T.transform_grid(grid, (torch.randn(n_theta, T.get_theta_dim(), generator=g) * 0.2).to(cuda))
    w = torch.randn(n_theta, 2, grid.shape[-1], generator=g).to(cuda)
    res = []
    for gs in (None, shape):
        th = theta.clone().requires_grad_(True)
        pts = grid.detach().clone().requires_grad_(True)
        out = TransformGrid.apply(pts, th, T.params, 1.0, store, gs)
        (out * w).sum().backward()
        res.append((out.detach(), th.grad, pts.grad))
    (o1, g1, p1), (o2, g2, p2) = res
    assert torch.equal(o1, o2)
    assert torch.allclose(g1, g2, rtol=1e-5, atol=1e-6 * g1.abs().max().item())
    assert torch.allclose(p1, p2, rtol=1e-5, atol=1e-6 * p1.abs().max().item())


def test_meshgrid_hint_reaches_transform_grid(cuda):
    """uniform_meshgrid tags 2D grids with their shape; transform_grid on such
    a grid (and on its warped output) takes the tiled order, with the same
    outputs as an untagged copy of the points."""
    T, b = make_cpab("2d_4x4")
    theta = (torch.randn(2, T.get_theta_dim(), generator=torch.Generator().manual_seed(2)) * 0.4).to(cuda)
    grid = T.uniform_meshgrid([100, 37])
    assert getattr(grid, "_cpab_grid_shape", None) == (100, 37)
    plain = grid.clone()
    out_tiled = T.transform_grid(grid, theta)
    out_plain = T.transform_grid(plain, theta)
    assert getattr(out_tiled, "_cpab_grid_shape", None) == (100, 37)
    assert torch.equal(out_tiled, out_plain)
    assert torch.equal(T.transform_grid(out_tiled, theta), T.transform_grid(out_plain, theta))

"""The benchmark configurations at their full sizes, checked through
properties that do not need the FP64 oracle to run at that size:
T_{-theta} o T_theta = id (CPA flows are invertible, the inverse is the flow
of the negated field), exact-flow samples, fused == unfused on a slice, the
gradient of a linear functional against a finite difference along a random
direction, and empty / one-point problems."""
import numpy as np
import pytest
import torch

import oracle as O
from libcpab_amd import Cpab

pytestmark = pytest.mark.gpu


def _theta(T, n, scale, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(n, T.get_theta_dim(), generator=g) * scale


@pytest.mark.parametrize("nc,n_theta,size,scale", [([4, 4], 256, [256, 256], 0.3), ([50], 4096, [1024], 0.5),
                                                   ([3, 3, 3], 16, [64, 64, 64], 0.2), ([4, 4], 256, [256, 256], 1.0),
                                                   ([3, 3, 3], 16, [64, 64, 64], 0.6), ([50], 4096, [1024], 1.0)])
def test_inverse_round_trip_at_benchmark_size(cuda, nc, n_theta, size, scale):
    T = Cpab(nc, backend="pytorch", device="gpu", zero_boundary=True)
    theta = _theta(T, n_theta, scale, 0).to(cuda)
    grid = T.uniform_meshgrid(size)
    fwd = T.transform_grid(grid, theta)
    back = T.transform_grid(fwd, -theta)
    err = (back - grid[None]).abs().max().item()
    # FP32 rounding of both flows, amplified by the inverse's Jacobian (the
    # FP64 flow of the FP32-rounded forward returns within 5e-8)
    assert err <= 5e-6, err
    # a sample of the forward against the exact FP64 flow
    p = T.get_params()
    idx = torch.randperm(grid.shape[1], generator=torch.Generator().manual_seed(1))[:512]
    for k in (0, n_theta - 1):
        A = O.expand_A(p.basis, theta[k:k + 1].double().cpu().numpy(), len(nc))
        ref = O.cf_forward(len(nc), p.nc, A, p.facet_planes, p.facet_nbr, grid[:, idx].double().cpu().numpy())
        assert np.abs(fwd[k:k + 1, :, idx].cpu().numpy() - ref).max() <= 1e-5


def _smooth_images(n, ch, size, g):
    """Sums of a few low-frequency sinusoids in [0, 1]: the bilinear
    interpolant of a smooth image has small slope jumps at the pixel edges,
    so a finite difference over a step of a few pixels measures the
    derivative (for noise images it would average over many kinks)."""
    y, x = torch.meshgrid(torch.linspace(0, 1, size[0]), torch.linspace(0, 1, size[1]), indexing="ij")
    fr = torch.rand(n, ch, 3, 2, generator=g) * 3.0 + 0.5
    ph = torch.rand(n, ch, 3, generator=g) * 6.2832
    arg = fr[..., 0, None, None] * x + fr[..., 1, None, None] * y  # [n, ch, 3, H, W]
    img = torch.sin(6.2832 * arg + ph[..., None, None]).mean(2)
    return 0.5 + 0.5 * img


def test_image_workload_gradient_along_a_direction(cuda):
    """d/deps L(theta + eps u) by central differences of the GPU forward vs
    <dL/dtheta, u> from the GPU backward, at the sizes of the image workload
    (256 thetas, 3 x 256 x 256 images, MSE), with smooth images."""
    T = Cpab([4, 4], backend="pytorch", device="gpu", zero_boundary=True)
    g = torch.Generator().manual_seed(3)
    theta = (torch.randn(256, T.get_theta_dim(), generator=g) * 0.3).to(cuda)
    data = _smooth_images(256, 3, [256, 256], g).to(cuda)
    target = _smooth_images(256, 3, [256, 256], g).to(cuda)
    u = torch.randn(theta.shape, generator=g).to(cuda)
    th = theta.clone().requires_grad_(True)
    loss = torch.nn.functional.mse_loss(T.transform_data(data, th, [256, 256]), target)
    loss.backward()
    analytic = (th.grad * u).sum().item()
    eps = 1e-3

    def loss64(th_):  # the loss of a perturbed theta, summed in FP64
        out = T.transform_data(data, th_, [256, 256])
        return ((out.double() - target.double()) ** 2).mean().item()

    with torch.no_grad():
        fd = (loss64(theta + eps * u) - loss64(theta - eps * u)) / (2 * eps)
    assert abs(fd - analytic) <= 1e-2 * abs(analytic) + 1e-7, (fd, analytic)


def test_fused_equals_unfused_on_a_slice_of_the_image_workload(cuda):
    T = Cpab([4, 4], backend="pytorch", device="gpu", zero_boundary=True)
    theta = _theta(T, 8, 0.3, 4).to(cuda)
    data = torch.rand(8, 3, 256, 256, generator=torch.Generator().manual_seed(5)).to(cuda)
    out1 = T.transform_data(data, theta, [256, 256])
    out2 = T.interpolate(data, T.transform_grid(T.uniform_meshgrid([256, 256]), theta), [256, 256])
    assert torch.allclose(out1, out2, atol=1e-6, rtol=0)


def test_empty_and_single_point_problems(cuda):
    T = Cpab([3, 3], backend="pytorch", device="gpu", zero_boundary=True)
    theta = _theta(T, 2, 0.5, 6).to(cuda).requires_grad_(True)
    empty = torch.zeros(2, 0, device=cuda)
    out = T.transform_grid(empty, theta)
    assert out.shape == (2, 2, 0)
    one = torch.tensor([[0.3], [0.7]], device=cuda)
    out = T.transform_grid(one, theta)
    out.sum().backward()
    assert out.shape == (2, 2, 1) and torch.isfinite(theta.grad).all()
    p = T.get_params()
    A = O.expand_A(p.basis, theta.detach().double().cpu().numpy(), 2)
    ref = O.cf_forward(2, p.nc, A, p.facet_planes, p.facet_nbr, one.double().cpu().numpy())
    assert np.abs(out.detach().cpu().numpy() - ref).max() <= 1e-5
    d1 = torch.rand(1, 1, 5, 4, device=cuda)
    assert T.transform_data(d1, theta[:1].detach(), [1, 1]).shape == (1, 1, 1, 1)

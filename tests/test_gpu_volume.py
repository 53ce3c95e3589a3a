"""Volume-preserving fields (Cpab(..., volume_perservation=True)) on the GPU.

Every cell's velocity matrix is trace-free, so inside a cell the flow's
Jacobian has determinant exp(t * trace) = 1, and a facet crossing multiplies
it by the saltation matrix I + (v_next - v_prev) n^T / (n . v_prev), whose
determinant (n . v_next) / (n . v_prev) is 1 because the field is continuous.
So det(dT/dx) = 1 at every point (Liouville).  The Jacobian comes from the
path's own gradient with respect to the points (one backward per output
component), so this checks the forward's trajectory, its carried Jacobian and
the saltation terms together.  Without the constraint the same fields change
volume by far more than the tolerance, which shows the check can fail."""
import pytest
import torch

from libcpab_amd import Cpab

pytestmark = pytest.mark.gpu


def _jacobian_det(T, theta, pts):
    """det(d transform_grid / d points) per (theta, point); pts [n, ndim, nP]."""
    x = pts.clone().requires_grad_(True)
    y = T.transform_grid(x, theta)
    nd = pts.shape[1]
    rows = []
    for a in range(nd):
        (g,) = torch.autograd.grad(y[:, a].sum(), x, retain_graph=a < nd - 1)
        rows.append(g)  # rows[a][n, b, p] = d y_a / d x_b
    J = torch.stack(rows, dim=1).permute(0, 3, 1, 2).double()  # [n, nP, a, b]
    return torch.linalg.det(J.cpu())


@pytest.mark.parametrize("nc,zb,scale", [([3, 3], True, 0.8), ([2, 2], False, 0.6), ([2, 2, 2], True, 0.5)])
def test_volume_preserving_flow_has_unit_jacobian(cuda, nc, zb, scale):
    print("case", nc, zb, scale, flush=True)
    g = torch.Generator().manual_seed(7)
    nd, n, nP = len(nc), 4, 3000
    pts = (torch.rand(n, nd, nP, generator=g) * 0.98 + 0.01).to(cuda)
    Tv = Cpab(nc, backend="pytorch", device="gpu", zero_boundary=zb, volume_perservation=True)
    th = (torch.randn(n, Tv.get_theta_dim(), generator=g) * scale).to(cuda)
    det = _jacobian_det(Tv, th, pts)
    assert (det - 1).abs().max().item() < 1e-4
    # the same check on unconstrained fields of the same size fails clearly
    T = Cpab(nc, backend="pytorch", device="gpu", zero_boundary=zb)
    th0 = (torch.randn(n, T.get_theta_dim(), generator=g) * scale).to(cuda)
    det0 = _jacobian_det(T, th0, pts)
    assert (det0 - 1).abs().max().item() > 1e-2

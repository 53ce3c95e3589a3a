"""The closed-form integrator's chunk budget (512 chunks and crossings per
point) on the GPU: a point that uses it up comes back as NaN -- point,
gradient -- and is counted, and raise_unfinished=True turns that into a
CpabError.  The field is set per cell through an identity basis (theta = A):
every cell of a [2,2] tessellation contracts towards its own centroid,
v = -k (x - c), so a point started near a centroid stays in its cell and
needs 2 sqrt(2) k chunks for t = 1 (|M|_F h <= 1/2)."""
import numpy as np
import pytest
import torch

from libcpab_amd import Cpab
from libcpab_amd._native import CpabError
from libcpab_amd.core.tessellation import make_tessellation

pytestmark = pytest.mark.gpu


def _contracting(k, cuda):
    t = make_tessellation([2, 2], zero_boundary=True)
    D = t.nC * 6
    T = Cpab([2, 2], backend="pytorch", device="gpu", zero_boundary=True, basis=np.eye(D))
    c = t.get_cell_centers()  # [nC, 2] x-space
    A = np.zeros((t.nC, 2, 3))
    A[:, 0, 0] = A[:, 1, 1] = -k
    A[:, :, 2] = k * c
    theta = torch.tensor(A.reshape(1, -1), dtype=torch.float32, device=cuda)
    pts = torch.tensor((c + 0.01).T, dtype=torch.float32, device=cuda)
    return T, theta, pts, c


def test_points_within_the_budget_are_finite_and_not_counted(cuda):
    k = 20.0
    T, theta, pts, c = _contracting(k, cuda)
    out = T.transform_grid(pts, theta)
    want = c.T + (pts.cpu().numpy() - c.T) * np.exp(-k)
    assert np.abs(out[0].cpu().numpy() - want).max() < 1e-6
    assert T.unfinished_points() == 0


def test_budget_exhaustion_is_nan_and_counted(cuda):
    T, theta, pts, c = _contracting(400.0, cuda)
    th = theta.clone().requires_grad_(True)
    out = T.transform_grid(pts, th)
    assert torch.isnan(out).all()
    out.nan_to_num(0.0).sum().backward()  # finite upstream gradient: the NaN comes from the saved state
    assert torch.isnan(th.grad).any()
    assert T.unfinished_points() == pts.shape[1]
    assert T.unfinished_points() == 0  # reset by the read
    # the fused transform_data counts too
    data = torch.rand(1, 1, 8, 8, device=cuda)
    T.transform_data(data, theta, [8, 8])
    assert T.unfinished_points() > 0
    T.set_solver_params(raise_unfinished=True)
    with pytest.raises(CpabError):
        T.transform_grid(pts, theta)


@pytest.mark.parametrize("name", ["2d_3x3", "2d_4x4", "3d_2x2x2", "3d_3x3x3"])
def test_domain_corners_stay_put_within_the_budget(cuda, name):
    """Under zero_boundary every corner of the domain is a fixed point; in
    FP32 its velocity is rounding noise, which must not carry the point back
    and forth across the facets through the corner until the budget runs out
    (it did, for one theta met while aligning with Adam)."""
    from conftest import make_cpab
    T, b = make_cpab(name)
    nd = len(b["nc"])
    g = torch.Generator().manual_seed(4)
    theta = (torch.randn(1024, T.get_theta_dim(), generator=g) * 0.8).to(cuda)
    corners = torch.tensor(np.array(np.meshgrid(*[[0.0, 1.0]] * nd, indexing="ij")).reshape(nd, -1),
                           dtype=torch.float32, device=cuda)
    T.unfinished_points()
    out = T.transform_grid(corners, theta)
    assert T.unfinished_points() == 0
    assert (out - corners[None]).abs().max().item() <= 1e-6

"""use_slow=True (the reference's pure-torch stepping path,
libcpab/pytorch/transformer.py:90-128) on the GPU against its FP64
restatement (oracle/slow_path.py): the forward, dL/dpoints and, through torch
autograd of matrix_exp and B theta (FP64, CPU), dL/dtheta.

The stepping map jumps where a step's cell changes (expm(A1/n) x and
expm(A2/n) x differ to second order in 1/n), so points whose oracle cell
sequence changes when the start moves by 1e-6 along an axis are left out of
the comparison (and of the functional whose gradient is compared); they are a
small minority."""
import numpy as np
import pytest
import torch

import oracle as O
from conftest import make_cpab
from libcpab_amd.pytorch.slow import trels_of

pytestmark = pytest.mark.gpu

NSTEP = 20


def _stable(S, pts, tr, nc):
    _, _, cs0 = S.slow_forward(pts, tr, nc, NSTEP, keep=True)
    ok = np.ones(cs0[0].shape, bool)
    for a in range(pts.shape[0]):
        for sgn in (1.0, -1.0):
            p = pts.copy()
            p[a] += sgn * 1e-6
            _, _, cs = S.slow_forward(p, tr, nc, NSTEP, keep=True)
            for x, y in zip(cs, cs0):
                ok &= x == y
    return ok


@pytest.mark.parametrize("name,scale", [("1d_10", 0.8), ("2d_3x3", 0.5), ("2d_2x2_nzb", 0.5), ("3d_2x2x2", 0.3),
                                        ("3d_2x2x2_nzb", 0.3)])
def test_use_slow_matches_its_fp64_restatement(cuda, name, scale):
    from oracle import slow_path as S
    T, b = make_cpab(name)
    T.set_solver_params(nstepsolver=NSTEP, use_slow=True)
    p = T.get_params()
    nd, nc = p.ndim, p.nc
    g = torch.Generator().manual_seed(41)
    theta = torch.randn(2, T.get_theta_dim(), generator=g) * scale
    lo, hi = (0.0, 1.0) if p.zero_boundary else (-0.1, 1.1)
    pts = (torch.rand(nd, 600, generator=g) * (hi - lo) + lo).float()
    th = theta.to(cuda).requires_grad_(True)
    pp = pts.to(cuda).requires_grad_(True)
    out = T.transform_grid(pp, th)
    # the oracle with the same FP32 Trels the GPU stepped with
    tr32 = trels_of(theta.to(cuda), p, torch.tensor(b["B"], dtype=torch.float32, device=cuda))
    tr = tr32.detach().double().cpu().numpy()
    p64 = pts.double().numpy()
    ref = S.slow_forward(p64, tr, nc, NSTEP)
    ok = _stable(S, p64, tr, nc).reshape(2, -1)
    print(name, "stable points", ok.mean())
    assert ok.mean() > 0.9
    err = np.abs(out.detach().cpu().numpy() - ref).max(axis=1)[ok]
    assert err.max() <= 1e-5, err.max()

    # gradients of <w, out> with w zero on the unstable points
    w = torch.randn(2, nd, 600, generator=g) * torch.tensor(ok[:, None, :], dtype=torch.float32)
    (out * w.to(cuda)).sum().backward()
    gT, gp = S.slow_backward(p64, tr, nc, NSTEP, w.double().numpy())
    rp = np.abs(pp.grad.cpu().numpy() - gp.sum(0)).max() / np.abs(gp.sum(0)).max()
    # dL/dtheta: dL/dTrels through matrix_exp and B theta by torch autograd in FP64
    th64 = theta.double().requires_grad_(True)
    tr64 = trels_of(th64, p, torch.tensor(b["B"], dtype=torch.float64))
    tr64.backward(torch.tensor(gT))
    gth = th64.grad.numpy()
    rt = np.abs(th.grad.cpu().double().numpy() - gth).max() / np.abs(gth).max()
    print(name, "dL/dpoints rel err", rp, "dL/dtheta rel err", rt)
    assert rp <= 1e-3
    assert rt <= 1e-3


def test_use_slow_dispatch_transform_data_and_sequential(cuda):
    """use_slow reaches transform_data (grid, then interpolation) and every
    stage of CpabSequential, whose thetas all get gradients (the slow path
    differentiates w.r.t. the points)."""
    from libcpab_amd import CpabSequential
    (T1, _), (T2, _) = make_cpab("2d_3x3"), make_cpab("2d_4x4")
    for T in (T1, T2):
        T.set_solver_params(nstepsolver=10, use_slow=True)
    g = torch.Generator().manual_seed(42)
    data = torch.rand(2, 3, 24, 20, generator=g).to(cuda)
    th = [(torch.randn(2, T.get_theta_dim(), generator=g) * 0.3).to(cuda).requires_grad_(True) for T in (T1, T2)]
    out = T1.transform_data(data, th[0], [24, 20])
    grid = T1.uniform_meshgrid([24, 20])
    ref = T1.interpolate(data, T1.transform_grid(grid, th[0]), [24, 20])
    assert torch.equal(out, ref)
    S = CpabSequential(T1, T2)
    S.transform_data(data, th, [24, 20]).square().sum().backward()
    assert all(t.grad is not None and t.grad.abs().max() > 0 for t in th)

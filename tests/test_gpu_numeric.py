"""method='numeric' (the reference's fixed-step scheme) on the GPU against the
golden vectors of the reference's own compiled CPU op, and against the exact
flow as nstepsolver grows."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from conftest import load_npz, make_cpab
from libcpab_amd import _native as N
from libcpab_amd.pytorch import hip_ops as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["1d_10", "2d_3x3", "3d_2x2x2"])
def test_ref_forward_backward_bitexact_with_reference_golden(cuda, name):
    f = load_npz("ref_step_%s.npz" % name)
    nc = [int(v) for v in f["nc"]]
    ndim = len(nc)
    pts = torch.tensor(f["points"], device=cuda)
    trels = torch.tensor(f["trels"], device=cuda)
    n_theta, nC = trels.shape[:2]
    nP = pts.shape[-1]
    out = torch.empty(n_theta, ndim, nP, device=cuda)
    st = H._stream(cuda)
    N.check(N.lib().cpab_ref_forward(H._ptr(out), H._ptr(pts), H._ptr(trels), int(f["nstep"]), H._int_arr(nc), ndim,
                                     nP, n_theta, 0, st))
    As = torch.tensor(f["As"], device=cuda)
    Bs = torch.tensor(f["Bs"], device=cuda)
    d = Bs.shape[0]
    J = torch.empty(d, n_theta, ndim, nP, device=cuda)
    N.check(N.lib().cpab_ref_backward(H._ptr(J), H._ptr(pts), H._ptr(As), H._ptr(Bs), int(f["nstep"]),
                                      H._int_arr(nc), n_theta, d, ndim, nP, nC, 0, st))
    torch.cuda.synchronize()
    fo = out.cpu().numpy()
    jb = J.cpu().numpy()
    print(name, "forward mismatches", int((fo != f["forward"]).sum()), "max", np.abs(fo - f["forward"]).max(),
          "backward max", np.abs(jb - f["backward"]).max())
    assert np.array_equal(fo, f["forward"])
    assert np.array_equal(jb, f["backward"])


@pytest.mark.parametrize("name", ["1d_10", "2d_3x3", "3d_2x2x2"])
def test_expm_matches_scipy(cuda, name):
    f = load_npz("ref_step_%s.npz" % name)
    ndim = len(f["nc"])
    A = torch.tensor(f["As"], device=cuda).reshape(-1, ndim, ndim + 1)
    T = torch.empty_like(A)
    N.check(N.lib().cpab_expm_affine(H._ptr(A), A.shape[0], ndim, 1.0 / 50, H._ptr(T), H._stream(cuda)))
    torch.cuda.synchronize()
    ref = f["trels"].reshape(-1, ndim, ndim + 1)
    assert np.abs(T.cpu().numpy() - ref).max() <= 2e-7


@pytest.mark.parametrize("name,scale", [("2d_3x3", 0.5), ("1d_50", 0.5)])
def test_numeric_method_through_cpab(cuda, name, scale):
    T, b = make_cpab(name)
    ndim = len(b["nc"])
    g = torch.Generator().manual_seed(3)
    theta = (torch.randn(2, T.get_theta_dim(), generator=g) * scale).to(cuda)
    pts = torch.rand(ndim, 300, generator=g).to(cuda)
    A = O.expand_A(b["B"], theta.double().cpu().numpy(), ndim)
    exact = O.cf_forward(ndim, list(b["nc"]), A, b["planes"], b["nbr"], pts.double().cpu().numpy())
    errs = []
    for nstep in (50, 500):
        T.set_solver_params(nstepsolver=nstep, method="numeric")
        out = T.transform_grid(pts, theta).cpu().numpy()
        errs.append(np.abs(out - exact).max())
    # FP32 stepping: the truncation error at 50 steps is already below the
    # rounding that accumulates over 500, so both sit near the exact flow
    assert max(errs) < 2e-4, errs
    # analytic (RK2) gradient vs the exact gradient: O(h) close, and vs finite differences of itself
    T.set_solver_params(nstepsolver=200, method="numeric")
    w = torch.randn(2, ndim, 300, generator=g).to(cuda)
    th = theta.clone().requires_grad_(True)
    (T.transform_grid(pts, th) * w).sum().backward()
    gA, _ = O.cf_backward(ndim, list(b["nc"]), A, b["planes"], b["nbr"], pts.double().cpu().numpy(),
                          w.double().cpu().numpy())
    gex = gA.reshape(2, -1) @ b["B"]
    rel = np.abs(th.grad.cpu().numpy() - gex).max() / np.abs(gex).max()
    assert rel < 2e-2
    T.set_solver_params(nstepsolver=200, numeric_grad=True, method="numeric")
    th2 = theta.clone().requires_grad_(True)
    (T.transform_grid(pts, th2) * w).sum().backward()
    rel2 = np.abs(th2.grad.cpu().numpy() - gex).max() / np.abs(gex).max()
    # forward differences with the reference's step h = 0.01: O(h) error
    assert rel2 < 0.15

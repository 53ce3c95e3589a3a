"""The workloads the path is measured on, at their real sizes, against the
FP64 oracle (run over point slices on a thread pool):

* plain [3,3] (zero boundary): one theta ~ N(0, 0.5^2 I), seed 0, the whole
  100 x 100 meshgrid and a 1 x 3 x 100 x 100 U(0,1) image -- warped grid,
  interpolation, dL/dtheta and dL/ddata against the oracle, dL/dtheta also
  against FP64 central differences of the oracle's own loss;
* the ragged mix of 1000 instances (1D [10..100] cells / 2D [2,2]..[8,8],
  64^2..512^2 pixels) through the graph-captured persistent launch --
  sampled instances (the largest 1D and 2D ones, every [8,8] one, a few
  drawn at random) against the oracle directly;
* CpabSequential of 4 x [4,4] on 64 x 3 x 256 x 256 images -- inverse round
  trip, sampled points of the composed flow against the oracle chained
  stage by stage, and the directional derivative of the loss;
* the time-series ([50], 4096 thetas x 1024 samples) and volume ([3,3,3],
  16 x 64^3) workloads: the directional derivative of the loss by FP64
  central differences of the GPU forward against the GPU gradient.
"""
import numpy as np
import pytest
import torch

import oracle as O
from conftest import make_cpab, par_cf_backward, par_cf_forward
from libcpab_amd import Cpab, CpabSequential

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


# ---------------------------------------------------------------------------
# plain [3,3], 100 x 100
# ---------------------------------------------------------------------------
def test_plain_3x3_on_the_100x100_meshgrid(cuda):
    T, b = make_cpab("2d_3x3")
    nc = [3, 3]
    g = torch.Generator().manual_seed(0)
    theta = torch.randn(1, T.get_theta_dim(), generator=g) * 0.5
    data = torch.rand(1, 3, 100, 100, generator=g)
    target = torch.rand(1, 3, 100, 100, generator=g)
    grid = T.uniform_meshgrid([100, 100])
    A = O.expand_A(b["B"], theta.double().numpy(), 2)
    g64 = grid.double().cpu().numpy()

    # warped grid: every one of the 10,000 points
    with torch.no_grad():
        warped = T.transform_grid(grid, theta.to(cuda)).cpu().numpy()
    exact = par_cf_forward(2, nc, A, b["planes"], b["nbr"], g64)
    err_pos = float(np.abs(warped - exact).max())
    print("max |grid - exact| =", err_pos)
    assert err_pos <= 1e-5

    # fused transform_data and its gradients, MSE to a U(0,1) target
    th = theta.to(cuda).requires_grad_(True)
    dd = data.to(cuda).requires_grad_(True)
    out = T.transform_data(dd, th, [100, 100])
    loss = torch.nn.functional.mse_loss(out, target.to(cuda))
    loss.backward()
    d64 = data.double().numpy()
    ref = O.interpolate(2, d64, exact, [100, 100])
    # the interpolant's slope is at most (size - 1) x |data step| <= 99 per axis
    tol = 2 * 99 * err_pos + 1e-6
    err_img = float(np.abs(out.detach().cpu().numpy() - ref).max())
    print("max |image - exact| =", err_img, "tol", tol)
    assert err_img <= tol
    # the interpolation alone at the GPU's own grid
    at_gpu = O.interpolate(2, d64, warped.astype(np.float64), [100, 100])
    assert np.abs(out.detach().cpu().numpy() - at_gpu).max() <= 1e-5

    gout = 2.0 * (ref - target.double().numpy()) / ref.size
    gdata, ggrid = O.interpolate_backward(2, d64, exact, [100, 100], gout)
    gA, _ = par_cf_backward(2, nc, A, b["planes"], b["nbr"], g64, ggrid)
    gtheta = gA.reshape(1, -1) @ b["B"]
    r_th = _rel(th.grad.cpu().double().numpy(), gtheta)
    r_da = _rel(dd.grad.cpu().double().numpy(), gdata)
    print("dL/dtheta rel err", r_th, "dL/ddata rel err", r_da)
    assert r_th <= 1e-3
    assert r_da <= 1e-3

    # dL/dtheta against FP64 central differences of the oracle's loss
    u = torch.randn(theta.shape, generator=g).double().numpy()
    tg = target.double().numpy()

    def loss64(t):
        At = O.expand_A(b["B"], t, 2)
        img = O.interpolate(2, d64, par_cf_forward(2, nc, At, b["planes"], b["nbr"], g64), [100, 100])
        return float(((img - tg) ** 2).mean())

    # a small step: the bilinear interpolant of a noise image has slope jumps
    # at pixel edges, and a larger step lets points cross them (at 1e-6 the
    # difference quotient is off by 1e-3, at 1e-7 by 1e-8 of the exact value)
    eps = 1e-7
    t0 = theta.double().numpy()
    fd = (loss64(t0 + eps * u) - loss64(t0 - eps * u)) / (2 * eps)
    an = float((th.grad.cpu().double().numpy() * u).sum())
    print("directional derivative: FD", fd, "GPU", an)
    assert abs(fd - an) <= 1e-3 * abs(fd)


# ---------------------------------------------------------------------------
# the ragged mix, 1000 instances, graph-captured persistent launch
# ---------------------------------------------------------------------------
def _ragged_batch(cuda):
    import bench
    from libcpab_amd.batch import CpabBatch
    probs = bench.ragged_problems(1000)
    b = CpabBatch(probs, device=cuda, save=True)
    g = torch.Generator().manual_seed(0)
    for i, (T, C, size, out) in enumerate(probs):
        b.theta_views[i].copy_(torch.randn(T.get_theta_dim(), generator=g) * 0.3)
        b.data_views[i].copy_(torch.rand([C] + size, generator=g))
        b.grad_out_views[i].copy_(torch.randn([C] + out, generator=g))
    return b, probs


def _sampled_instances(probs):
    nd = [T.params.ndim for T, _, _, _ in probs]
    npts = [int(np.prod(o)) for _, _, _, o in probs]
    one = [i for i in range(len(probs)) if nd[i] == 1]
    two = [i for i in range(len(probs)) if nd[i] == 2]
    pick = {max(one, key=lambda i: npts[i]), max(two, key=lambda i: npts[i])}
    pick |= {i for i in two if probs[i][0].params.nc == [8, 8]}
    rng = np.random.default_rng(7)
    pick |= set(int(i) for i in rng.choice(len(probs), 6, replace=False))
    return sorted(pick)


def test_ragged_mix_against_the_oracle(cuda):
    b, probs = _ragged_batch(cuda)
    b.set_launch(2, 1024)
    b.capture()
    b.replay_forward()
    b.replay_backward()
    torch.cuda.synchronize()
    picked = _sampled_instances(probs)
    assert any(probs[i][0].params.nc == [8, 8] for i in picked)
    worst = {"fwd": 0.0, "theta": 0.0, "data": 0.0}
    for i in picked:
        T, C, size, osz = probs[i]
        p = T.get_params()
        ndim = p.ndim
        theta = b.theta_views[i].detach().double().cpu().numpy()[None]
        A = O.expand_A(p.basis, theta, ndim)
        grid = O.uniform_meshgrid(osz)
        data = b.data_views[i].detach().double().cpu().numpy()[None]
        gout = b.grad_out_views[i].detach().double().cpu().numpy()[None]
        exact = par_cf_forward(ndim, p.nc, A, p.facet_planes, p.facet_nbr, grid)
        # the forward's saved warped points (output order: last axis fastest)
        warped = b.saved_views[i][:ndim].detach().double().cpu().numpy()
        exact_t = exact[0].reshape((ndim,) + tuple(osz[::-1])).transpose((0,) + tuple(range(ndim, 0, -1)))
        e = float(np.abs(warped - exact_t.reshape(ndim, -1)).max())
        worst["fwd"] = max(worst["fwd"], e)
        assert e <= 1e-5, (i, p.nc, osz, e)
        # the interpolated output at those points against the FP64 restatement,
        # sample coordinates x * (size - 1) formed in FP32 as the reference's
        # float32 interpolation does (at 512 samples an FP32 product is off by
        # up to 3e-5 pixel, which moves a point across a pixel edge)
        got = b.out_views[i].detach().cpu().numpy()
        w_pts = warped.reshape((ndim,) + tuple(osz)).transpose((0,) + tuple(range(ndim, 0, -1)))
        w_pts = w_pts.reshape(1, ndim, -1)
        ref = O.interpolate(ndim, data, w_pts, osz, index32=True)[0]
        assert np.abs(got - ref).max() <= 1e-5, (i, p.nc, osz)
        # dL/dgrid where the GPU's points are, with the same FP32 sample
        # coordinates (the bilinear interpolant of a noise image changes slope
        # at every pixel edge: the side of an edge a point lies on must be the
        # GPU's), then the oracle's exact adjoint of the integration
        gdata, ggrid = O.interpolate_backward(ndim, data, w_pts, osz, gout, index32=True)
        gA, _ = par_cf_backward(ndim, p.nc, A, p.facet_planes, p.facet_nbr, grid, ggrid)
        gtheta = (gA.reshape(1, -1) @ p.basis)[0]
        rt = _rel(b.grad_theta_views[i].detach().double().cpu().numpy(), gtheta)
        rd = _rel(b.grad_data_views[i].detach().double().cpu().numpy(), gdata[0])
        worst["theta"] = max(worst["theta"], rt)
        worst["data"] = max(worst["data"], rd)
        assert rt <= 1e-3, (i, p.nc, osz, rt)
        assert rd <= 1e-3, (i, p.nc, osz, rd)
    print("instances", picked, "worst", worst)


# ---------------------------------------------------------------------------
# CpabSequential, 4 x [4,4] on 64 x 3 x 256 x 256
# ---------------------------------------------------------------------------
def _smooth(n, ch, size, g):
    """Sums of low-frequency sinusoids in [0, 1] (a finite difference over a
    few pixels then measures the derivative, not an average over kinks)."""
    axes = torch.meshgrid(*[torch.linspace(0, 1, s) for s in size], indexing="ij")
    nd = len(size)
    fr = torch.rand(n, ch, 3, nd, generator=g) * 3.0 + 0.5
    ph = torch.rand(n, ch, 3, generator=g) * 6.2832
    arg = sum(fr[..., a].reshape(n, ch, 3, *([1] * nd)) * axes[a] for a in range(nd))
    return 0.5 + 0.5 * torch.sin(6.2832 * arg + ph.reshape(n, ch, 3, *([1] * nd))).mean(2)


def test_sequential_4_stages_on_the_image_batch(cuda):
    Ts = [Cpab([4, 4], backend="pytorch", device="gpu", zero_boundary=True) for _ in range(4)]
    S = CpabSequential(*Ts)
    g = torch.Generator().manual_seed(0)
    thetas = [(torch.randn(64, T.get_theta_dim(), generator=g) * 0.3).to(cuda) for T in Ts]
    grid = S.uniform_meshgrid([256, 256])

    with torch.no_grad():
        fwd = S.transform_grid(grid, thetas)
        # inverse round trip: the inverse of the chain is the chain of the
        # negated fields in reverse order
        inv = CpabSequential(*Ts[::-1])
        back = inv.transform_grid(fwd, [-t for t in thetas[::-1]])
    err = (back - grid[None]).abs().max().item()
    print("4-stage inverse round trip", err)
    assert err <= 2e-5

    # sampled points of the composed flow against the oracle, stage by stage
    idx = torch.randperm(grid.shape[1], generator=torch.Generator().manual_seed(1))[:1024]
    for k in (0, 63):
        pts = grid[:, idx].double().cpu().numpy()
        for T, th in zip(Ts, thetas):
            p = T.get_params()
            A = O.expand_A(p.basis, th[k:k + 1].double().cpu().numpy(), 2)
            pts = O.cf_forward(2, p.nc, A, p.facet_planes, p.facet_nbr, pts)[0]
        e = np.abs(fwd[k, :, idx].cpu().numpy() - pts).max()
        print("theta", k, "composed flow vs oracle", e)
        assert e <= 4e-6

    # directional derivative of the loss w.r.t. all four thetas
    data = _smooth(64, 3, [256, 256], g).to(cuda)
    target = _smooth(64, 3, [256, 256], g).to(cuda)
    us = [torch.randn(t.shape, generator=g).to(cuda) for t in thetas]
    ths = [t.clone().requires_grad_(True) for t in thetas]
    loss = torch.nn.functional.mse_loss(S.transform_data(data, ths, [256, 256]), target)
    loss.backward()
    analytic = sum((t.grad * u).sum().item() for t, u in zip(ths, us))
    assert all(torch.isfinite(t.grad).all() and t.grad.abs().max() > 0 for t in ths)

    def loss64(ts):
        out = S.transform_data(data, ts, [256, 256])
        return ((out.double() - target.double()) ** 2).mean().item()

    eps = 1e-3
    with torch.no_grad():
        fd = (loss64([t + eps * u for t, u in zip(thetas, us)]) -
              loss64([t - eps * u for t, u in zip(thetas, us)])) / (2 * eps)
    print("sequential directional derivative: FD", fd, "GPU", analytic)
    assert abs(fd - analytic) <= 1e-2 * abs(analytic)


# ---------------------------------------------------------------------------
# time series and volumes: directional derivatives at full size
# ---------------------------------------------------------------------------
def _blobs(n, size, g, nblob=8):
    """Sums of 8 random Gaussian blobs, [n, 1, *size]."""
    axes = torch.meshgrid(*[torch.linspace(0, 1, s) for s in size], indexing="ij")
    out = torch.zeros((n, 1) + tuple(size))
    for _ in range(nblob):
        c = torch.rand(n, len(size), generator=g)
        s = torch.rand(n, generator=g) * 0.15 + 0.1
        r2 = sum((axes[a][None] - c[:, a].reshape(n, *([1] * len(size)))) ** 2 for a in range(len(size)))
        out[:, 0] += torch.exp(-r2 / (2 * s.reshape(n, *([1] * len(size))) ** 2))
    return out


@pytest.mark.parametrize("workload", ["ts1d", "vol3d"])
def test_directional_derivative_at_full_size(cuda, workload):
    g = torch.Generator().manual_seed(3)
    if workload == "ts1d":
        T = Cpab([50], backend="pytorch", device="gpu", zero_boundary=True)
        n, size, scale = 4096, [1024], 0.5
        # smooth signals: sums of three random sinusoids (no noise, so the
        # finite difference sees the derivative of the interpolant)
        data = _smooth(n, 1, size, g)
        target = _smooth(n, 1, size, g)
    else:
        T = Cpab([3, 3, 3], backend="pytorch", device="gpu", zero_boundary=True)
        n, size, scale = 16, [64, 64, 64], 0.2
        data = _blobs(n, size, g)
        target = _blobs(n, size, g)
    theta = (torch.randn(n, T.get_theta_dim(), generator=g) * scale).to(cuda)
    data, target = data.to(cuda), target.to(cuda)
    u = torch.randn(theta.shape, generator=g).to(cuda)
    th = theta.clone().requires_grad_(True)
    loss = torch.nn.functional.mse_loss(T.transform_data(data, th, size), target)
    loss.backward()
    analytic = (th.grad * u).sum().item()

    def loss64(t):
        out = T.transform_data(data, t, size)
        return ((out.double() - target.double()) ** 2).mean().item()

    eps = 1e-3
    with torch.no_grad():
        fd = (loss64(theta + eps * u) - loss64(theta - eps * u)) / (2 * eps)
    print(workload, "directional derivative: FD", fd, "GPU", analytic)
    assert abs(fd - analytic) <= 1e-2 * abs(analytic)

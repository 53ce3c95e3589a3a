"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Bars: warped points within 1e-5 of the exact FP64 flow; gradients within
1e-3 relative (of the gradient's max magnitude) of the exact FP64 gradient;
interpolation within 1e-5 of the FP64 restatement at the same grid.  Without
zero boundary the points start inside and outside the domain (3D: in the
regions outside it, each with the map of the tetrahedron the reference's
lookup gives the pushed point).
"""
import numpy as np
import pytest
import torch

import oracle as O
from conftest import make_cpab

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5
GRAD_RTOL = 1e-3

CASES = [  # tessellation, n_theta, theta scale
    ("1d_10", 3, 0.8),
    ("1d_50", 3, 0.5),
    ("2d_3x3", 3, 0.5),
    ("2d_4x4", 3, 0.3),
    ("2d_2x2_nzb", 3, 0.5),
    ("3d_2x2x2", 2, 0.3),
    ("3d_3x3x3", 2, 0.2),
    ("3d_2x2x2_nzb", 2, 0.5),
]


def _theta(d, n, scale, seed):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(n, d, generator=g) * scale).float()


def _points(ndim, nP, seed, lo=0.0, hi=1.0):
    g = torch.Generator().manual_seed(seed + 17)
    return (torch.rand(ndim, nP, generator=g) * (hi - lo) + lo).float()


def _range(b):
    return (0.0, 1.0) if bool(b["zero_boundary"]) else (-0.2, 1.2)


def _cf_forward(b, A, pts):
    return O.cf_forward(len(b["nc"]), [int(v) for v in b["nc"]], A, b["planes"], b["nbr"], pts, amap=b.get("amap"))


def _cf_backward(b, A, pts, w):
    return O.cf_backward(len(b["nc"]), [int(v) for v in b["nc"]], A, b["planes"], b["nbr"], pts, w,
                         amap=b.get("amap"))


def _A64(B, theta, ndim):
    return O.expand_A(B, theta.double().numpy(), ndim)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-12)


@pytest.mark.parametrize("name,n_theta,scale", CASES)
def test_transform_grid_matches_exact(cuda, name, n_theta, scale):
    T, b = make_cpab(name)
    ndim = len(b["nc"])
    theta = _theta(T.get_theta_dim(), n_theta, scale, 1)
    pts = _points(ndim, 2000, 2, *_range(b))
    # include a meshgrid: its points sit on facets and on the domain boundary
    side = {1: 101, 2: 23, 3: 9}[ndim]
    mesh = T.uniform_meshgrid([side] * ndim).cpu()
    grid = torch.cat([pts, mesh], dim=1)
    out = T.transform_grid(grid.to(cuda), theta.to(cuda)).cpu().numpy()
    ref = _cf_forward(b, _A64(b["B"], theta, ndim), grid.double().numpy())
    err = np.abs(out - ref).max()
    print(name, "max |gpu - exact| =", err)
    assert err <= POS_TOL


@pytest.mark.parametrize("name,n_theta,scale", CASES)
def test_transform_grid_per_theta_grid(cuda, name, n_theta, scale):
    T, b = make_cpab(name)
    ndim = len(b["nc"])
    theta = _theta(T.get_theta_dim(), n_theta, scale, 3)
    grid = torch.stack([_points(ndim, 500, 10 + i, *_range(b)) for i in range(n_theta)])
    out = T.transform_grid(grid.to(cuda), theta.to(cuda)).cpu().numpy()
    ref = _cf_forward(b, _A64(b["B"], theta, ndim), grid.double().numpy())
    assert np.abs(out - ref).max() <= POS_TOL


@pytest.mark.parametrize("name,n_theta,scale", CASES)
@pytest.mark.parametrize("store", [True, False])
def test_transform_grid_gradients(cuda, name, n_theta, scale, store):
    T, b = make_cpab(name)
    T.set_solver_params(store_jacobian=store)
    ndim = len(b["nc"])
    theta = _theta(T.get_theta_dim(), n_theta, scale, 4)
    pts = _points(ndim, 400, 5, *_range(b))
    g = torch.Generator().manual_seed(6)
    w = torch.randn(n_theta, ndim, pts.shape[1], generator=g)
    th = theta.to(cuda).requires_grad_(True)
    pp = pts.to(cuda).requires_grad_(True)
    out = T.transform_grid(pp, th)
    (out * w.to(cuda)).sum().backward()
    A = _A64(b["B"], theta, ndim)
    gA, gp = _cf_backward(b, A, pts.double().numpy(), w.double().numpy())
    gtheta = gA.reshape(n_theta, -1) @ b["B"]
    rt = _rel(th.grad.cpu().double().numpy(), gtheta)
    rp = _rel(pp.grad.cpu().double().numpy(), gp.sum(0))
    print(name, "grad theta rel err", rt, "grad points rel err", rp)
    assert rt <= GRAD_RTOL
    assert rp <= GRAD_RTOL


@pytest.mark.parametrize("ndim,size,outsize", [(1, [40], [57]), (2, [13, 17], [11, 19]), (3, [7, 6, 5], [5, 8, 6])])
def test_interpolate_matches_restatement(cuda, ndim, size, outsize):
    g = torch.Generator().manual_seed(7)
    N, C = 3, 2
    data = torch.rand((N, C) + tuple(size), generator=g)
    nP = int(np.prod(outsize))
    grid = torch.rand(N, ndim, nP, generator=g) * 1.2 - 0.1  # includes clamped corners
    gout = torch.randn((N, C) + tuple(outsize), generator=g)
    dd = data.to(cuda).requires_grad_(True)
    gg = grid.to(cuda).requires_grad_(True)
    from libcpab_amd.pytorch.functions import interpolate
    out = interpolate(ndim, dd, gg, outsize)
    (out * gout.to(cuda)).sum().backward()
    ref = O.interpolate(ndim, data.double().numpy(), grid.double().numpy(), outsize)
    gd, ggr = O.interpolate_backward(ndim, data.double().numpy(), grid.double().numpy(), outsize, gout.double().numpy())
    assert np.abs(out.detach().cpu().numpy() - ref).max() <= 1e-5
    assert np.abs(dd.grad.cpu().numpy() - gd).max() <= 1e-4
    assert _rel(gg.grad.cpu().double().numpy(), ggr) <= 1e-4


@pytest.mark.parametrize("name,n_theta,scale,size", [("1d_10", 3, 0.8, [64]), ("2d_3x3", 3, 0.5, [24, 20]),
                                                      ("3d_2x2x2", 2, 0.3, [9, 8, 7]),
                                                      ("3d_2x2x2", 2, 0.3, [12, 9, 10]),  # 3D tiles
                                                      ("3d_2x2x2_nzb", 2, 0.5, [9, 8, 7])])
@pytest.mark.parametrize("store", [True, False])
def test_transform_data_fused_matches_unfused(cuda, name, n_theta, scale, size, store):
    """Fused meshgrid->integrate->interpolate equals transform_grid + interpolate,
    forward and backward, and the warped grid equals the exact flow."""
    T, b = make_cpab(name)
    T.set_solver_params(store_jacobian=store)
    ndim = len(b["nc"])
    theta = _theta(T.get_theta_dim(), n_theta, scale, 8)
    g = torch.Generator().manual_seed(9)
    data = torch.rand((n_theta, 2) + tuple(size), generator=g)
    gout = torch.randn((n_theta, 2) + tuple(size), generator=g)
    th1 = theta.to(cuda).requires_grad_(True)
    d1 = data.to(cuda).requires_grad_(True)
    out1 = T.transform_data(d1, th1, size)
    (out1 * gout.to(cuda)).sum().backward()
    th2 = theta.to(cuda).requires_grad_(True)
    d2 = data.to(cuda).requires_grad_(True)
    grid = T.uniform_meshgrid(size)
    gt = T.transform_grid(grid, th2)
    out2 = T.interpolate(d2, gt, size)
    (out2 * gout.to(cuda)).sum().backward()
    assert torch.allclose(out1, out2, atol=1e-6, rtol=0)
    assert torch.allclose(d1.grad, d2.grad, atol=1e-5, rtol=0)
    assert _rel(th1.grad.cpu().double().numpy(), th2.grad.cpu().double().numpy()) <= 1e-4
    # against the exact flow + FP64 interpolation at the exact grid
    ref_grid = _cf_forward(b, _A64(b["B"], theta, ndim), grid.cpu().double().numpy())
    assert np.abs(gt.detach().cpu().numpy() - ref_grid).max() <= POS_TOL
    ref = O.interpolate(ndim, data.double().numpy(), ref_grid, size)
    # interpolation amplifies grid error by at most |grad data| ~ (size-1)
    assert np.abs(out1.detach().cpu().numpy() - ref).max() <= POS_TOL * max(size) * 2 + 1e-6


@pytest.mark.parametrize("nc", [[3, 2, 2], [1, 2, 3]])
def test_free_boundary_3d_far_outside_and_uneven_tessellations(cuda, nc):
    """3D without zero boundary on tessellations of unequal sides, points up
    to 4 domain lengths outside (the outside regions are unbounded): warped
    points and both gradients against the FP64 oracle with the same
    outside regions."""
    from libcpab_amd import Cpab
    T = Cpab(nc, backend="pytorch", device="gpu", zero_boundary=False)
    p = T.get_params()
    b = {"nc": np.array(nc), "planes": p.facet_planes, "nbr": p.facet_nbr, "amap": p.facet_amap,
         "zero_boundary": False, "B": p.basis}
    theta = _theta(T.get_theta_dim(), 2, 0.3, 41)
    g = torch.Generator().manual_seed(42)
    pts = torch.cat([torch.rand(3, 600, generator=g) * 1.4 - 0.2, torch.rand(3, 200, generator=g) * 9.0 - 4.0], 1)
    th = theta.to(cuda).requires_grad_(True)
    pp = pts.to(cuda).requires_grad_(True)
    out = T.transform_grid(pp, th)
    A = _A64(b["B"], theta, 3)
    ref = _cf_forward(b, A, pts.double().numpy())
    assert np.abs(out.detach().cpu().numpy() - ref).max() <= POS_TOL * max(1.0, np.abs(ref).max())
    w = torch.randn(2, 3, pts.shape[1], generator=g)
    (out * w.to(cuda)).sum().backward()
    gA, gp = _cf_backward(b, A, pts.double().numpy(), w.double().numpy())
    assert _rel(th.grad.cpu().double().numpy(), gA.reshape(2, -1) @ b["B"]) <= GRAD_RTOL
    assert _rel(pp.grad.cpu().double().numpy(), gp.sum(0)) <= GRAD_RTOL


@pytest.mark.parametrize("dsize,outsize", [([512, 512], [64, 64]), ([40, 50], [100, 90]), ([33, 7], [50, 5]),
                                           ([64, 64], [64, 64]), ([96, 96], [64, 64])])
def test_transform_data_2d_tiles_and_lds_box(cuda, dsize, outsize):
    """The 2D fused backward accumulates dL/ddata of each output tile in an LDS
    box when its footprint fits, with the footprint's data staged beside it
    when that fits too (96 -> 64: the bins fit, the data does not; 512 -> 64
    overflows the bins: global atomics; a second axis below 8 runs untiled).
    Gradients equal the unfused path's."""
    T, b = make_cpab("2d_4x4")
    theta = _theta(T.get_theta_dim(), 3, 0.3, 21)
    g = torch.Generator().manual_seed(22)
    data = torch.rand([3, 3] + dsize, generator=g)
    gout = torch.randn([3, 3] + outsize, generator=g)
    th1 = theta.to(cuda).requires_grad_(True)
    d1 = data.to(cuda).requires_grad_(True)
    out1 = T.transform_data(d1, th1, outsize)
    (out1 * gout.to(cuda)).sum().backward()
    th2 = theta.to(cuda).requires_grad_(True)
    d2 = data.to(cuda).requires_grad_(True)
    gt = T.transform_grid(T.uniform_meshgrid(outsize), th2)
    out2 = T.interpolate(d2, gt, outsize)
    (out2 * gout.to(cuda)).sum().backward()
    assert torch.allclose(out1, out2, atol=1e-6, rtol=0)
    assert torch.allclose(d1.grad, d2.grad, atol=1e-5, rtol=0)
    assert _rel(th1.grad.cpu().double().numpy(), th2.grad.cpu().double().numpy()) <= 1e-4


@pytest.mark.parametrize("n_theta,size", [(64, [256, 256]), (128, [256, 256]), (1400, [90, 40])])
def test_transform_data_backward_with_several_tiles_per_block(cuda, n_theta, size):
    """With enough thetas the fused 2D backward runs 2 (64 thetas of 256^2)
    or 4 (128 of 256^2; 1400 of 90 x 40, whose 3 x 3 tiles leave a partial
    group of tile rows and partial tiles) output tiles per block, down one
    tile column, into one row of partials; both gradients equal the
    unfused path's."""
    T, b = make_cpab("2d_4x4")
    theta = _theta(T.get_theta_dim(), n_theta, 0.3, 51)
    g = torch.Generator().manual_seed(52)
    data = torch.rand([n_theta, 3] + size, generator=g).to(cuda)
    gout = torch.randn([n_theta, 3] + size, generator=g).to(cuda)
    th1 = theta.to(cuda).requires_grad_(True)
    d1 = data.clone().requires_grad_(True)
    (T.transform_data(d1, th1, size) * gout).sum().backward()
    th2 = theta.to(cuda).requires_grad_(True)
    d2 = data.clone().requires_grad_(True)
    gt = T.transform_grid(T.uniform_meshgrid(size), th2)
    (T.interpolate(d2, gt, size) * gout).sum().backward()
    assert torch.allclose(d1.grad, d2.grad, atol=1e-5, rtol=0)
    assert _rel(th1.grad.cpu().double().numpy(), th2.grad.cpu().double().numpy()) <= 1e-4


def test_identity_theta_is_exact_identity(cuda):
    T, b = make_cpab("2d_4x4")
    grid = T.uniform_meshgrid([64, 48])
    th = T.identity(4)
    out = T.transform_grid(grid, th)
    assert torch.equal(out, grid[None].expand(4, -1, -1))
    data = torch.rand(4, 3, 64, 48, device=cuda)
    # x * (W-1) of a float32 linspace value is within an ulp of the pixel
    # index, so the identity reproduces the data to ~1e-6
    assert torch.allclose(T.transform_data(data, th, [64, 48]), data, atol=1e-5, rtol=0)


def test_theta_gradient_is_bit_reproducible(cuda):
    """dL/dtheta adds in a fixed order (wave sums in a fixed DPP order, one
    lane per wave adding into that wave's own accumulator, tiles summed in
    order): two runs on the same inputs give the same bits, for the fused
    image path and the grid path."""
    T, b = make_cpab("2d_4x4")
    theta = _theta(T.get_theta_dim(), 6, 0.5, 31).to(cuda)
    g = torch.Generator().manual_seed(32)
    data = torch.rand(6, 3, 96, 80, generator=g).to(cuda)
    grid = T.uniform_meshgrid([96, 80])
    grads = []
    for _ in range(2):
        th = theta.clone().requires_grad_(True)
        (T.transform_data(data, th, [96, 80]) ** 2).sum().backward()
        th2 = theta.clone().requires_grad_(True)
        (T.transform_grid(grid, th2) ** 2).sum().backward()
        grads.append((th.grad.clone(), th2.grad.clone()))
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])


def _lookup_points(nd, nc, zb, rng):
    """Random points (outside the domain too where cells extend there), the
    meshgrid through every vertex and cell centre, and points on the cell
    diagonals and on the facets between the simplices of a square / cube."""
    lo, hi = (-0.2, 1.2) if (nd < 3 or not zb) else (0.0, 1.0)
    pts = [rng.uniform(lo, hi, size=(nd, 4000))]
    pts.append(O.uniform_meshgrid([2 * n + 1 for n in nc]))
    k = rng.integers(0, nc, size=(1500, nd)).T
    t = rng.uniform(0, 1, size=(1, 1500))
    diag = (k + t) / np.array(nc)[:, None]
    pts.append(diag)
    if nd >= 2:
        anti = diag.copy()
        anti[1] = (k[1] + 1 - t[0]) / nc[1]
        pts.append(anti)
    return np.concatenate(pts, axis=1).astype(np.float32)


@pytest.mark.parametrize("name", ["1d_10", "1d_50", "2d_3x3", "2d_4x4", "2d_2x2_nzb", "3d_2x2x2", "3d_3x3x3",
                                  "3d_2x2x2_nzb"])
def test_findcell_is_exact_except_within_1e6_of_a_facet(cuda, name):
    """cpab_findcell (FP32) against the oracle's FP64 numpy-semantics lookup:
    every mismatch lies within 1e-6 (x-space) of the region of the cell the
    GPU returned, measured with the facet table (q = x * nc space, unit
    normals; the signed x-space distance is G / |n * nc|).  3D points
    outside the domain are looked up after the reference's push onto the
    boundary (restated in oracle/numpy_path.py and pinned there), so for them
    the pushed point is measured.  At most 1% of the random points (inside
    the domain in 3D) may differ at all; the points on diagonals and grid
    lines are chosen to sit on facets."""
    T, b = make_cpab(name)
    nc = [int(v) for v in b["nc"]]
    nd = len(nc)
    zb = bool(b["zero_boundary"])
    rng = np.random.default_rng(11)
    pts = _lookup_points(nd, nc, zb, rng)
    if nd == 3:
        pts = np.concatenate([pts, rng.uniform(-0.2, 1.2, size=(3, 2000)).astype(np.float32)], axis=1)
    idx = T.backend.findcellidx(nd, torch.tensor(pts, device=cuda), nc).cpu().numpy()
    p64 = pts.astype(np.float64)
    ref = O.findcell(nd, nc, p64)
    bad = np.flatnonzero(idx != ref)
    inside = np.all((p64 >= 0) & (p64 <= 1), axis=0)
    if nd == 3:
        from oracle.numpy_path import push_inside
        p64 = push_inside(p64, nc)
    q = p64.T * np.array(nc, dtype=np.float64)
    planes, nbr = b["planes"], b["nbr"]
    for i in bad:
        c = idx[i]
        # signed x-space distance to each facet plane n.q + e = 0, q = nc * x
        live = nbr[c] != -2  # 1D/2D domain walls (unused slots) bound no region
        G = (planes[c, live, :nd] @ q[i] + planes[c, live, 3]) / np.linalg.norm(planes[c, live, :nd] * nc, axis=1)
        assert G.max() <= 1e-6, (i, pts[:, i], idx[i], ref[i], G.max())
    rand = np.zeros(pts.shape[1], bool)
    rand[:4000] = True
    if nd == 3:
        rand &= inside
    frac = float((idx != ref)[rand].mean())
    print(name, "points", pts.shape[1], "mismatches", bad.size, "of the random ones", frac)
    assert frac <= 0.01


@pytest.mark.parametrize("name", ["1d_10", "2d_3x3", "2d_2x2_nzb", "3d_2x2x2", "3d_3x3x3", "3d_2x2x2_nzb"])
def test_vectorfield_matches_the_oracle_at_every_point(cuda, name):
    """calc_vectorfield against the oracle's FP64 field at every point.
    Inside the domain the field is continuous, so a point whose FP32 lookup
    lands in the neighbouring cell gets the same velocity up to rounding.  A
    3D point outside the domain takes the map of its pushed point's
    tetrahedron, which jumps where the push rule changes; there the GPU's
    velocity is the FP64 field of the cell its FP32 lookup chose (which the
    lookup test bounds to within 1e-6 of that cell)."""
    T, b = make_cpab(name)
    nc = [int(v) for v in b["nc"]]
    nd = len(nc)
    pts = _lookup_points(nd, nc, bool(b["zero_boundary"]), np.random.default_rng(12))
    theta = _theta(T.get_theta_dim(), 1, 0.5, 12)
    tp = torch.tensor(pts, device=cuda)
    v = T.calc_vectorfield(tp, theta.to(cuda)).cpu().numpy()
    p64 = pts.astype(np.float64)
    A = _A64(b["B"], theta, nd)[0]
    vr = O.vectorfield(nd, nc, A, p64)
    outside = np.any((p64 < 0) | (p64 > 1), axis=0) if nd == 3 else np.zeros(pts.shape[1], bool)
    scale = max(1.0, np.abs(vr).max())
    assert np.abs(v - vr)[:, ~outside].max() <= 1e-5 * scale
    if outside.any():
        idx = T.backend.findcellidx(nd, tp, nc).cpu().numpy()[outside]
        po = p64[:, outside]
        vc = np.einsum("nij,jn->in", A[idx, :, :nd], po) + A[idx, :, nd].T
        assert np.abs(v[:, outside] - vc).max() <= 1e-5 * scale
        assert (idx != O.findcell(nd, nc, po)).mean() <= 0.01


def test_prior_samples_follow_the_prior_covariance(cuda):
    """sample_transformation_with_prior draws theta ~ N(mean, prior_covariance)
    on the GPU: the empirical covariance of 40,000 draws is within 5% (Frobenius,
    relative) of it, and a given mean shifts the draws."""
    T, b = make_cpab("2d_3x3")
    torch.manual_seed(5)
    s = T.sample_transformation_with_prior(40000, length_scale=0.3, output_variance=0.7)
    assert s.shape == (40000, T.get_theta_dim()) and s.device.type == "cuda"
    emp = torch.cov(s.double().t()).cpu().numpy()
    want = T.prior_covariance(0.3, 0.7).numpy()
    assert np.linalg.norm(emp - want) <= 0.05 * np.linalg.norm(want)
    mean = torch.full((T.get_theta_dim(),), 0.25, device=cuda)
    s2 = T.sample_transformation_with_prior(40000, mean=mean, length_scale=0.3, output_variance=0.7)
    assert np.abs(s2.double().mean(0).cpu().numpy() - 0.25).max() < 0.05

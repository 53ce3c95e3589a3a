"""Pins the CPU oracle before anything is compared against it.

Each check is chosen so that a plausible mistake in the oracle (a dropped
series term, a wrong neighbour or facet sign, a transposed adjoint, a missing
saltation term) fails it:
  * the restated fixed-step integrator against the reference's own compiled
    op (golden fixtures; plus the live compiled op where /root/reference is
    mounted): bit-exact,
  * the exact flow against closed forms: the 1D piecewise-exponential
    solution, and expm(A~) for globally affine fields (scipy),
  * flow invariants: identity at theta = 0, semigroup, inverse,
  * the reference's scheme converging to the exact flow as nstep grows,
  * cell sequences against brute-force fine-step integration,
  * the exact gradient against FP64 central finite differences, including a
    discontinuous field where only the saltation (hit-time) terms make it
    right,
  * interpolation against torch.nn.functional.grid_sample and finite
    differences.
"""
import itertools

import numpy as np
import pytest
import torch
from scipy.linalg import expm

import oracle as O
from conftest import load_basis, load_npz
from libcpab_amd.core.tessellation import make_tessellation

RNG = np.random.default_rng(12345)


def _setup(name, n_theta, scale, seed=0):
    b = load_basis(name)
    nc = [int(v) for v in b["nc"]]
    ndim = len(nc)
    rng = np.random.default_rng(seed)
    theta = rng.standard_normal((n_theta, b["B"].shape[1])) * scale
    return b, nc, ndim, O.expand_A(b["B"], theta, ndim), theta


# --------------------------------------------------------------------------
# reference restatement
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["1d_10", "2d_3x3", "3d_2x2x2"])
def test_step_restatement_matches_reference_golden(name):
    f = load_npz("ref_step_%s.npz" % name)
    nc = [int(v) for v in f["nc"]]
    ndim = len(nc)
    fwd = O.step_forward_f32(ndim, nc, int(f["nstep"]), f["points"], f["trels"])
    assert np.array_equal(fwd, f["forward"])
    bwd = O.step_backward_f32(ndim, nc, int(f["nstep"]), f["points"], f["As"], f["Bs"])
    assert np.array_equal(bwd, f["backward"])


@pytest.mark.skipif(O.ref_lib() is None, reason="reference sources not mounted (oracle/_ref not built)")
@pytest.mark.parametrize("name", ["1d_50", "2d_4x4", "2d_2x2_nzb", "3d_2x2x2"])
def test_step_restatement_matches_compiled_reference(name):
    b, nc, ndim, A, theta = _setup(name, 2, 0.6, seed=3)
    pts = RNG.uniform(-0.1, 1.1, size=(ndim, 300)).astype(np.float32)
    tr = O.trels(A, 50).astype(np.float32)
    assert np.array_equal(O.ref_forward(ndim, nc, 50, pts, tr), O.step_forward_f32(ndim, nc, 50, pts, tr))
    Bs = O.bases_Bs(b["B"].astype(np.float32), A.shape[1], ndim)
    A32 = A.astype(np.float32)
    assert np.array_equal(O.ref_backward(ndim, nc, 50, pts[:, :15], A32, Bs),
                          O.step_backward_f32(ndim, nc, 50, pts[:, :15], A32, Bs))


@pytest.mark.skipif(O.ref_lib() is None, reason="reference sources not mounted (oracle/_ref not built)")
@pytest.mark.parametrize("nc", [[7], [3, 3], [4, 2], [2, 3, 2]])
def test_lookup_restatement_matches_compiled_reference(nc):
    import ctypes
    ndim = len(nc)
    pts = RNG.uniform(-0.3, 1.3, size=(ndim, 3000)).astype(np.float32)
    mine = O.findcell_f32(ndim, nc, pts)
    arr = np.array(nc + [1] * (3 - ndim), dtype=np.int32)
    ref = np.array([O.ref_lib().ref_findcellidx(ndim, np.ascontiguousarray(pts[:, i]).ctypes.data_as(ctypes.c_void_p),
                                                arr.ctypes.data_as(ctypes.c_void_p)) for i in range(pts.shape[1])])
    assert np.array_equal(mine, ref)


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nc,zb,d", [([10], True, 9), ([50], True, 49), ([3, 3], True, 34), ([4, 4], True, 62),
                                     ([2, 2], False, 20), ([2, 2, 2], True, 27), ([3, 3, 3], True, 96),
                                     ([2, 2, 2], False, 48)])
def test_basis_dimension_and_null_space(nc, zb, d):
    t = make_tessellation(nc, zero_boundary=zb)
    assert t.B.shape[1] == d
    assert np.abs(t.L @ t.B).max() < 1e-9
    assert np.allclose(t.B.T @ t.B, np.eye(d), atol=1e-10)


@pytest.mark.parametrize("name", ["1d_10", "2d_3x3", "2d_4x4", "2d_2x2_nzb", "3d_2x2x2", "3d_3x3x3", "3d_2x2x2_nzb"])
def test_fixture_basis_spans_the_rebuilt_null_space(name):
    b = load_basis(name)
    t = make_tessellation([int(v) for v in b["nc"]], zero_boundary=bool(b["zero_boundary"]))
    P1 = b["B"] @ b["B"].T
    P2 = t.B @ t.B.T
    assert np.abs(P1 - P2).max() < 1e-9
    planes, nbr = t.facet_table()
    assert np.array_equal(nbr, b["nbr"]) and np.allclose(planes, b["planes"])
    if "amap" in b:
        assert np.array_equal(t.amap, b["amap"])


@pytest.mark.parametrize("nc,zb", [([6], True), ([3, 3], True), ([2, 4], False), ([2, 2, 2], True),
                                   ([2, 2, 2], False), ([3, 2, 1], False)])
def test_facet_table_agrees_with_lookup(nc, zb):
    """Every point lies inside the facet region of the cell the reference's
    lookup assigns it (also outside the domain where the field is valid
    there; in 3D outside the domain: of the region containing it), neighbours
    are symmetric, a facet shared by two cells carries opposite planes on
    its two sides, and the domain boundary has no walls without zero
    boundary."""
    from libcpab_amd.core.exterior3d import build_regions, region_of
    t = make_tessellation(nc, zero_boundary=zb)
    planes, nbr = t.facet_table()
    ndim = len(nc)
    lo, hi = (-0.4, 1.4) if (ndim < 3 or not zb) else (0.0, 1.0)
    pts = RNG.uniform(lo, hi, size=(ndim, 20000))
    c = O.findcell(ndim, nc, pts)
    if ndim == 3 and not zb:
        c = np.where(np.any((pts < 0) | (pts > 1), axis=0), region_of(pts, nc, build_regions(t)[0], t.nC), c)
        assert (c >= t.nC).mean() > 0.5
        assert np.array_equal(t.amap[c[c >= t.nC] - t.nC], O.findcell(3, nc, pts[:, c >= t.nC]))
    q = pts.T * np.array(nc)
    G = np.einsum("pfa,pa->pf", planes[c, :, :ndim], q) + planes[c, :, 3]
    G[nbr[c] == -2] = -1
    assert G.max() < 1e-9
    for ci in range(nbr.shape[0]):
        for f in range(nbr.shape[1]):
            n = nbr[ci, f]
            if n >= 0:
                back = [g for g in range(nbr.shape[1]) if nbr[n, g] == ci]
                assert back
                assert min(np.abs(planes[ci, f] + planes[n, g]).max() for g in back) < 1e-12
    assert zb or (nbr != -1).all()


@pytest.mark.parametrize("nc", [[2, 2, 2], [3, 2, 2], [1, 2, 3]])
def test_exterior_regions_take_the_map_of_the_pushed_lookup(nc):
    """3D without zero boundary: a point outside [0,1]^3 uses the map of the
    tetrahedron the reference's lookup gives it after pushing it onto the
    boundary (libcpab/numpy/findcellidx.py:84-109).  Two restatements of
    that lookup (the oracle's C one, the geometry module's numpy one) agree,
    and the region containing each point names the same tetrahedron."""
    from libcpab_amd.core.exterior3d import build_regions, lookup3, region_of
    t = make_tessellation(nc, zero_boundary=False)
    t.facet_table()
    xplanes = build_regions(t)[0]
    pts = RNG.uniform(-0.8, 1.8, size=(3, 30000))
    pts = pts[:, np.any((pts < 0) | (pts > 1), axis=0)]
    ref = O.findcell(3, nc, pts)
    assert np.array_equal(lookup3(pts, nc), ref)
    reg = region_of(pts, nc, xplanes, t.nC)
    assert np.array_equal(t.amap[reg - t.nC], ref)


# --------------------------------------------------------------------------
# exact flow
# --------------------------------------------------------------------------
def _flow1d_analytic(x, A, nx, T=1.0):
    """Piecewise exact 1D flow from the closed form x(t) = (x0 + b/a) e^{at} - b/a
    and hit times t = log(v(xb)/v(x0))/a, written independently of the oracle."""
    c = min(max(int(np.floor(x * nx)), 0), nx - 1)
    rem = T
    for _ in range(10000):
        a, b = A[c]
        v = a * x + b
        if v == 0:
            return x
        right = v > 0
        xb = (c + 1) / nx if right else c / nx
        at_end = (right and c == nx - 1) or (not right and c == 0)
        th = np.inf
        if not at_end:
            vb = a * xb + b
            if vb / v > 0:
                th = np.log(vb / v) / a if a != 0 else (xb - x) / v
        if th >= rem:
            return (x + b / a) * np.exp(a * rem) - b / a if a != 0 else x + b * rem
        x, rem, c = xb, rem - th, c + (1 if right else -1)
    raise AssertionError("no convergence")


@pytest.mark.parametrize("name,scale", [("1d_10", 0.8), ("1d_50", 1.5)])
def test_exact_1d_matches_closed_form(name, scale):
    b, nc, ndim, A, theta = _setup(name, 3, scale, seed=5)
    pts = np.concatenate([RNG.uniform(0, 1, 400), np.linspace(0, 1, 51)])[None]
    out = O.cf_forward(1, nc, A, b["planes"], b["nbr"], pts)
    ref = np.array([[_flow1d_analytic(x, A[k, :, 0, :], nc[0]) for x in pts[0]] for k in range(A.shape[0])])
    assert np.abs(out[:, 0] - ref).max() < 1e-12


@pytest.mark.parametrize("nc", [[4], [3, 3], [2, 2], [2, 2, 2]])
def test_exact_affine_field_matches_expm(nc):
    """A globally affine field lies in span(B) without the zero-boundary
    constraint; its flow is expm(A~) for every point."""
    ndim = len(nc)
    zb = ndim == 3  # 3D: use a field with a fixed point at the centre so points stay inside
    t = make_tessellation(nc, zero_boundary=False) if ndim < 3 else make_tessellation(nc, zero_boundary=True)
    planes, nbr = t.facet_table()
    rng = np.random.default_rng(7)
    M = rng.standard_normal((ndim, ndim)) * 0.7
    bvec = rng.standard_normal(ndim) * 0.3
    if zb:
        # a rotation-free linear field around the centre does not satisfy zero
        # normal velocity; use an unconstrained basis for 3D instead
        t = make_tessellation(nc, zero_boundary=True)
    Avec = np.tile(np.concatenate([M, bvec[:, None]], 1).ravel(), t.nC)
    if ndim == 3:
        pytest.skip("3D affine-field check uses test_exact_affine_3d")
    theta = t.B.T @ Avec
    A = O.expand_A(t.B, theta[None], ndim)
    assert np.abs(A.ravel() - Avec).max() < 1e-10
    pts = rng.uniform(-0.2, 1.2, size=(ndim, 500))
    out = O.cf_forward(ndim, nc, A, planes, nbr, pts)
    Sq = np.zeros((ndim + 1, ndim + 1))
    Sq[:ndim] = np.concatenate([M, bvec[:, None]], 1)
    ref = (expm(Sq) @ np.vstack([pts, np.ones((1, 500))]))[:ndim]
    assert np.abs(out[0] - ref).max() < 1e-12


def test_exact_affine_field_3d_free_boundary_matches_expm():
    """3D without zero boundary: a globally affine field lies in span(B); its
    flow is expm(A~) for points anywhere, through the regions outside the
    domain (each using a copy of the same map)."""
    b = load_basis("3d_2x2x2_nzb")
    rng = np.random.default_rng(19)
    M = rng.standard_normal((3, 3)) * 0.6
    bvec = rng.standard_normal(3) * 0.4
    Avec = np.tile(np.concatenate([M, bvec[:, None]], 1).ravel(), 40)
    theta = b["B"].T @ Avec
    A = O.expand_A(b["B"], theta[None], 3)
    assert np.abs(A.ravel() - Avec).max() < 1e-10
    pts = rng.uniform(-0.4, 1.4, size=(3, 600))
    out = O.cf_forward(3, [2, 2, 2], A, b["planes"], b["nbr"], pts, amap=b["amap"])
    Sq = np.zeros((4, 4))
    Sq[:3] = np.concatenate([M, bvec[:, None]], 1)
    ref = (expm(Sq) @ np.vstack([pts, np.ones((1, 600))]))[:3]
    assert np.abs(out[0] - ref).max() < 1e-12
    assert np.any((ref < 0) | (ref > 1), axis=0).mean() > 0.5


def test_exact_flow_outside_3d_matches_brute_force():
    """3D without zero boundary, points starting inside and outside the
    domain: the exact flow through the outside regions equals a fine RK4
    integration whose velocity looks the reference's own cell up (with its
    push onto the boundary) at every stage, and the exact integrator's cells,
    each outside region mapped to the tetrahedron whose map it uses, are the
    brute-force cell sequence."""
    b = load_basis("3d_2x2x2_nzb")
    nc, amap = [2, 2, 2], b["amap"]
    rng = np.random.default_rng(13)
    theta = rng.standard_normal((1, b["B"].shape[1])) * 0.5
    A = O.expand_A(b["B"], theta, 3)
    n = 200
    pts = rng.uniform(-0.3, 1.3, size=(3, n))
    exact, seq, dur, nseg = O.cf_forward(3, nc, A, b["planes"], b["nbr"], pts, max_segs=64, amap=amap)
    Ac = A[0]

    def vel(x):
        c = O.findcell(3, nc, x)
        return np.einsum("nij,jn->in", Ac[c, :, :3], x) + Ac[c, :, 3].T, c

    N = 10000
    h = 1.0 / N
    x = pts.copy()
    runs = [[] for _ in range(n)]
    for _ in range(N):
        k1, c = vel(x)
        for i in range(n):
            if runs[i] and runs[i][-1][0] == c[i]:
                runs[i][-1][1] += 1
            else:
                runs[i].append([int(c[i]), 1])
        k2, _ = vel(x + h / 2 * k1)
        k3, _ = vel(x + h / 2 * k2)
        k4, _ = vel(x + h * k3)
        x = x + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert np.abs(x - exact[0]).max() < 2e-6
    crossed_outside = 0
    for i in range(n):
        cells = [int(seq[0, i, s]) for s in range(nseg[0, i]) if dur[0, i, s] >= 3 * h]
        mapped = [c if c < 40 else int(amap[c - 40]) for c in cells]
        merged = [c for k, c in enumerate(mapped) if k == 0 or c != mapped[k - 1]]
        brute = [c for c, m in runs[i] if m >= 3]
        brute = [c for k, c in enumerate(brute) if k == 0 or c != brute[k - 1]]
        assert merged == brute, (i, merged, brute)
        crossed_outside += sum(1 for u, v in zip(cells[:-1], cells[1:]) if u >= 40 and v >= 40)
    assert crossed_outside > 50


def test_exact_affine_3d():
    """3D: an affine field with a fixed point at the cube centre, given directly
    per cell (the oracle takes A; continuity holds since the field is global)."""
    nc = [2, 2, 2]
    t = make_tessellation(nc, zero_boundary=True)
    planes, nbr = t.facet_table()
    rng = np.random.default_rng(8)
    M = rng.standard_normal((3, 3)) * 0.25
    M -= np.eye(3) * 0.6  # contracting: trajectories stay in the cube
    bvec = -M @ np.full(3, 0.5)
    A = np.tile(np.concatenate([M, bvec[:, None]], 1), (t.nC, 1, 1))[None]
    pts = rng.uniform(0.1, 0.9, size=(3, 400))
    out = O.cf_forward(3, nc, A, planes, nbr, pts)
    Sq = np.zeros((4, 4))
    Sq[:3] = np.concatenate([M, bvec[:, None]], 1)
    ref = (expm(Sq) @ np.vstack([pts, np.ones((1, 400))]))[:3]
    assert np.abs(out[0] - ref).max() < 1e-12


@pytest.mark.parametrize("name,scale", [("1d_10", 0.8), ("2d_3x3", 0.5), ("2d_2x2_nzb", 0.5), ("3d_2x2x2", 0.3),
                                        ("3d_2x2x2_nzb", 0.6)])
def test_exact_flow_invariants(name, scale):
    b, nc, ndim, A, theta = _setup(name, 2, scale, seed=9)
    am = b.get("amap")
    lo, hi = (0.0, 1.0) if am is None else (-0.3, 1.3)
    pts = RNG.uniform(lo, hi, size=(ndim, 300))
    zero = O.cf_forward(ndim, nc, A * 0, b["planes"], b["nbr"], pts, amap=am)
    assert np.array_equal(zero[0], pts)
    full = O.cf_forward(ndim, nc, A, b["planes"], b["nbr"], pts, amap=am)
    half = O.cf_forward(ndim, nc, A, b["planes"], b["nbr"], pts, T=0.5, amap=am)
    twice = np.stack([O.cf_forward(ndim, nc, A[k:k + 1], b["planes"], b["nbr"], half[k], T=0.5, amap=am)[0]
                      for k in range(A.shape[0])])
    assert np.abs(twice - full).max() < 1e-11
    back = np.stack([O.cf_forward(ndim, nc, -A[k:k + 1], b["planes"], b["nbr"], full[k], amap=am)[0]
                     for k in range(A.shape[0])])
    assert np.abs(back - pts).max() < 1e-11
    if b["zero_boundary"]:
        assert full.min() > -1e-12 and full.max() < 1 + 1e-12


@pytest.mark.parametrize("name,scale", [("1d_10", 0.8), ("2d_3x3", 0.5), ("3d_2x2x2", 0.3)])
def test_reference_scheme_converges_to_exact(name, scale):
    b, nc, ndim, A, theta = _setup(name, 1, scale, seed=11)
    pts = RNG.uniform(0, 1, size=(ndim, 200))
    exact = O.cf_forward(ndim, nc, A, b["planes"], b["nbr"], pts)
    errs = [np.abs(O.step_forward_f64(ndim, nc, n, pts, O.trels(A, n)) - exact).max() for n in (50, 500, 5000)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-7
    assert errs[0] / errs[2] > 300  # second order in 1/nstep


@pytest.mark.parametrize("name,scale", [("2d_3x3", 0.5), ("1d_10", 0.8), ("3d_2x2x2", 0.3), ("2d_4x4", 0.3)])
def test_reference_gradient_converges_to_the_exact_gradient(name, scale):
    """The reference's own gradient -- midpoint-RK2 sensitivities over nstep
    steps (libcpab/core/cpab_ops.cpp:292-367, restated bit-exactly by
    step_backward_f32) -- agrees with the exact gradient within its
    discretisation error at the default 50 steps and approaches it as nstep
    grows.  Its steps take the cell at their start, so steps that straddle a
    facet make the convergence about first order and uneven."""
    b, nc, ndim, A, theta = _setup(name, 1, scale, seed=21)
    rng = np.random.default_rng(21)
    rng.standard_normal((1, b["B"].shape[1]))
    pts = rng.uniform(0.05, 0.95, size=(ndim, 60))
    w = rng.standard_normal((1, ndim, 60))
    gA, _ = O.cf_backward(ndim, nc, A, b["planes"], b["nbr"], pts, w)
    exact = (gA.reshape(1, -1) @ b["B"])[0]
    Bs = O.bases_Bs(b["B"], A.shape[1], ndim)
    errs = {}
    for n in (12, 50, 400):
        J = O.step_backward_f32(ndim, nc, n, pts, A, Bs).astype(np.float64)
        g = (J[:, 0] * w[0][None]).sum((1, 2))
        errs[n] = np.abs(g - exact).max() / np.abs(exact).max()
    assert errs[50] < 2e-2, errs
    assert errs[400] < errs[12] / 10, errs


@pytest.mark.parametrize("name,scale", [("2d_3x3", 0.5), ("3d_2x2x2", 0.3), ("1d_50", 0.6)])
def test_cell_sequences_match_brute_force(name, scale):
    """Cells visited by the exact integrator equal those of a fine RK4
    integration that looks the cell up at every step (segments shorter than the
    RK4 step are not resolvable by brute force and are ignored)."""
    b, nc, ndim, A, theta = _setup(name, 1, scale, seed=13)
    pts = RNG.uniform(0.02, 0.98, size=(ndim, 40))
    _, seq, dur, nseg = O.cf_forward(ndim, nc, A, b["planes"], b["nbr"], pts, max_segs=64)
    h = 1.0 / 20000
    Ac = A[0]
    n = pts.shape[1]

    def vel(x):
        c = O.findcell(ndim, nc, x)
        return np.einsum("nij,jn->in", Ac[c, :, :ndim], x) + Ac[c, :, ndim].T, c

    x = pts.copy()
    runs = [[] for _ in range(n)]
    for _ in range(20000):
        k1, c = vel(x)
        for i in range(n):
            if runs[i] and runs[i][-1][0] == c[i]:
                runs[i][-1][1] += 1
            else:
                runs[i].append([int(c[i]), 1])
        k2, _ = vel(x + h / 2 * k1)
        k3, _ = vel(x + h / 2 * k2)
        k4, _ = vel(x + h * k3)
        x = x + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    for i in range(n):
        brute = [c for c, m in runs[i] if m >= 3]
        exact = [int(seq[0, i, s]) for s in range(nseg[0, i]) if dur[0, i, s] >= 3 * h]
        assert brute == exact, (i, brute, exact)


# --------------------------------------------------------------------------
# exact gradient
# --------------------------------------------------------------------------
def _fd_theta(b, nc, ndim, theta, pts, w, ks, eps=1e-6):
    out = []
    am = b.get("amap")
    for k in ks:
        e = np.zeros_like(theta)
        e[:, k] = eps
        fp = O.cf_forward(ndim, nc, O.expand_A(b["B"], theta + e, ndim), b["planes"], b["nbr"], pts, amap=am)
        fm = O.cf_forward(ndim, nc, O.expand_A(b["B"], theta - e, ndim), b["planes"], b["nbr"], pts, amap=am)
        out.append(((fp - fm) * w).sum(axis=(1, 2)) / (2 * eps))
    return np.stack(out, 1)


@pytest.mark.parametrize("name,scale", [("1d_10", 0.8), ("2d_3x3", 0.5), ("2d_2x2_nzb", 0.5), ("3d_2x2x2", 0.3),
                                        ("3d_2x2x2_nzb", 0.6)])
def test_exact_gradient_matches_finite_differences(name, scale):
    """Outside the domain (3D without zero boundary) the regions' maps jump
    across their facets, so there the saltation terms are needed too."""
    b, nc, ndim, A, theta = _setup(name, 2, scale, seed=15)
    am = b.get("amap")
    lo, hi = (0.0, 1.0) if am is None else (-0.3, 1.3)
    pts = RNG.uniform(lo, hi, size=(ndim, 60))
    w = RNG.standard_normal((2, ndim, 60))
    gA, gp = O.cf_backward(ndim, nc, A, b["planes"], b["nbr"], pts, w, amap=am)
    gtheta = gA.reshape(2, -1) @ b["B"]
    ks = list(range(0, theta.shape[1], max(1, theta.shape[1] // 7)))
    fd = _fd_theta(b, nc, ndim, theta, pts, w, ks)
    assert np.abs(gtheta[:, ks] - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max())
    if am is not None:
        gA0, _ = O.cf_backward(ndim, nc, A, b["planes"], b["nbr"], pts, w, saltation=False, amap=am)
        assert np.abs((gA0.reshape(2, -1) @ b["B"])[:, ks] - fd).max() > 1e-3 * np.abs(fd).max()
    eps = 1e-6
    for a in range(ndim):
        e = np.zeros_like(pts)
        e[a, 7] = eps
        fp = O.cf_forward(ndim, nc, A, b["planes"], b["nbr"], pts + e, amap=am)
        fm = O.cf_forward(ndim, nc, A, b["planes"], b["nbr"], pts - e, amap=am)
        fd_p = ((fp - fm) * w).sum(axis=(1, 2)) / (2 * eps)
        assert np.abs(gp[:, a, 7] - fd_p).max() <= 1e-6 * max(1.0, np.abs(fd_p).max())


def test_saltation_terms_needed_for_discontinuous_field():
    """With A perturbed independently per cell the field jumps across facets;
    then the exact gradient needs the hit-time / cell-change (saltation) terms,
    and finite differences confirm them."""
    b = load_basis("2d_3x3")
    nc = [3, 3]
    rng = np.random.default_rng(17)
    theta = rng.standard_normal((1, b["B"].shape[1])) * 0.5
    A = O.expand_A(b["B"], theta, 2)
    A = A + rng.standard_normal(A.shape) * 0.05
    pts = rng.uniform(0.05, 0.95, size=(2, 40))
    w = rng.standard_normal((1, 2, 40))
    gA, _ = O.cf_backward(2, nc, A, b["planes"], b["nbr"], pts, w, saltation=True)
    gA0, _ = O.cf_backward(2, nc, A, b["planes"], b["nbr"], pts, w, saltation=False)
    eps = 1e-7
    fd = np.zeros_like(A)
    for idx in np.ndindex(*A.shape[1:]):
        e = np.zeros_like(A)
        e[(0,) + idx] = eps
        fp = O.cf_forward(2, nc, A + e, b["planes"], b["nbr"], pts)
        fm = O.cf_forward(2, nc, A - e, b["planes"], b["nbr"], pts)
        fd[(0,) + idx] = ((fp - fm) * w).sum() / (2 * eps)
    scale = np.abs(fd).max()
    assert np.abs(gA - fd).max() <= 1e-5 * scale
    assert np.abs(gA0 - fd).max() > 1e-3 * scale


# --------------------------------------------------------------------------
# interpolation
# --------------------------------------------------------------------------
def test_interpolation_matches_grid_sample_2d():
    """Inside the domain, libcpab's bilinear interpolation equals grid_sample
    (align_corners=True) after swapping axes: data[n, c, x, y] is sampled at
    grid (x, y) in [0,1]^2."""
    rng = np.random.default_rng(19)
    data = rng.uniform(size=(2, 3, 9, 13))
    outsize = [7, 5]
    grid = rng.uniform(0, 1, size=(2, 2, 35))
    ref = O.interpolate(2, data, grid, outsize)  # [N, C, 7, 5]
    img = torch.tensor(data).permute(0, 1, 3, 2)  # [N, C, H=13(y), W=9(x)]
    g = torch.tensor(grid).permute(0, 2, 1) * 2 - 1  # [N, nP, 2] (x, y) in [-1, 1]
    s = torch.nn.functional.grid_sample(img, g[:, None], mode="bilinear", align_corners=True,
                                        padding_mode="border")[:, :, 0]  # [N, C, nP]
    mine = ref.transpose(0, 1, 3, 2).reshape(2, 3, -1)  # point order x fastest
    assert np.abs(mine - s.numpy()).max() < 1e-12


@pytest.mark.parametrize("nd,size,outsize", [(1, [40], [57]), (2, [13, 17], [11, 19]), (3, [7, 6, 5], [5, 8, 6])])
def test_torch_interpolation_reference_equals_the_oracle(nd, size, outsize):
    """The plain-torch restatement the fused GPU path is tested against gives
    the oracle's values and both gradients (data, grid), in FP64, also for
    grids outside [0,1] (both corners clamp)."""
    from torch_interp_ref import torch_interpolate
    rng = np.random.default_rng(23 + nd)
    data = rng.random((2, 3, *size))
    grid = rng.uniform(-0.2, 1.2, (2, nd, int(np.prod(outsize))))
    ref = O.interpolate(nd, data, grid, outsize)
    go = rng.standard_normal(ref.shape)
    gd_ref, gg_ref = O.interpolate_backward(nd, data, grid, outsize, go)
    dt = torch.tensor(data, requires_grad=True)
    gt = torch.tensor(grid, requires_grad=True)
    out = torch_interpolate(dt, gt, outsize)
    (out * torch.tensor(go)).sum().backward()
    assert np.abs(out.detach().numpy() - ref).max() < 1e-12
    assert np.abs(dt.grad.numpy() - gd_ref).max() < 1e-12
    assert np.abs(gt.grad.numpy() - gg_ref).max() < 1e-12


def test_interpolation_identity_and_linear():
    rng = np.random.default_rng(21)
    data = rng.uniform(size=(1, 2, 6, 5, 4))
    grid = O.uniform_meshgrid([6, 5, 4])[None]
    assert np.allclose(O.interpolate(3, data, grid, [6, 5, 4]), data, atol=1e-14)
    # data linear in the index is reproduced exactly at any inside point
    xs = np.arange(11.0)
    lin = (3 * xs + 1)[None, None]
    pts = rng.uniform(0, 1, size=(1, 1, 50))
    v = O.interpolate(1, lin, pts, [50])
    assert np.allclose(v[0, 0], 3 * pts[0, 0] * 10 + 1, atol=1e-12)


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_interpolation_backward_matches_finite_differences(ndim):
    rng = np.random.default_rng(23 + ndim)
    size = {1: [9], 2: [6, 7], 3: [4, 5, 3]}[ndim]
    outsize = {1: [11], 2: [3, 4], 3: [2, 3, 2]}[ndim]
    data = rng.uniform(size=(2, 2) + tuple(size))
    nP = int(np.prod(outsize))
    grid = rng.uniform(0.01, 0.99, size=(2, ndim, nP))
    go = rng.standard_normal((2, 2) + tuple(outsize))
    gd, gg = O.interpolate_backward(ndim, data, grid, outsize, go)
    eps = 1e-7
    for _ in range(6):
        n, a, p = rng.integers(2), rng.integers(ndim), rng.integers(nP)
        e = np.zeros_like(grid)
        e[n, a, p] = eps
        fd = ((O.interpolate(ndim, data, grid + e, outsize) - O.interpolate(ndim, data, grid - e, outsize)) * go).sum()
        assert abs(gg[n, a, p] - fd / (2 * eps)) < 1e-5 * max(1, abs(fd / (2 * eps)))
    # linear in data: the data gradient is exact
    e = rng.standard_normal(data.shape)
    lhs = (O.interpolate(ndim, e, grid, outsize) * go).sum()
    assert abs(lhs - (gd * e).sum()) < 1e-10 * max(1, abs(lhs))


def test_touching_mass_bounds_a_weight_perturbation():
    """touching_mass (the per-cell bound the GPU dL/ddata check uses) is the
    sum of |grad_out| over the points reading each cell: it equals the data
    gradient of <|grad_out|, interpolate> with every corner weight set to 1,
    checked here against a direct loop."""
    from torch_interp_ref import touching_mass
    rng = np.random.default_rng(29)
    for nd, size, outsize in [(1, [9], [7]), (2, [5, 4], [3, 6]), (3, [3, 4, 2], [2, 3, 4])]:
        nP = int(np.prod(outsize))
        grid = rng.uniform(-0.2, 1.2, (2, nd, nP))
        g = rng.standard_normal((2, 3, *outsize))
        got = touching_mass(torch.tensor(grid), size, outsize, torch.tensor(np.abs(g))).numpy()
        want = np.zeros((2, 3, *size))
        gp = np.abs(g).transpose(0, 1, *range(nd + 1, 1, -1)).reshape(2, 3, nP)  # x fastest
        for n in range(2):
            for p in range(nP):
                lo = [min(max(int(np.floor(grid[n, a, p] * (size[a] - 1))), 0), size[a] - 1) for a in range(nd)]
                hi = [min(max(int(np.floor(grid[n, a, p] * (size[a] - 1))) + 1, 0), size[a] - 1) for a in range(nd)]
                for bits in itertools.product((0, 1), repeat=nd):
                    idx = tuple(hi[a] if bits[a] else lo[a] for a in range(nd))
                    want[(n, slice(None)) + idx] += gp[n, :, p]
        assert np.allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("nc,zb", [([3, 3], True), ([2, 2], False), ([2, 2, 2], True)])
def test_volume_preserving_basis_is_trace_free(nc, zb):
    """volume_perservation=True adds one zero-trace row per cell
    (libcpab/core/tesselation.py:155-162): every cell's velocity matrix of
    every theta in span(B) is trace-free, B stays an orthonormal null-space
    basis, and the space is strictly smaller than without the constraint."""
    t = make_tessellation(nc, zero_boundary=zb, volume_perservation=True)
    t0 = make_tessellation(nc, zero_boundary=zb)
    ndim = len(nc)
    assert 0 < t.B.shape[1] < t0.B.shape[1]
    assert np.abs(t.L @ t.B).max() < 1e-9
    assert np.allclose(t.B.T @ t.B, np.eye(t.B.shape[1]), atol=1e-10)
    theta = RNG.standard_normal((5, t.B.shape[1]))
    A = (t.B @ theta.T).T.reshape(5, -1, ndim, ndim + 1)
    assert np.abs(np.trace(A[..., :ndim], axis1=2, axis2=3)).max() < 1e-12
    A0 = (t0.B @ RNG.standard_normal((5, t0.B.shape[1])).T).T.reshape(5, -1, ndim, ndim + 1)
    assert np.abs(np.trace(A0[..., :ndim], axis1=2, axis2=3)).max() > 1e-3


# --------------------------------------------------------------------------
# the numpy-backend restatement timed as the "ref-numpy" CPU baseline
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nc", [[7], [50], [3, 3], [4, 2], [2, 2, 2], [3, 3, 3]])
def test_numpy_path_lookup_equals_the_oracle_lookup(nc):
    from oracle import numpy_path as NP
    ndim = len(nc)
    pts = RNG.uniform(-0.3, 1.3, size=(ndim, 20000))
    # points on grid lines, vertices and diagonals
    k = RNG.integers(0, nc, size=(2000, ndim)).T
    t = RNG.uniform(0, 1, size=(1, 2000))
    grid = O.uniform_meshgrid([2 * n + 1 for n in nc])
    pts = np.concatenate([pts, (k + t) / np.array(nc)[:, None], grid], axis=1)
    assert np.array_equal(NP.findcellidx(ndim, pts, nc), O.findcell(ndim, nc, pts))


@pytest.mark.parametrize("name", ["1d_10", "2d_3x3", "2d_2x2_nzb", "3d_2x2x2", "3d_2x2x2_nzb"])
def test_numpy_path_transformer_equals_the_fp64_step_restatement(name):
    from oracle import numpy_path as NP
    b, nc, ndim, A, theta = _setup(name, 2, 0.5, seed=8)
    lo, hi = (0.0, 1.0) if bool(b["zero_boundary"]) else (-0.1, 1.1)
    pts = RNG.uniform(lo, hi, size=(ndim, 500))
    got = NP.transformer(pts, theta, b["B"], nc, 50)
    ref = O.step_forward_f64(ndim, nc, 50, pts, O.trels(A, 50))
    assert np.abs(got - ref).max() < 1e-12


@pytest.mark.parametrize("ndim,size,outsize", [(1, [40], [57]), (2, [13, 17], [11, 19]), (3, [7, 6, 5], [5, 8, 6])])
def test_numpy_path_interpolation_equals_the_nc_restatement(ndim, size, outsize):
    """NWHC layout of libcpab/numpy/interpolation.py against the NC layout of
    the torch backend's interpolation, restated independently."""
    from oracle import numpy_path as NP
    data = RNG.uniform(size=(2, 3) + tuple(size))
    grid = RNG.uniform(-0.1, 1.1, size=(2, ndim, int(np.prod(outsize))))
    nwhc = np.moveaxis(data, 1, -1)
    got = NP.interpolate(ndim, nwhc, grid, outsize)
    ref = np.moveaxis(O.interpolate(ndim, data, grid, outsize), 1, -1)
    assert np.abs(got - ref).max() < 1e-12


# --------------------------------------------------------------------------
# velocity field
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nc", [[5], [3, 3], [2, 2, 2]])
def test_vectorfield_of_a_global_affine_field_is_exact(nc):
    """Every cell carrying the same affine map: v = M x + b at every point,
    whatever cell the lookup picks (outside the domain too)."""
    ndim = len(nc)
    rng = np.random.default_rng(21)
    M = rng.standard_normal((ndim, ndim))
    bvec = rng.standard_normal(ndim)
    nC = O.n_cells(ndim, nc)
    A = np.tile(np.concatenate([M, bvec[:, None]], 1), (nC, 1, 1))
    pts = rng.uniform(-0.3, 1.3, size=(ndim, 2000))
    v = O.vectorfield(ndim, nc, A, pts)
    assert np.abs(v - (M @ pts + bvec[:, None])).max() < 1e-13


@pytest.mark.parametrize("name,scale", [("1d_10", 0.8), ("2d_3x3", 0.5), ("2d_2x2_nzb", 0.5), ("3d_2x2x2", 0.3)])
def test_vectorfield_is_the_time_derivative_of_the_flow(name, scale):
    """(phi_eps(x) - x) / eps -> v(x): the field of the oracle's lookup is the
    one its exact integrator follows (first-order difference, error O(eps))."""
    b, nc, ndim, A, theta = _setup(name, 1, scale, seed=22)
    lo, hi = (0.02, 0.98) if bool(b["zero_boundary"]) else (-0.1, 1.1)
    pts = RNG.uniform(lo, hi, size=(ndim, 400))
    v = O.vectorfield(ndim, nc, A[0], pts)
    errs = []
    for eps in (1e-4, 1e-5):
        d = (O.cf_forward(ndim, nc, A, b["planes"], b["nbr"], pts, T=eps)[0] - pts) / eps
        errs.append(np.abs(d - v).max())
    scale_v = max(np.abs(v).max(), 1.0)
    assert errs[1] < 1e-4 * scale_v
    assert errs[1] < 0.2 * errs[0] + 1e-9  # shrinks with eps


def test_fp32_sample_coordinates_pick_the_float32_cell():
    """index32 forms x * (size - 1) in FP32, as torch float32 tensors do:
    equal to the FP64 interpolation away from pixel edges, and for a float32
    point just below sample 9 of 600 whose FP32 product rounds onto 9 it takes
    the slope of cell [9, 10] (the FP64 product stays in [8, 9])."""
    rng = np.random.default_rng(3)
    data = rng.uniform(size=(1, 2, 600))
    grid = rng.uniform(0, 1, size=(1, 1, 300))
    a = O.interpolate(1, data, grid, [300])
    b = O.interpolate(1, data, grid, [300], index32=True)
    assert np.abs(a - b).max() < 1e-4
    x = np.nextafter(np.float32(9 / 599), np.float32(0))
    assert np.float32(x) * np.float32(599) == 9.0 and float(x) * 599 < 9.0
    pt = np.array([[[float(x)]]])
    go = np.ones((1, 2, 1))
    _, g64 = O.interpolate_backward(1, data, pt, [1], go)
    _, g32 = O.interpolate_backward(1, data, pt, [1], go, index32=True)
    slope = lambda i: (data[0, :, i + 1] - data[0, :, i]).sum() * 599
    assert abs(g64[0, 0, 0] - slope(8)) < 1e-9 and abs(g32[0, 0, 0] - slope(9)) < 1e-9


# --------------------------------------------------------------------------
# the use_slow path (FP64 restatement and its adjoint)
# --------------------------------------------------------------------------
def _slow_setup(name, seed, scale=0.6, nstep=20):
    from oracle import slow_path as S
    b, nc, ndim, A, theta = _setup(name, 2, scale, seed=seed)
    tr = O.trels(A, nstep)
    lo, hi = (0.02, 0.98) if bool(b["zero_boundary"]) else (-0.1, 1.1)
    pts = np.random.default_rng(seed).uniform(lo, hi, size=(ndim, 60))
    return S, b, nc, ndim, tr, pts


@pytest.mark.parametrize("name", ["1d_10", "2d_3x3", "2d_2x2_nzb", "3d_2x2x2", "3d_2x2x2_nzb"])
def test_slow_path_forward_is_the_numpy_stepping_with_the_torch_lookup(name):
    """Away from the domain's upper faces (where the torch 3D lookup clamps at
    1 - 1e-4) the slow path's steps are the numpy backend's."""
    from oracle import numpy_path as NP
    S, b, nc, ndim, tr, pts = _slow_setup(name, 30)
    got = S.slow_forward(pts, tr, nc, 20)
    _, _, _, A, theta = _setup(name, 2, 0.6, seed=30)
    ref = NP.transformer(pts, theta, b["B"], nc, 20)
    assert np.abs(got - ref).max() < 1e-12


@pytest.mark.parametrize("name", ["1d_10", "2d_3x3", "2d_2x2_nzb", "3d_2x2x2", "3d_2x2x2_nzb"])
def test_slow_path_adjoint_matches_central_differences(name):
    """The reverse sweep against FP64 central differences of slow_forward
    w.r.t. every Trel entry and every point coordinate.  Points whose cell
    sequence changes under one of the 1e-7 steps (the map is only piecewise
    smooth) are left out of the functional; most points stay."""
    S, b, nc, ndim, tr, pts = _slow_setup(name, 31)
    eps = 1e-7
    _, _, cs0 = S.slow_forward(pts, tr, nc, 20, keep=True)
    n_theta, nP = tr.shape[0], pts.shape[1]
    steps = []
    for idx in np.ndindex(*tr.shape):
        for sgn in (1.0, -1.0):
            t_ = tr.copy()
            t_[idx] += sgn * eps
            steps.append((t_, pts))
    for a in range(ndim):
        for sgn in (1.0, -1.0):
            p_ = pts.copy()
            p_[a] += sgn * eps
            steps.append((tr, p_))
    stable = np.ones(n_theta * nP, bool)
    for t_, p_ in steps:
        _, _, cs = S.slow_forward(p_, t_, nc, 20, keep=True)
        for a, b_ in zip(cs, cs0):
            stable &= a == b_
    assert stable.mean() > 0.8
    w = np.random.default_rng(32).standard_normal((n_theta, ndim, nP))
    w *= stable.reshape(n_theta, 1, nP)
    gT, gp = S.slow_backward(pts, tr, nc, 20, w)
    fd = np.zeros_like(tr)
    for idx in np.ndindex(*tr.shape):
        tp, tm = tr.copy(), tr.copy()
        tp[idx] += eps
        tm[idx] -= eps
        fd[idx] = ((S.slow_forward(pts, tp, nc, 20) - S.slow_forward(pts, tm, nc, 20)) * w).sum() / (2 * eps)
    assert np.abs(gT - fd).max() <= 1e-6 * max(np.abs(fd).max(), 1.0)
    fdp = np.zeros((n_theta, ndim, nP))
    for a in range(ndim):
        pp, pm = pts.copy(), pts.copy()
        pp[a] += eps
        pm[a] -= eps
        fdp[:, a] = ((S.slow_forward(pp, tr, nc, 20) - S.slow_forward(pm, tr, nc, 20)) * w).sum(1) / (2 * eps)
    assert np.abs(gp - fdp).max() <= 1e-6 * max(np.abs(fdp).max(), 1.0)

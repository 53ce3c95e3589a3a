"""The device integrator (libcpab_amd/csrc/cpab_device.hpp) compiled for the
host and compared with the FP64 oracle on CPU, over meshgrids whose points sit
on facets, vertices and the domain boundary, random interior points and strong
fields.  It is the same source the GPU kernels inline, so its logic (facet
brackets, zero-length crossings, the dip/graze cases) is tested without a GPU.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT
from libcpab_amd.core.tessellation import make_tessellation

SRC = os.path.join(ROOT, "tests", "host_integrator.cpp")
OUT = os.path.join(ROOT, "tests", "_build", "libhostint.so")


def _lib():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    deps = [SRC, os.path.join(ROOT, "libcpab_amd", "csrc", "cpab_device.hpp")]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        if not os.path.exists(hipcc):
            pytest.skip("hipcc not available")
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        subprocess.check_call([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-fPIC", "-shared", "-o", OUT, SRC])
    return ctypes.CDLL(OUT)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def host_transform(nc, B, planes, nbr, theta, grid, T=1.0, want_jac=False):
    nd = len(nc)
    A = O.expand_A(B, theta, nd).astype(np.float32)
    Aq = A.copy()
    for r in range(nd):
        for c in range(nd):
            Aq[..., r, c] = A[..., r, c] * nc[r] / nc[c]
        Aq[..., r, nd] = A[..., r, nd] * nc[r]
    Aq = np.ascontiguousarray(Aq.reshape(A.shape[0], -1))
    grid = np.ascontiguousarray(grid, dtype=np.float32)
    out = np.zeros((theta.shape[0], nd, grid.shape[-1]), np.float32)
    jac = np.zeros((theta.shape[0], nd * nd, grid.shape[-1]), np.float32) if want_jac else None
    ncarr = np.array(list(nc) + [1] * (3 - nd), np.int32)
    _lib().host_transform(nd, _ptr(ncarr), nbr.shape[0], _ptr(Aq), _ptr(np.ascontiguousarray(planes, np.float32)),
                          _ptr(np.ascontiguousarray(nbr, np.int32)), _ptr(grid), grid.shape[-1], theta.shape[0],
                          int(grid.ndim == 3), ctypes.c_float(T), ctypes.c_float(2e-6 * max(nc)), _ptr(out),
                          _ptr(jac) if want_jac else None)
    return (out, jac) if want_jac else out


CASES = [([10], True, 1.0), ([50], True, 0.7), ([3, 3], True, 0.7), ([4, 4], True, 0.5), ([2, 2], False, 0.7),
         ([1, 1], False, 0.8), ([5, 3], True, 0.6), ([2, 2, 2], True, 0.4), ([3, 2, 2], True, 0.3), ([3, 3], False, 0.6),
         ([4, 4], True, 2.0), ([2, 2, 2], True, 1.5)]


@pytest.mark.parametrize("nc,zb,scale", CASES)
def test_host_build_of_device_integrator_matches_exact(nc, zb, scale):
    t = make_tessellation(nc, zero_boundary=zb)
    planes, nbr = t.facet_table()
    nd = len(nc)
    rng = np.random.default_rng(sum(nc) * 7 + int(scale * 10))
    theta = rng.standard_normal((3, t.B.shape[1])) * scale
    side = {1: 201, 2: 31, 3: 11}[nd]
    lo, hi = (0.0, 1.0) if zb else (-0.2, 1.2)
    grid = np.concatenate([O.uniform_meshgrid([side] * nd), rng.uniform(lo, hi, size=(nd, 500))], 1)
    got = host_transform(nc, t.B, planes, nbr, theta, grid)
    ref = O.cf_forward(nd, nc, O.expand_A(t.B, theta, nd), planes, nbr, grid.astype(np.float32).astype(np.float64))
    err = np.abs(got - ref).max(1) / np.maximum(1.0, np.abs(ref).max(1))
    assert err.max() <= 1e-5, err.max()


@pytest.mark.parametrize("nc,zb,scale", [([4, 4], True, 0.5), ([3, 3], False, 0.6), ([2, 2, 2], True, 0.4),
                                          ([3, 2, 2], True, 0.3)])
def test_host_build_jacobian_matches_finite_differences(nc, zb, scale):
    """The per-chunk exp(hM) of the Jacobian (Cayley-Hamilton series) and the
    saltation updates: dq(T)/dq0 against central differences of the FP64 flow."""
    t = make_tessellation(nc, zero_boundary=zb)
    planes, nbr = t.facet_table()
    nd = len(nc)
    rng = np.random.default_rng(5 + sum(nc))
    theta = rng.standard_normal((2, t.B.shape[1])) * scale
    grid = rng.uniform(0.05, 0.95, size=(nd, 400))
    _, jac = host_transform(nc, t.B, planes, nbr, theta, grid, want_jac=True)
    A = O.expand_A(t.B, theta, nd)
    eps = 1e-5
    S = np.array(nc, dtype=np.float64)
    fd = np.zeros((2, nd, nd, grid.shape[1]))
    for c in range(nd):
        gp, gm = grid.copy(), grid.copy()
        gp[c] += eps
        gm[c] -= eps
        d = (O.cf_forward(nd, nc, A, planes, nbr, gp) - O.cf_forward(nd, nc, A, planes, nbr, gm)) / (2 * eps)
        fd[:, :, c, :] = d  # dx_r / dx_c
    # q-space Jacobian -> x-space: dx_r/dx_c = J_q[r][c] * nc[c] / nc[r]
    Jq = jac.reshape(2, nd, nd, -1).astype(np.float64)
    Jx = Jq * S[None, None, :, None] / S[None, :, None, None]
    err = np.abs(Jx - fd).max(axis=(1, 2)) / np.maximum(1.0, np.abs(fd).max(axis=(1, 2)))
    # points whose trajectory passes within eps of a vertex see a kink in the differences
    assert np.quantile(err, 0.98) <= 1e-3, np.quantile(err, 0.98)


def _facet_lines_2d(nx, ny):
    """Interior facets of the 2D tessellation as q-space segments (p0, p1) with
    unit normals: both diagonals of every cell square, and the interior edges
    between squares."""
    lines = []
    r2 = 1.0 / np.sqrt(2.0)
    for i in range(nx):
        for j in range(ny):
            lines.append(((i, j), (i + 1, j + 1), (r2, -r2)))
            lines.append(((i, j + 1), (i + 1, j), (r2, r2)))
    for i in range(1, nx):
        for j in range(ny):
            lines.append(((i, j), (i, j + 1), (1.0, 0.0)))
    for j in range(1, ny):
        for i in range(nx):
            lines.append(((i, j), (i + 1, j), (0.0, 1.0)))
    return lines


@pytest.mark.parametrize("nc,zb,scale", [([4, 4], True, 0.3), ([3, 3], False, 0.5), ([5, 3], True, 0.4)])
def test_host_starts_on_facets_with_shallow_dips(nc, zb, scale):
    """Starts on a facet near the point where the normal velocity vanishes:
    the trajectory dips into the cell by ~G'(0)^2 / (4 G''(0)/2), far less
    than the rounding of G(0) in FP32, and leaves again through the same
    facet (or grazes it).  The crossing back must be found, or the point goes
    on with the wrong cell's field for the rest of the time."""
    t = make_tessellation(nc, zero_boundary=zb)
    planes, nbr = t.facet_table()
    rng = np.random.default_rng(11 + sum(nc))
    theta = rng.standard_normal((3, t.B.shape[1])) * scale
    A = O.expand_A(t.B, theta, 2)
    S = np.array(nc, dtype=np.float64)[:, None]
    offsets = np.array([0.0, 1e-6, -1e-6, 1e-5, -1e-5, 1e-4, -1e-4, 1e-3, -1e-3])
    for k in range(theta.shape[0]):
        pts = []
        for p0, p1, n in _facet_lines_2d(*nc):
            p0, p1, n = np.array(p0, float), np.array(p1, float), np.array(n)
            u = np.array([0.25, 0.75])
            q = p0[:, None] + u[None] * (p1 - p0)[:, None]
            g = n @ (O.vectorfield(2, nc, A[k], q / S) * S)  # normal velocity, linear along the facet
            if g[0] == g[1]:
                continue
            ustar = 0.25 - g[0] * 0.5 / (g[1] - g[0])
            uu = ustar + offsets
            uu = uu[(uu > 0.02) & (uu < 0.98)]
            pts.append((p0[:, None] + uu[None] * (p1 - p0)[:, None]) / S)
        grid = np.concatenate(pts, 1).astype(np.float32)
        got = host_transform(nc, t.B, planes, nbr, theta[k:k + 1], grid)
        ref = O.cf_forward(2, nc, A[k:k + 1], planes, nbr, grid.astype(np.float64))
        err = np.abs(got - ref).max()
        assert err <= 1e-5, (k, err, grid[:, np.abs(got - ref).max(1).argmax()])


def test_host_exit_through_the_facet_entered_by_a_zero_length_crossing():
    """A point that entered its cell by a zero-length crossing, spent a whole
    chunk (|M|_F h = 1/2) inside, and starts the next chunk within tol0 of the
    facet it came in by while moving out: it must be allowed to leave through
    that facet (the no-return guard covers only a run of zero-length
    crossings).  The state is built by flowing back from a point on the facet
    for one chunk plus a step that leaves it 2e-6 inside."""
    nc = [4, 4]
    t = make_tessellation(nc, zero_boundary=True)
    planes, nbr = t.facet_table()
    S = np.array(nc, dtype=np.float64)
    tol0 = 2e-6 * max(nc)
    lines = _facet_lines_2d(*nc)
    rng = np.random.default_rng(21)
    done = 0
    for _ in range(400):
        theta = rng.standard_normal((1, t.B.shape[1])) * 0.5
        A = O.expand_A(t.B, theta, 2)
        p0, p1, n = (np.array(v, float) for v in lines[rng.integers(len(lines))])
        qr = p0 + rng.uniform(0.2, 0.8) * (p1 - p0)  # crossing point, q-space
        v = O.vectorfield(2, nc, A[0], (qr / S)[:, None])[:, 0] * S
        vn = float(n @ v)
        if np.linalg.norm(v) < 0.05 or abs(vn) < 0.3 * np.linalg.norm(v):
            continue
        side = np.sign(vn) * n  # points into the cell the trajectory exits to
        X = int(O.findcell(2, nc, ((qr + 1e-3 * side) / S)[:, None])[0])
        Y = int(O.findcell(2, nc, ((qr - 1e-3 * side) / S)[:, None])[0])
        Aq = A[0, Y, :, :2].astype(np.float32)
        h = np.float32(0.5) / np.sqrt(np.float32((Aq * Aq).sum()))  # the kernel's chunk in cell Y
        delta = 2e-6 / abs(vn)
        back, seq, dur, nseg = O.cf_forward(2, nc, -A, planes, nbr, (qr / S)[:, None], T=float(h) + delta, max_segs=4)
        if nseg[0, 0] != 1 or seq[0, 0, 0] != Y:
            continue  # the backward path must stay inside Y
        x0 = back[0, :, 0].astype(np.float32)
        T = float(h) + delta + 0.25
        ref = O.cf_forward(2, nc, A, planes, nbr, x0.astype(np.float64)[:, None], T=T)[0, :, 0]
        Aqf = A[0] * np.concatenate([S[:, None] / S[None, :], S[:, None]], 1)  # q-space: nc_r/nc_c, b * nc_r
        Aqf = np.ascontiguousarray(Aqf.reshape(-1, 6).astype(np.float32))
        q = (x0 * np.array(nc, np.float32)).astype(np.float32)
        ncarr = np.array(nc + [1], np.int32)
        H_lib = _lib()
        H_lib.host_integrate_state2(_ptr(ncarr), nbr.shape[0], _ptr(Aqf), _ptr(np.ascontiguousarray(planes, np.float32)),
                                    _ptr(np.ascontiguousarray(nbr, np.int32)), _ptr(q), Y, ctypes.c_float(T), X, 1,
                                    ctypes.c_float(tol0))
        err = np.abs(q / S - ref).max()
        assert err <= 1e-5, (done, err)
        done += 1
        if done == 20:
            break
    assert done == 20


@pytest.mark.parametrize("nc,scale", [([3, 3, 3], 0.6), ([2, 2, 2], 0.8), ([3, 2, 2], 0.6)])
def test_host_3d_starts_on_faces_near_tangency(nc, scale):
    """Starts on the interior faces between cubes, moved along the face to
    where the normal velocity vanishes (plus small offsets).  There the
    normal speed is below the FP32 rounding of M q + b, and the two cells'
    fields may both round it to a small exit; the side the point goes to must
    come from the curvature instead."""
    t = make_tessellation(nc, zero_boundary=True)
    planes, nbr = t.facet_table()
    rng = np.random.default_rng(3 + sum(nc))
    theta = rng.standard_normal((2, t.B.shape[1])) * scale
    A = O.expand_A(t.B, theta, 3)
    S = np.array(nc, dtype=np.float64)[:, None]
    for k in range(theta.shape[0]):
        pts = []
        for a in range(3):
            b = (a + 1) % 3
            for i in range(1, nc[a]):
                q = rng.uniform(0.05, 0.95, size=(3, 300)) * S
                q[a] = i
                v = O.vectorfield(3, nc, A[k], q / S)[a] * nc[a]
                q2 = q.copy()
                q2[b] += 1e-3
                dv = (O.vectorfield(3, nc, A[k], q2 / S)[a] * nc[a] - v) / 1e-3
                with np.errstate(divide="ignore", invalid="ignore"):
                    shift = -v / dv
                keep = np.isfinite(shift) & (np.abs(shift) < 0.05)
                q = q[:, keep]
                q[b] += shift[keep]
                for off in (0.0, 1e-6, -1e-5, 1e-4):
                    qq = q.copy()
                    qq[b] += off
                    pts.append(qq / S)
        grid = np.clip(np.concatenate(pts, 1), 0.0, 1.0).astype(np.float32)
        got = host_transform(nc, t.B, planes, nbr, theta[k:k + 1], grid)
        ref = O.cf_forward(3, nc, A[k:k + 1], planes, nbr, grid.astype(np.float64))
        err = np.abs(got - ref).max()
        assert err <= 1e-5, (k, err, grid[:, np.abs(got - ref).max(1).argmax()])


@pytest.mark.parametrize("nc,zb,scale", [([4, 4], True, 0.3), ([4, 4], True, 1.5), ([3, 3], False, 0.6),
                                          ([2, 2, 2], True, 0.5), ([3, 3, 3], True, 1.0)])
def test_host_replay_of_the_chunk_log_is_bit_exact(nc, zb, scale):
    """The backward replays the first chunks the forward logged (duration and
    crossed facet) instead of searching for them again; every replayed chunk,
    and the end state, must equal the forward's bit for bit.  Points on
    facets, vertices and the boundary included."""
    t = make_tessellation(nc, zero_boundary=zb)
    planes, nbr = t.facet_table()
    nd = len(nc)
    rng = np.random.default_rng(7 + sum(nc))
    theta = rng.standard_normal((2, t.B.shape[1])) * scale
    side = {2: 41, 3: 13}[nd]
    grid = np.concatenate([O.uniform_meshgrid([side] * nd), rng.uniform(0, 1, size=(nd, 500))], 1)
    grid = np.ascontiguousarray(grid, dtype=np.float32)
    S = np.array(nc, dtype=np.float64)
    for k in range(theta.shape[0]):
        A = O.expand_A(t.B, theta[k:k + 1], nd)[0]
        Aq = A * np.concatenate([S[:, None] / S[None, :], S[:, None]], 1)
        Aq = np.ascontiguousarray(Aq.reshape(-1, nd * (nd + 1)).astype(np.float32))
        ncarr = np.array(list(nc) + [1] * (3 - nd), np.int32)
        logged = ctypes.c_int(0)
        bad = _lib().host_replay_mismatches(nd, _ptr(ncarr), nbr.shape[0], _ptr(Aq),
                                            _ptr(np.ascontiguousarray(planes, np.float32)),
                                            _ptr(np.ascontiguousarray(nbr, np.int32)), _ptr(grid), grid.shape[1],
                                            ctypes.c_float(1.0), ctypes.c_float(2e-6 * max(nc)), ctypes.byref(logged))
        assert bad == 0, bad
        assert logged.value >= grid.shape[1]


@pytest.mark.parametrize("k,finished", [(20.0, True), (400.0, False)])
def test_host_chunk_budget_exhaustion_is_nan_and_counted(k, finished):
    """Every cell of a [2,2] tessellation contracts towards its own centroid,
    q' = -k (q - c) in q-space, and a point started near the centroid stays
    in the cell.  Chunks obey |M|_F h <= 1/2, so T = 1 takes 2 sqrt(2) k of
    them: k = 20 finishes (q(T) = c + (q0 - c) e^-k), k = 400 would need about
    1130 > 512 and must come back as NaN -- point and Jacobian -- and be
    counted, not returned short of its end."""
    nc = [2, 2]
    t = make_tessellation(nc, zero_boundary=True)
    planes, nbr = t.facet_table()
    cq = t.get_cell_centers() * np.array(nc)  # centroids in q-space
    Aq = np.zeros((t.nC, 2, 3), np.float32)
    Aq[:, 0, 0] = Aq[:, 1, 1] = -k
    Aq[:, :, 2] = k * cq
    pts = (cq + 0.01).T / np.array(nc)[:, None]  # one point per cell, x-space
    out = np.zeros((1, 2, t.nC), np.float32)
    jac = np.zeros((1, 4, t.nC), np.float32)
    ncarr = np.array(nc + [1], np.int32)
    n_bad = _lib().host_transform(2, _ptr(ncarr), t.nC, _ptr(np.ascontiguousarray(Aq.reshape(1, -1))),
                                  _ptr(np.ascontiguousarray(planes, np.float32)),
                                  _ptr(np.ascontiguousarray(nbr, np.int32)),
                                  _ptr(np.ascontiguousarray(pts, np.float32)), t.nC, 1, 0, ctypes.c_float(1.0),
                                  ctypes.c_float(2e-6 * 2), _ptr(out), _ptr(jac))
    if finished:
        assert n_bad == 0
        want = (cq.T + (pts * np.array(nc)[:, None] - cq.T) * np.exp(-k)) / np.array(nc)[:, None]
        assert np.abs(out[0] - want).max() < 1e-6
        assert np.isfinite(jac).all()
    else:
        assert n_bad == t.nC
        assert np.isnan(out).all() and np.isnan(jac).all()

"""CPU oracle for the CPA integration path -- TEST INFRASTRUCTURE ONLY.

Only tests/, the package smoke check and bench.py's cpu_baseline leg import
this package, and only as the checker; the product path (libcpab_amd) never
imports, links or runs anything from here.

Pieces (C functions live in oracle/cpab_oracle.c, built by oracle/Makefile):
  findcell / findcell_f32   cell lookup, libcpab/numpy/findcellidx.py and
                            libcpab/core/cpab_ops.cpp:26-190
  step_forward_* / step_backward_f32
                            the reference's fixed-step integrator,
                            libcpab/core/cpab_ops.cpp:224-372
  trels                     Trel = expm(A~ / nstep) per cell,
                            libcpab/numpy/transformer.py:40-47
  cf_forward / cf_backward  exact FP64 integration and its gradient; with
                            amap (3D without zero boundary) also through the
                            regions outside the domain, each with the map of
                            the tetrahedron libcpab/numpy/findcellidx.py:84-109
                            gives its pushed points
  interpolate / interpolate_backward
                            libcpab/pytorch/interpolation.py:18-171 in FP64
  ref_forward / ref_backward
                            the reference's own compiled op (oracle/_ref)
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "_build", "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libcpab_ref.so")

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p


def build(force=False):
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    if force or not os.path.exists(_LIB) or (
            os.path.getmtime(_LIB) < os.path.getmtime(os.path.join(HERE, "cpab_oracle.c"))):
        subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(_LIB)
    return _lib


def ref_lib():
    """The reference's own CPU op compiled from its sources, or None."""
    global _ref
    if _ref is None and os.path.exists(_REF):
        _ref = ctypes.CDLL(_REF)
    return _ref


def _ptr(a):
    return a.ctypes.data_as(_p)


def _nc(nc):
    return np.ascontiguousarray(list(nc) + [1] * (3 - len(nc)), dtype=np.int32)


def n_cells(ndim, nc):
    return int(np.prod(nc)) * {1: 1, 2: 4, 3: 5}[ndim]


# --------------------------------------------------------------------------
# lookup
# --------------------------------------------------------------------------
def findcell(ndim, nc, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    nP = pts.shape[-1]
    idx = np.zeros(nP, dtype=np.int32)
    lib().oracle_findcell(_i(ndim), _ptr(_nc(nc)), _ptr(pts), _i(nP), _ptr(idx))
    return idx


def findcell_f32(ndim, nc, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    nP = pts.shape[-1]
    idx = np.zeros(nP, dtype=np.int32)
    lib().oracle_findcell_f32(_i(ndim), _ptr(_nc(nc)), _ptr(pts), _i(nP), _ptr(idx))
    return idx


# --------------------------------------------------------------------------
# parameters
# --------------------------------------------------------------------------
def expand_A(B, theta, ndim):
    """A = (B theta^T)^T as [n_theta, nC, ndim, ndim+1]
    (libcpab/pytorch/transformer.py:146-148)."""
    theta = np.atleast_2d(theta)
    A = (np.asarray(B, dtype=np.float64) @ np.asarray(theta, dtype=np.float64).T).T
    return A.reshape(theta.shape[0], -1, ndim, ndim + 1)


def trels(A, nstep):
    """expm(A~ / nstep)[:ndim] per cell (libcpab/numpy/transformer.py:40-47)."""
    from scipy.linalg import expm
    A = np.asarray(A, dtype=np.float64)
    n_theta, nC, ndim, _ = A.shape
    sq = np.zeros((n_theta * nC, ndim + 1, ndim + 1))
    sq[:, :ndim] = A.reshape(-1, ndim, ndim + 1)
    T = np.stack([expm(m / nstep) for m in sq])
    return T[:, :ndim].reshape(n_theta, nC, ndim, ndim + 1)


# --------------------------------------------------------------------------
# reference fixed-step integrator (restated)
# --------------------------------------------------------------------------
def _bcast(points, n_theta):
    return int(points.ndim == 3 and points.shape[0] == n_theta)


def step_forward_f32(ndim, nc, nstep, points, trel):
    points = np.ascontiguousarray(points, dtype=np.float32)
    trel = np.ascontiguousarray(trel, dtype=np.float32)
    n_theta = trel.shape[0]
    nP = points.shape[-1]
    out = np.zeros((n_theta, ndim, nP), dtype=np.float32)
    lib().oracle_step_forward_f32(_i(ndim), _ptr(_nc(nc)), _i(nstep), _ptr(points), _ptr(trel), _i(nP),
                                  _i(n_theta), _i(_bcast(points, n_theta)), _ptr(out))
    return out


def step_forward_f64(ndim, nc, nstep, points, trel):
    points = np.ascontiguousarray(points, dtype=np.float64)
    trel = np.ascontiguousarray(trel, dtype=np.float64)
    n_theta = trel.shape[0]
    nP = points.shape[-1]
    out = np.zeros((n_theta, ndim, nP), dtype=np.float64)
    lib().oracle_step_forward_f64(_i(ndim), _ptr(_nc(nc)), _i(nstep), _ptr(points), _ptr(trel), _i(nP),
                                  _i(n_theta), _i(_bcast(points, n_theta)), _ptr(out))
    return out


def bases_Bs(B, nC, ndim):
    """Bs = B^T viewed as [d, nC, ndim, ndim+1] (libcpab/pytorch/transformer.py:174)."""
    return np.ascontiguousarray(np.asarray(B).T.reshape(-1, nC, ndim, ndim + 1))


def step_backward_f32(ndim, nc, nstep, points, A, Bs):
    points = np.ascontiguousarray(points, dtype=np.float32)
    A = np.ascontiguousarray(A, dtype=np.float32)
    Bs = np.ascontiguousarray(Bs, dtype=np.float32)
    n_theta, d = A.shape[0], Bs.shape[0]
    nP = points.shape[-1]
    grad = np.zeros((d, n_theta, ndim, nP), dtype=np.float32)
    lib().oracle_step_backward_f32(_i(ndim), _ptr(_nc(nc)), _i(nstep), _ptr(points), _ptr(A), _ptr(Bs),
                                   _i(nP), _i(n_theta), _i(d), _i(_bcast(points, n_theta)), _ptr(grad))
    return grad


def ref_forward(ndim, nc, nstep, points, trel):
    """The reference's compiled cpab_forward_op (None if oracle/_ref is absent)."""
    r = ref_lib()
    if r is None:
        return None
    points = np.ascontiguousarray(points, dtype=np.float32)
    trel = np.ascontiguousarray(trel, dtype=np.float32)
    n_theta = trel.shape[0]
    nP = points.shape[-1]
    out = np.zeros((n_theta, ndim, nP), dtype=np.float32)
    r.ref_cpab_forward(_ptr(out), _ptr(points), _ptr(trel), _i(nstep), _ptr(_nc(nc)), _i(ndim), _i(nP),
                       _i(n_theta), _i(_bcast(points, n_theta)))
    return out


def ref_backward(ndim, nc, nstep, points, A, Bs):
    """The reference's compiled cpab_backward_op: [d, n_theta, ndim, nP]."""
    r = ref_lib()
    if r is None:
        return None
    points = np.ascontiguousarray(points, dtype=np.float32)
    A = np.ascontiguousarray(A, dtype=np.float32)
    Bs = np.ascontiguousarray(Bs, dtype=np.float32)
    n_theta, nC, d = A.shape[0], A.shape[1], Bs.shape[0]
    nP = points.shape[-1]
    grad = np.zeros((d, n_theta, ndim, nP), dtype=np.float32)
    r.ref_cpab_backward(_ptr(grad), _ptr(points), _ptr(A), _ptr(Bs), _i(nstep), _ptr(_nc(nc)), _i(n_theta),
                        _i(d), _i(ndim), _i(nP), _i(nC), _i(_bcast(points, n_theta)))
    return grad


# --------------------------------------------------------------------------
# exact integrator
# --------------------------------------------------------------------------
def _virtual(ndim, nc, A, planes, points, amap):
    """With virtual cells (amap non-empty: regions outside the domain, each
    using the map of simplex amap[v]): A extended to every cell of the facet
    table, and the start cell of each point outside [0,1]^ndim -- the region
    whose largest facet value G = n.q + e is the smallest (the one that
    contains it); -1 (the lookup) inside."""
    if amap is None or len(amap) == 0:
        return A, 0, None
    amap = np.asarray(amap, dtype=np.int64)
    nC = A.shape[1]
    A_all = np.ascontiguousarray(np.concatenate([A, A[:, amap]], axis=1))
    p = np.asarray(points, dtype=np.float64)
    q = p * np.asarray(nc, dtype=np.float64).reshape((ndim, 1))
    vp = np.asarray(planes, dtype=np.float64)[nC:]
    used = np.any(vp[:, :, :ndim] != 0, axis=2)
    G = np.einsum("vfa,...an->...vfn", vp[:, :, :ndim], q) + vp[:, :, 3][..., None]
    G = np.where(used[:, :, None], G, -np.inf)
    region = nC + np.argmin(G.max(axis=-2), axis=-2)
    outside = np.any((p < 0.0) | (p > 1.0), axis=-2)
    c0 = np.ascontiguousarray(np.where(outside, region, -1).astype(np.int32))
    return A_all, A_all.shape[1], c0


def cf_forward(ndim, nc, A, planes, nbr, points, T=1.0, max_segs=0, amap=None):
    """Exact flow of the CPA field for time T.  A [n_theta, nC, ndim, ndim+1].
    Returns out [n_theta, ndim, nP] (and cell sequences if max_segs > 0:
    seq [n_theta, nP, max_segs], dur, nseg).  amap: the simplex of each
    virtual cell of the facet table (3D without zero boundary)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    planes = np.ascontiguousarray(planes, dtype=np.float64)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    points = np.ascontiguousarray(points, dtype=np.float64)
    A, nC_all, c0 = _virtual(ndim, nc, A, planes, points, amap)
    n_theta = A.shape[0]
    nP = points.shape[-1]
    out = np.zeros((n_theta, ndim, nP))
    if max_segs:
        seq = np.zeros((n_theta, nP, max_segs), dtype=np.int32)
        dur = np.zeros((n_theta, nP, max_segs))
        nseg = np.zeros((n_theta, nP), dtype=np.int32)
        args = (_i(max_segs), _ptr(seq), _ptr(dur), _ptr(nseg))
    else:
        args = (_i(0), None, None, None)
    lib().oracle_cf_forward(_i(ndim), _ptr(_nc(nc)), _ptr(A), _ptr(planes), _ptr(nbr), _ptr(points), _i(nP),
                            _i(n_theta), _i(_bcast(points, n_theta)), _d(T), _ptr(out), *args, _i(nC_all),
                            _ptr(c0) if c0 is not None else None)
    if max_segs:
        return out, seq, dur, nseg
    return out


def cf_backward(ndim, nc, A, planes, nbr, points, grad_out, T=1.0, saltation=True, amap=None):
    """Exact gradient of <grad_out, psi(points)> w.r.t. A and the points; a
    virtual cell's share goes to the simplex whose map it uses."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    planes = np.ascontiguousarray(planes, dtype=np.float64)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    points = np.ascontiguousarray(points, dtype=np.float64)
    grad_out = np.ascontiguousarray(grad_out, dtype=np.float64)
    nC = A.shape[1]
    A_all, nC_all, c0 = _virtual(ndim, nc, A, planes, points, amap)
    n_theta = A.shape[0]
    nP = points.shape[-1]
    gA = np.zeros_like(A_all)
    gp = np.zeros((n_theta, ndim, nP))
    rc = lib().oracle_cf_backward(_i(ndim), _ptr(_nc(nc)), _ptr(A_all), _ptr(planes), _ptr(nbr), _ptr(points),
                                  _i(nP), _i(n_theta), _i(_bcast(points, n_theta)), _d(T), _ptr(grad_out),
                                  _ptr(gA), _ptr(gp), _i(int(saltation)), _i(nC_all),
                                  _ptr(c0) if c0 is not None else None)
    if rc != 0:
        raise RuntimeError("oracle_cf_backward failed: %d" % rc)
    if nC_all:
        g = gA[:, :nC].copy()
        np.add.at(g, (slice(None), np.asarray(amap, dtype=np.int64)), gA[:, nC:])
        gA = g
    return gA, gp


def vectorfield(ndim, nc, A, pts):
    A = np.ascontiguousarray(A, dtype=np.float64)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    nP = pts.shape[-1]
    v = np.zeros((ndim, nP))
    lib().oracle_vectorfield(_i(ndim), _ptr(_nc(nc)), _ptr(A), _ptr(pts), _i(nP), _ptr(v))
    return v


# --------------------------------------------------------------------------
# grid and interpolation (FP64 restatements)
# --------------------------------------------------------------------------
def uniform_meshgrid(n_points, dtype=np.float64):
    """libcpab/pytorch/functions.py:102-108: x varies fastest."""
    lin = [np.linspace(0.0, 1.0, n) for n in n_points]
    mesh = np.meshgrid(*lin[::-1], indexing="ij")
    return np.stack([m.ravel() for m in mesh[::-1]]).astype(dtype)


def _corners(x, size, index32=False):
    """Corner indices and fraction of coordinate x on an axis of `size`
    samples.  index32: x * (size - 1) is formed in FP32, as the reference's
    float32 tensors (libcpab/pytorch/interpolation.py:75-82) and the GPU do,
    so that a point within FP32 rounding of a sample picks the same cell."""
    if index32:
        x = (np.asarray(x, dtype=np.float32) * np.float32(size - 1)).astype(np.float64)
    else:
        x = x * (size - 1)
    x0 = np.floor(x).astype(np.int64)
    x1 = x0 + 1
    x0 = np.clip(x0, 0, size - 1)
    x1 = np.clip(x1, 0, size - 1)
    return x0, x1, x - x0


def interpolate(ndim, data, grid, outsize, index32=False):
    """Linear interpolation of data [N, C, *size] at grid [N, ndim, nP]
    (libcpab/pytorch/interpolation.py:18-171).  Returns [N, C, *outsize].
    index32: sample coordinates x * (size - 1) in FP32 (see _corners)."""
    data = np.asarray(data, dtype=np.float64)
    grid = np.asarray(grid, dtype=np.float64)
    N, C = data.shape[:2]
    size = data.shape[2:]
    out = np.zeros((N, C, grid.shape[-1]))
    for n in range(N):
        idx, w = _weights(ndim, grid[n], size, index32)
        for ci, wi in zip(idx, w):
            out[n] += data[n][(slice(None),) + tuple(ci)] * wi
    return _to_outsize(out, outsize)


def _weights(ndim, g, size, index32=False):
    cs = [_corners(g[a], size[a], index32) for a in range(ndim)]
    idx, w = [], []
    for bits in np.ndindex(*([2] * ndim)):
        ci = [cs[a][1] if bits[a] else cs[a][0] for a in range(ndim)]
        wi = np.ones(g.shape[-1])
        for a in range(ndim):
            wi = wi * (cs[a][2] if bits[a] else 1 - cs[a][2])
        idx.append(ci)
        w.append(wi)
    return idx, w


def _to_outsize(v, outsize):
    # points are x-fastest; output dims are [out_0, out_1, ...] with out_0 = x
    N, C = v.shape[:2]
    return v.reshape((N, C) + tuple(outsize[::-1])).transpose((0, 1) + tuple(range(len(outsize) + 1, 1, -1)))


def _from_outsize(g, outsize):
    nd = len(outsize)
    N, C = g.shape[:2]
    return g.transpose((0, 1) + tuple(range(nd + 1, 1, -1))).reshape(N, C, -1)


def interpolate_backward(ndim, data, grid, outsize, grad_out, index32=False):
    """Gradients of <grad_out, interpolate(data, grid)> w.r.t. data and grid,
    as torch autograd gives them for libcpab/pytorch/interpolation.py.
    index32: sample coordinates in FP32 (see _corners)."""
    data = np.asarray(data, dtype=np.float64)
    grid = np.asarray(grid, dtype=np.float64)
    go = _from_outsize(np.asarray(grad_out, dtype=np.float64), outsize)
    N, C = data.shape[:2]
    size = data.shape[2:]
    gdata = np.zeros_like(data)
    ggrid = np.zeros_like(grid)
    for n in range(N):
        cs = [_corners(grid[n, a], size[a], index32) for a in range(ndim)]
        for bits in np.ndindex(*([2] * ndim)):
            ci = tuple(cs[a][1] if bits[a] else cs[a][0] for a in range(ndim))
            wi = np.ones(grid.shape[-1])
            for a in range(ndim):
                wi = wi * (cs[a][2] if bits[a] else 1 - cs[a][2])
            vals = data[n][(slice(None),) + ci]  # [C, nP]
            for c in range(C):
                np.add.at(gdata[n, c], ci, go[n, c] * wi)
            dot = np.sum(vals * go[n], axis=0)
            for a in range(ndim):
                dw = np.ones(grid.shape[-1]) * (1.0 if bits[a] else -1.0)
                for b in range(ndim):
                    if b != a:
                        dw = dw * (cs[b][2] if bits[b] else 1 - cs[b][2])
                ggrid[n, a] += dot * dw * (size[a] - 1)
    return gdata, ggrid

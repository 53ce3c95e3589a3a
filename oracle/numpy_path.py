"""The reference's numpy backend restated in vectorised numpy -- TEST AND
BASELINE INFRASTRUCTURE ONLY (bench.py's CPU baseline times it as the
"ref-numpy" path; tests/test_oracle.py pins it).

  findcellidx    cell lookup, libcpab/numpy/findcellidx.py:16-149 (FP64)
  transformer    fixed-step integration, libcpab/numpy/transformer.py:22-61:
                 Trel = expm(A~ / nstep) per cell with scipy.linalg.expm, one
                 matrix at a time as the reference does, then nstep steps of
                 x <- Trel[cell(x)] x~ for all points at once
  interpolate    NWHC linear interpolation, libcpab/numpy/interpolation.py:18-173

The numpy backend has no gradient, so there is no backward here.
"""
import numpy as np
from scipy.linalg import expm


def _round_min(a, b):
    # the reference's mymin: a where a < b, else round(b)
    return np.where(a < b, a, np.round(b))


def _lookup1(p, nx):
    return np.clip(np.floor(p[0] * nx), 0.0, nx - 1).astype(np.int64)


def _lookup2(p, nx, ny):
    ix, iy = 1.0 / nx, 1.0 / ny
    x0 = np.minimum(nx * ix - 1e-8, np.maximum(0.0, p[0]))
    y0 = np.minimum(ny * iy - 1e-8, np.maximum(0.0, p[1]))
    xm, ym = np.fmod(x0, ix), np.fmod(y0, iy)
    u, v = xm / ix, ym / iy
    base = 4 * (_round_min(nx - 1, (x0 - xm) / ix) + _round_min(ny - 1, (y0 - ym) / iy) * nx)
    px, py = p[0], p[1]
    # exterior strips and corner quadrants, then the four triangles of a square
    # (0 top, 1 right, 2 bottom, 3 left), first matching rule wins
    left = px <= 0
    right = px >= nx * ix
    below = py <= 0
    above = py >= ny * iy
    rules = [
        (left & below & (py / iy < px / ix), 0),
        (left & above & (py / iy - ny > -px / ix), 2),
        (left, 3),
        (right & below & (-py / iy > px / ix - nx), 0),
        (right & above & (py / iy - ny > px / ix - nx), 2),
        (right, 1),
        (above, 2),
        ((u < v) & (1 - u < v), 2),
        (u < v, 3),
        ((u >= v) & (1 - u < v), 1),
    ]
    off = np.select([r for r, _ in rules], [k for _, k in rules], default=0)
    return (base + off).astype(np.int64)


def push_inside(p, nc):
    """3D points outside [0,1]^3 moved onto the domain boundary the way the
    3D lookup does before it looks them up (coordinates beyond the faces are
    set to the face, or half a cell inside it for the pushed axis); points
    inside are returned unchanged."""
    p = np.array(p, dtype=np.float64, copy=True)
    inc = 1.0 / np.asarray(nc, dtype=np.float64)
    out = np.any((p < 0.0) | (p > 1.0), axis=0)
    if out.any():
        c = p[:, out] - 0.5
        a = np.abs(c)
        push = np.stack([0.5 * inc[0] * ((a[0] < a[1]) & (a[0] < a[2])),
                         0.5 * inc[1] * ((a[1] < a[0]) & (a[0] < a[2])),
                         0.5 * inc[2] * ((a[2] < a[0]) & (a[0] < a[1]))])
        beyond = a > 0.5
        c = np.where(beyond, np.copysign(0.5 - push, c), c)
        p[:, out] = c + 0.5
    return p


def _lookup3(p, nx, ny, nz, margin=1e-8):
    p = push_inside(p, (nx, ny, nz))
    inc = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    lim = np.array([nx, ny, nz]) * inc - margin
    q = np.minimum(lim[:, None], np.maximum(0.0, p))
    m = np.mod(q, inc[:, None])
    ijk = np.stack([_round_min(n - 1, (q[a] - m[a]) / inc[a]) for a, n in enumerate((nx, ny, nz))])
    i, j, k = ijk
    base = 5 * (i + j * nx + k * nx * ny)
    u, v, w = m / inc[:, None]
    # cubes of odd parity are the even ones turned by a quarter about z
    flip = ((i + j + k) % 2) == 1
    u, v = np.where(flip, v, u), np.where(flip, 1 - u, v)
    rules = [(-u - v + w >= 0, 1), (u + v + w - 2 >= 0, 2), (-u + v - w >= 0, 3), (u - v - w >= 0, 4)]
    off = np.select([r for r, _ in rules], [k_ for _, k_ in rules], default=0)
    return (base + off).astype(np.int64)


def findcellidx(ndim, p, nc, margin3=1e-8):
    """Cell index of each point of p [ndim, nP] (numpy-backend rules; the 3D
    upper clamp 1 - margin3 is 1 - 1e-4 in the torch backend's lookup,
    libcpab/pytorch/findcellidx.py:112-114)."""
    p = np.asarray(p, dtype=np.float64)
    if ndim == 3:
        return _lookup3(p, *nc, margin=margin3)
    return (_lookup1, _lookup2)[ndim - 1](p, *nc)


def trels(As, nstep):
    """expm(A~ / nstep) of every cell matrix A [n, ndim, ndim+1], one scipy
    call per matrix (libcpab/numpy/transformer.py:40-47)."""
    n, nd, _ = As.shape
    sq = np.zeros((n, nd + 1, nd + 1))
    sq[:, :nd] = As
    return np.stack([expm(m / nstep) for m in sq])


def transformer(points, theta, B, nc, nstep=50):
    """points [ndim, nP] or [n_theta, ndim, nP], theta [n_theta, d], B [D, d]
    -> [n_theta, ndim, nP], FP64 throughout."""
    ndim = len(nc)
    theta = np.atleast_2d(np.asarray(theta, dtype=np.float64))
    n_theta = theta.shape[0]
    nC = B.shape[0] // (ndim * (ndim + 1))
    A = (B @ theta.T).T.reshape(n_theta * nC, ndim, ndim + 1)
    T = trels(A, nstep)  # [n_theta * nC, ndim+1, ndim+1]
    pts = np.asarray(points, dtype=np.float64)
    if pts.ndim == 2:
        pts = np.broadcast_to(pts, (n_theta,) + pts.shape)
    nP = pts.shape[-1]
    # homogeneous points, one row per (theta, point)
    x = np.concatenate([pts, np.ones((n_theta, 1, nP))], axis=1).transpose(0, 2, 1).reshape(-1, ndim + 1)
    offs = np.repeat(np.arange(n_theta) * nC, nP)
    for _ in range(nstep):
        c = findcellidx(ndim, x[:, :ndim].T, nc) + offs
        x = np.einsum("pij,pj->pi", T[c], x)
    return x[:, :ndim].reshape(n_theta, nP, ndim).transpose(0, 2, 1)


def interpolate(ndim, data, grid, outsize):
    """data NWHC [N, s0(, s1(, s2)), C], grid [N, ndim, nP] -> [N, out0, ..., C]
    (x-fastest points, output axes in data order)."""
    data = np.asarray(data, dtype=np.float64)
    N = data.shape[0]
    size = data.shape[1:1 + ndim]
    C = data.shape[-1]
    nb = np.repeat(np.arange(N), grid.shape[-1])
    lo, hi, fr = [], [], []
    for a in range(ndim):
        s = grid[:, a].ravel() * (size[a] - 1)
        f = np.floor(s).astype(np.int64)
        i0 = np.clip(f, 0, size[a] - 1)
        i1 = np.clip(f + 1, 0, size[a] - 1)
        lo.append(i0)
        hi.append(i1)
        fr.append((s - i0)[:, None])
    val = np.zeros((nb.size, C))
    for corner in np.ndindex(*([2] * ndim)):
        idx = tuple(hi[a] if corner[a] else lo[a] for a in range(ndim))
        w = np.ones((nb.size, 1))
        for a in range(ndim):
            w = w * (fr[a] if corner[a] else 1 - fr[a])
        val += data[(nb,) + idx] * w
    out = val.reshape((N,) + tuple(outsize[::-1]) + (C,))
    return out.transpose((0,) + tuple(range(ndim, 0, -1)) + (ndim + 1,))

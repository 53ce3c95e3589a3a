"""The reference's use_slow path restated in FP64 numpy -- TEST
INFRASTRUCTURE ONLY (tests/ use it as the checker of the GPU's use_slow
kernels; tests/test_oracle.py pins it).

  slow_forward   libcpab/pytorch/transformer.py:90-128: nstep applications of
                 x <- Trel[cell(x)] [x; 1], the cell from the torch lookup
                 (libcpab/pytorch/findcellidx.py: the numpy rules with the 3D
                 clamp at 1 - 1e-4)
  slow_backward  the gradient torch autograd takes through that loop: the
                 adjoint of the discrete map, by a reverse sweep over the
                 stored states, w.r.t. the Trels and the points
"""
import numpy as np

from .numpy_path import findcellidx


def _cells(ndim, x, nc):
    return findcellidx(ndim, x, nc, margin3=1e-4)


def slow_forward(points, trels, nc, nstep, keep=False):
    """points [ndim, nP] or [n_theta, ndim, nP]; trels [n_theta, nC, ndim,
    ndim+1] -> [n_theta, ndim, nP] (and the states x_k and cells c_k of every
    step with keep=True)."""
    trels = np.asarray(trels, dtype=np.float64)
    n_theta, nC, ndim, _ = trels.shape
    pts = np.asarray(points, dtype=np.float64)
    if pts.ndim == 2:
        pts = np.broadcast_to(pts, (n_theta,) + pts.shape)
    nP = pts.shape[-1]
    x = pts.transpose(0, 2, 1).reshape(-1, ndim).copy()  # [(theta, point), ndim]
    offs = np.repeat(np.arange(n_theta) * nC, nP)
    T = trels.reshape(n_theta * nC, ndim, ndim + 1)
    xs, cs = [], []
    for _ in range(nstep):
        c = _cells(ndim, x.T, nc) + offs
        if keep:
            xs.append(x.copy())
            cs.append(c)
        x = np.einsum("pij,pj->pi", T[c, :, :ndim], x) + T[c, :, ndim]
    out = x.reshape(n_theta, nP, ndim).transpose(0, 2, 1)
    return (out, xs, cs) if keep else out


def slow_backward(points, trels, nc, nstep, grad_out):
    """Gradient of <grad_out, slow_forward(points, trels)> w.r.t. the trels
    [n_theta, nC, ndim, ndim+1] and the points [n_theta, ndim, nP]."""
    trels = np.asarray(trels, dtype=np.float64)
    n_theta, nC, ndim, _ = trels.shape
    _, xs, cs = slow_forward(points, trels, nc, nstep, keep=True)
    T = trels.reshape(n_theta * nC, ndim, ndim + 1)
    lam = np.asarray(grad_out, dtype=np.float64).transpose(0, 2, 1).reshape(-1, ndim)
    g = np.zeros_like(T)
    for k in range(nstep - 1, -1, -1):
        c, x = cs[k], xs[k]
        np.add.at(g, c, np.concatenate([lam[:, :, None] * x[:, None, :], lam[:, :, None]], axis=2))
        lam = np.einsum("pji,pj->pi", T[c, :, :ndim], lam)
    nP = lam.shape[0] // n_theta
    return g.reshape(trels.shape), lam.reshape(n_theta, nP, ndim).transpose(0, 2, 1)

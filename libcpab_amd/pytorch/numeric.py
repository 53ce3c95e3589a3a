"""method='numeric': the reference's fixed-step integrator on the GPU.

Trels = expm(A~ / nstepsolver) per cell (libcpab/pytorch/transformer.py:146-155),
nstepsolver steps of the cell-local Trel (libcpab/core/cpab_ops.cpp:224-260)
and, for the gradient, the midpoint-RK2 sensitivity integration
(libcpab/core/cpab_ops.cpp:262-372) followed by the reduction of
libcpab/pytorch/transformer.py:199-202.  Like the reference, the analytic
backward returns no gradient w.r.t. the points.  numeric_grad=True gives the
finite-difference gradient the reference intended
(libcpab/pytorch/transformer.py:205-295, which does not run as written).
"""
import ctypes

import torch
from torch.autograd.function import once_differentiable

from .. import _native as N
from . import hip_ops as H


def _trels(A, ndim, nstep):
    """A [n, ndim, ndim+1] -> expm(A~/nstep)[:ndim]."""
    A = A.contiguous()
    T = torch.empty_like(A)
    N.check(N.lib().cpab_expm_affine(H._ptr(A), A.shape[0], ndim, 1.0 / nstep, H._ptr(T), H._stream(A.device)))
    return T


def _expand(theta, dev, params):
    As = (dev.B @ theta.t()).t().contiguous()  # [n_theta, D]
    return As.view(theta.shape[0], params.nC, params.ndim, params.ndim + 1)


def _forward(points, theta, params):
    dev = H.device_tess(params, theta.device)
    ndim, n_theta = params.ndim, theta.shape[0]
    As = _expand(theta, dev, params)
    trels = _trels(As.view(-1, ndim, ndim + 1), ndim, params.nstepsolver).view_as(As)
    broadcast = int(points.dim() == 3 and points.shape[0] == n_theta)
    nP = points.shape[-1]
    out = torch.empty(n_theta, ndim, nP, device=theta.device, dtype=torch.float32)
    N.check(N.lib().cpab_ref_forward(H._ptr(out), H._ptr(points), H._ptr(trels), int(params.nstepsolver),
                                     H._int_arr(params.nc), ndim, nP, n_theta, broadcast, H._stream(theta.device)))
    return out, As, broadcast


class NumericAnalyticGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, theta, params):
        H._require_gpu_f32("grid", points)
        H._require_gpu_f32("theta", theta)
        points = points.contiguous()
        theta = theta.contiguous()
        out, As, broadcast = _forward(points, theta, params)
        ctx.params = params
        ctx.broadcast = broadcast
        ctx.save_for_backward(points, theta, As)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        points, theta, As = ctx.saved_tensors
        p = ctx.params
        dev = H.device_tess(p, theta.device)
        n_theta, nP = theta.shape[0], points.shape[-1]
        Bs = dev.B.t().contiguous()  # [d, D] = [d, nC, ndim, ndim+1]
        J = torch.empty(dev.d, n_theta, p.ndim, nP, device=theta.device, dtype=torch.float32)
        N.check(N.lib().cpab_ref_backward(H._ptr(J), H._ptr(points), H._ptr(As.contiguous()), H._ptr(Bs),
                                          int(p.nstepsolver), H._int_arr(p.nc), n_theta, dev.d, p.ndim, nP, p.nC,
                                          ctx.broadcast, H._stream(theta.device)))
        g = J.mul_(grad.contiguous()).sum(dim=(2, 3))
        return None, g.t(), None


class NumericFiniteDiff(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, theta, params):
        H._require_gpu_f32("grid", points)
        H._require_gpu_f32("theta", theta)
        points = points.contiguous()
        theta = theta.contiguous()
        out, _, _ = _forward(points, theta, params)
        ctx.params = params
        ctx.save_for_backward(points, theta, out)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        points, theta, out = ctx.saved_tensors
        h = 0.01
        cols = []
        for k in range(theta.shape[1]):
            t2 = theta.clone()
            t2[:, k] += h
            o2, _, _ = _forward(points, t2, ctx.params)
            cols.append(((o2 - out) / h * grad).sum(dim=(1, 2)))
        return None, torch.stack(cols, dim=1), None


def transformer_numeric(points, theta, params):
    if getattr(params, "numeric_grad", False):
        return NumericFiniteDiff.apply(points, theta, params)
    return NumericAnalyticGrad.apply(points, theta, params)

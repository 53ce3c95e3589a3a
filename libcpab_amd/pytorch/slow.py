"""use_slow=True: the reference's pure-torch integrator
(libcpab/pytorch/transformer.py:90-128) with its full-autograd gradient.

Trels = expm(A~ / nstepsolver) per cell comes from torch.linalg.matrix_exp
(differentiable, as the reference's own expm), then nstepsolver steps
x <- Trel[cell(x)] x~ run as a HIP kernel (cpab_slow_forward).  The backward
kernel (cpab_slow_backward) is the exact adjoint of those steps -- the
gradient torch autograd computes for the reference's loop -- w.r.t. the
points and the Trels; torch autograd carries dL/dTrels through matrix_exp and
A = B theta.  Unlike the reference's fast path, this one returns dL/dpoints,
as the reference's slow path does.
"""
import torch
from torch.autograd.function import once_differentiable

from .. import _native as N
from . import hip_ops as H


class SlowSteps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, trels, params):
        """points [ndim, nP] or [n_theta, ndim, nP]; trels [n_theta, nC, ndim, ndim+1]."""
        H._require_gpu_f32("grid", points)
        H._require_gpu_f32("trels", trels)
        points = points.contiguous()
        trels = trels.contiguous()
        ndim, n_theta = params.ndim, trels.shape[0]
        per_theta = int(points.dim() == 3)
        nP = points.shape[-1]
        out = torch.empty(n_theta, ndim, nP, device=trels.device, dtype=torch.float32)
        N.check(N.lib().cpab_slow_forward(H._ptr(points), H._ptr(trels), int(params.nstepsolver),
                                          H._int_arr(params.nc), ndim, nP, n_theta, per_theta, H._ptr(out),
                                          H._stream(trels.device)))
        ctx.params = params
        ctx.per_theta = per_theta
        ctx.save_for_backward(points, trels)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        points, trels = ctx.saved_tensors
        p = ctx.params
        ndim, n_theta, nP = p.ndim, trels.shape[0], points.shape[-1]
        grad = grad.contiguous()
        # inverse linear part of every Trel (pass 2 carries the adjoint forward)
        tinv = torch.linalg.inv(trels[..., :ndim].double()).float().contiguous()
        gtrels = torch.zeros_like(trels)
        gpts = None
        if ctx.needs_input_grad[0]:
            gpts = torch.empty(n_theta, ndim, nP, device=trels.device, dtype=torch.float32)
        N.check(N.lib().cpab_slow_backward(H._ptr(points), H._ptr(trels), H._ptr(tinv), int(p.nstepsolver),
                                           H._int_arr(p.nc), ndim, nP, n_theta, ctx.per_theta, H._ptr(grad),
                                           H._ptr(gpts), H._ptr(gtrels), H._stream(trels.device)))
        if gpts is not None and not ctx.per_theta:
            gpts = gpts.sum(0)
        return gpts, gtrels, None


def trels_of(theta, params, basis):
    """expm(A~ / nstepsolver)[:ndim] per cell, [n_theta, nC, ndim, ndim+1],
    differentiable w.r.t. theta (libcpab/pytorch/transformer.py:103-112)."""
    ndim, nC = params.ndim, params.nC
    n_theta = theta.shape[0]
    As = (basis @ theta.t()).t().reshape(n_theta * nC, ndim, ndim + 1)
    sq = torch.cat([As, torch.zeros(n_theta * nC, 1, ndim + 1, device=theta.device, dtype=As.dtype)], dim=1)
    T = torch.linalg.matrix_exp(sq / params.nstepsolver)
    return T[:, :ndim, :].reshape(n_theta, nC, ndim, ndim + 1)


def transformer_slow(points, theta, params):
    H._require_gpu_f32("grid", points)
    H._require_gpu_f32("theta", theta)
    dev = H.device_tess(params, theta.device)
    trels = trels_of(theta, params, dev.B)
    return SlowSteps.apply(points, trels, params)

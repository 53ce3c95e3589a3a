"""torch.autograd Functions over the libcpab_hip C ABI.

These replace the reference's _CPABFunction_AnalyticGrad
(libcpab/pytorch/transformer.py:136-202) and the torch-autograd interpolation
(libcpab/pytorch/interpolation.py): every launch goes to the caller's current
stream, buffers come from torch's caching allocator, and nothing synchronises
the host.
"""
import collections
import ctypes

import numpy as np
import torch
from torch.autograd.function import once_differentiable

from .. import _native as N


# --------------------------------------------------------------------------
# per-device tessellation tables
# --------------------------------------------------------------------------
class DeviceTess:
    """B, facet planes and neighbours of one tessellation on one device, plus
    the cpab_tess_t describing them.  Uploaded once (the reference copies B to
    the device on every call, libcpab/pytorch/transformer.py:146)."""

    def __init__(self, params, device):
        self.device = device
        self.ndim = int(params.ndim)
        self.nc = [int(v) for v in params.nc]
        self.nC = int(params.nC)
        self.d = int(params.d)
        self.B = torch.as_tensor(np.ascontiguousarray(params.basis), dtype=torch.float32, device=device).contiguous()
        self.planes = torch.as_tensor(np.ascontiguousarray(params.facet_planes), dtype=torch.float32,
                                      device=device).contiguous()
        self.nbr = torch.as_tensor(np.ascontiguousarray(params.facet_nbr), dtype=torch.int32, device=device).contiguous()
        # points that used up the chunk budget in a forward (their outputs are NaN)
        self.unfinished = torch.zeros(1, dtype=torch.int32, device=device)
        amap = np.asarray(getattr(params, "facet_amap", np.zeros(0)), dtype=np.int32)
        if amap.size and (amap.min() < 0 or amap.max() >= self.nC or params.facet_nbr.shape[0] != self.nC + amap.size):
            raise ValueError("facet_amap must name a simplex for each of the facet table's cells after the nC simplices")
        self.amap = torch.as_tensor(amap if amap.size else np.zeros(1, np.int32), device=device).contiguous()
        t = N.CpabTess()
        t.ndim = self.ndim
        t.nc = N.int3(self.nc)
        t.nC = self.nC
        t.d = self.d
        t.B = self.B.data_ptr()
        t.planes = self.planes.data_ptr()
        t.nbr = self.nbr.data_ptr()
        t.unfinished = self.unfinished.data_ptr()
        t.nV = int(amap.size)
        t.amap = self.amap.data_ptr()
        self.struct = t
        self.ref = ctypes.byref(t)

    def workspace(self, n_theta, nP, outsize=None):
        """Workspace of a transform call on n_theta x nP points, or (outsize
        given) of a transform_data call onto that meshgrid."""
        if outsize is None:
            nbytes = int(N.lib().cpab_workspace_bytes(self.ref, n_theta, nP))
        else:
            nbytes = int(N.lib().cpab_data_workspace_bytes(self.ref, n_theta, _int_arr(outsize)))
        ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
        return ws, nbytes


def device_tess(params, device):
    cache = getattr(params, "_device_tess", None)
    if cache is None:
        cache = {}
        params._device_tess = cache
    key = (device.type, device.index if device.index is not None else torch.cuda.current_device())
    if key not in cache:
        cache[key] = DeviceTess(params, torch.device("cuda", key[1]))
    return cache[key]


def unfinished_points(params, reset=True):
    """Points that used up the chunk budget (NaN outputs) in the forward calls
    of this tessellation since the last reset, summed over devices.  Reads
    device counters: synchronises with the devices' current streams."""
    total = 0
    for dev in getattr(params, "_device_tess", {}).values():
        total += int(dev.unfinished.item())
        if reset:
            dev.unfinished.zero_()
    return total


def _check_unfinished(params):
    if getattr(params, "raise_unfinished", False):
        n = unfinished_points(params)
        if n:
            raise N.CpabError("%d point(s) used up the integrator's chunk budget (512 chunks and crossings per "
                              "point); their outputs are NaN -- the field is too strong for t_end" % n)


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _require_gpu_f32(name, t):
    if not t.is_cuda:
        raise NotImplementedError("%s is on the CPU: this build implements the GPU path only "
                                  "(Cpab(..., backend='pytorch', device='gpu'))" % name)
    if t.dtype != torch.float32:
        raise TypeError("%s must be float32, got %s" % (name, t.dtype))


def _int_arr(vals):
    arr = (ctypes.c_int * 3)()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


# --------------------------------------------------------------------------
# transform_grid
# --------------------------------------------------------------------------
class TransformGrid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, theta, params, t_end, store_jac=None, grid_shape=None):
        """grid_shape (optional [n0, n1]): the 2D points are a meshgrid of that
        shape in uniform_meshgrid order, walked in compact tiles."""
        _require_gpu_f32("grid", points)
        _require_gpu_f32("theta", theta)
        dev = device_tess(params, theta.device)
        points = points.contiguous()
        theta = theta.contiguous()
        n_theta = theta.shape[0]
        ndim = dev.ndim
        if theta.dim() != 2 or theta.shape[1] != dev.d:
            raise ValueError("theta must be [n_theta, %d], got %s" % (dev.d, tuple(theta.shape)))
        if points.dim() not in (2, 3) or points.shape[-2] != ndim:
            raise ValueError("grid must be [%d, nP] or [n_theta, %d, nP], got %s" % (ndim, ndim, tuple(points.shape)))
        if points.dim() == 3 and points.shape[0] != n_theta:
            raise ValueError("a 3D grid needs one grid per theta")
        per_theta = int(points.dim() == 3)
        nP = points.shape[-1]
        out = torch.empty(n_theta, ndim, nP, device=theta.device, dtype=torch.float32)
        need = any(ctx.needs_input_grad[:2])
        jac = None
        store = getattr(params, "store_jacobian", True) if store_jac is None else store_jac
        if need and store:
            jac = torch.empty(int(N.lib().cpab_jac_floats(ndim, n_theta, nP)), device=theta.device,
                              dtype=torch.float32)
        ws, nb = dev.workspace(n_theta, nP)
        shape = _int_arr(grid_shape) if grid_shape is not None else None
        N.check(N.lib().cpab_transform_fwd(dev.ref, _ptr(theta), n_theta, _ptr(points), nP, per_theta, shape,
                                           float(t_end), _ptr(out), _ptr(jac), _ptr(ws), nb, _stream(theta.device)))
        _check_unfinished(params)
        ctx.params = params
        ctx.t_end = t_end
        ctx.per_theta = per_theta
        ctx.grid_shape = grid_shape
        ctx.save_for_backward(points, theta, jac)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        points, theta, jac = ctx.saved_tensors
        dev = device_tess(ctx.params, theta.device)
        grad = grad.contiguous()
        n_theta = theta.shape[0]
        nP = points.shape[-1]
        gtheta = torch.empty(n_theta, dev.d, device=theta.device, dtype=torch.float32)
        gpts = None
        if ctx.needs_input_grad[0]:
            gpts = torch.empty(n_theta, dev.ndim, nP, device=theta.device, dtype=torch.float32)
        ws, nb = dev.workspace(n_theta, nP)
        shape = _int_arr(ctx.grid_shape) if ctx.grid_shape is not None else None
        N.check(N.lib().cpab_transform_bwd(dev.ref, _ptr(theta), n_theta, _ptr(points), nP, ctx.per_theta, shape,
                                           float(ctx.t_end), _ptr(grad), _ptr(jac), _ptr(gtheta), _ptr(gpts),
                                           _ptr(ws), nb, _stream(theta.device)))
        if gpts is not None and not ctx.per_theta:
            gpts = gpts.sum(0)
        return gpts, (gtheta if ctx.needs_input_grad[1] else None), None, None, None, None


# --------------------------------------------------------------------------
# interpolate
# --------------------------------------------------------------------------
def cell_sequences(points, theta, params, max_segs=8, t_end=1.0):
    """Validation: the cells each trajectory passes through (segments of
    positive duration), cells [n_theta, nP, max_segs] (-1 padded) and the
    segment count nseg [n_theta, nP], from the same integration as
    TransformGrid."""
    _require_gpu_f32("points", points)
    _require_gpu_f32("theta", theta)
    dev = device_tess(params, theta.device)
    theta = theta.contiguous()
    points = points.contiguous()
    n_theta, nP = theta.shape[0], points.shape[-1]
    cells = torch.empty(n_theta, nP, max_segs, device=theta.device, dtype=torch.int32)
    nseg = torch.empty(n_theta, nP, device=theta.device, dtype=torch.int32)
    ws, nb = dev.workspace(n_theta, nP)
    N.check(N.lib().cpab_transform_cells(dev.ref, _ptr(theta), n_theta, _ptr(points), nP,
                                         int(points.dim() == 3), float(t_end), max_segs, _ptr(cells), _ptr(nseg),
                                         _ptr(ws), nb, _stream(theta.device)))
    return cells, nseg


class Interpolate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, data, grid, ndim, outsize):
        _require_gpu_f32("data", data)
        _require_gpu_f32("grid", grid)
        data = data.contiguous()
        grid = grid.contiguous()
        Nb, C = data.shape[:2]
        if data.dim() != ndim + 2 or len(outsize) != ndim:
            raise ValueError("data must be [n_batch, C] + %d spatial dims and outsize %d long" % (ndim, ndim))
        if grid.dim() != 3 or grid.shape[0] != Nb or grid.shape[1] != ndim or grid.shape[2] != int(np.prod(outsize)):
            raise ValueError("grid must be [%d, %d, prod(outsize)=%d], got %s"
                             % (Nb, ndim, int(np.prod(outsize)), tuple(grid.shape)))
        out = torch.empty((Nb, C) + tuple(outsize), device=data.device, dtype=torch.float32)
        N.check(N.lib().cpab_interpolate_fwd(ndim, _ptr(data), Nb, C, _int_arr(data.shape[2:]), _ptr(grid),
                                             _int_arr(outsize), _ptr(out), _stream(data.device)))
        ctx.ndim = ndim
        ctx.outsize = tuple(outsize)
        ctx.save_for_backward(data, grid)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        data, grid = ctx.saved_tensors
        grad = grad.contiguous()
        Nb, C = data.shape[:2]
        gdata = torch.zeros_like(data) if ctx.needs_input_grad[0] else None
        ggrid = torch.empty_like(grid) if ctx.needs_input_grad[1] else None
        N.check(N.lib().cpab_interpolate_bwd(ctx.ndim, _ptr(data), Nb, C, _int_arr(data.shape[2:]), _ptr(grid),
                                             _int_arr(ctx.outsize), _ptr(grad), _ptr(gdata), _ptr(ggrid),
                                             _stream(data.device)))
        return gdata, ggrid, None, None


# --------------------------------------------------------------------------
# fused transform_data
# --------------------------------------------------------------------------
_AXES = collections.OrderedDict()


def meshgrid_axes(outsize, device):
    """The axis coordinates of libcpab/pytorch/functions.py:102-108, concatenated.
    Built once per (outsize, device) and kept (read-only) for later calls.
    Under graph capture an uncached tensor is built on the capturing stream
    instead (no host synchronisation, which capture forbids; its memory then
    belongs to the graph, which recomputes it on every replay)."""
    key = (tuple(int(n) for n in outsize), device.type, device.index)
    lin = _AXES.get(key)
    if lin is None:
        lin = torch.cat([torch.linspace(0, 1, int(n), device=device) for n in outsize]).contiguous()
        if lin.is_cuda:
            if torch.cuda.is_current_stream_capturing():
                return lin
            torch.cuda.current_stream(device).synchronize()  # usable from any stream from now on
        _AXES[key] = lin
        if len(_AXES) > 64:
            _AXES.popitem(last=False)
    return lin


class TransformData(torch.autograd.Function):
    @staticmethod
    def forward(ctx, data, theta, params, outsize, t_end):
        _require_gpu_f32("data", data)
        _require_gpu_f32("theta", theta)
        dev = device_tess(params, theta.device)
        data = data.contiguous()
        theta = theta.contiguous()
        n_theta = theta.shape[0]
        C = data.shape[1]
        if theta.dim() != 2 or theta.shape[1] != dev.d:
            raise ValueError("theta must be [n_theta, %d], got %s" % (dev.d, tuple(theta.shape)))
        if data.dim() != dev.ndim + 2 or data.shape[0] != n_theta or len(outsize) != dev.ndim:
            raise ValueError("data must be [n_theta, C] + %d spatial dims" % dev.ndim)
        nP = int(np.prod(outsize))
        lin = meshgrid_axes(outsize, theta.device)
        out = torch.empty((n_theta, C) + tuple(outsize), device=theta.device, dtype=torch.float32)
        saved = None
        if any(ctx.needs_input_grad[:2]) and getattr(params, "store_jacobian", True):
            saved = torch.empty(int(N.lib().cpab_saved_floats(dev.ndim, n_theta, nP)), device=theta.device,
                                dtype=torch.float32)
        plain = getattr(params, "launch", "batched") == "plain"
        ws, nb = dev.workspace(1 if plain else n_theta, nP, outsize)
        fn = N.lib().cpab_transform_data_fwd_plain if plain else N.lib().cpab_transform_data_fwd
        N.check(fn(dev.ref, _ptr(theta), n_theta, _ptr(data), C, _int_arr(data.shape[2:]), _int_arr(outsize),
                   _ptr(lin), float(t_end), _ptr(out), _ptr(saved), _ptr(ws), nb, _stream(theta.device)))
        _check_unfinished(params)
        ctx.params = params
        ctx.outsize = tuple(outsize)
        ctx.t_end = t_end
        ctx.save_for_backward(data, theta, lin, saved)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        data, theta, lin, saved = ctx.saved_tensors
        dev = device_tess(ctx.params, theta.device)
        grad = grad.contiguous()
        n_theta = theta.shape[0]
        C = data.shape[1]
        nP = int(np.prod(ctx.outsize))
        gdata = torch.zeros_like(data) if ctx.needs_input_grad[0] else None
        gtheta = torch.empty(n_theta, dev.d, device=theta.device, dtype=torch.float32)
        plain = getattr(ctx.params, "launch", "batched") == "plain"
        ws, nb = dev.workspace(1 if plain else n_theta, nP, ctx.outsize)
        fn = N.lib().cpab_transform_data_bwd_plain if plain else N.lib().cpab_transform_data_bwd
        N.check(fn(dev.ref, _ptr(theta), n_theta, _ptr(data), C, _int_arr(data.shape[2:]), _int_arr(ctx.outsize),
                   _ptr(lin), float(ctx.t_end), _ptr(grad), _ptr(saved), _ptr(gdata), _ptr(gtheta), _ptr(ws), nb,
                   _stream(theta.device)))
        return gdata, (gtheta if ctx.needs_input_grad[1] else None), None, None, None

"""libcpab's interpolation (libcpab/pytorch/interpolation.py:18-171) in plain
torch ops, the fp32 reference the fused HIP transform_data is tested against.
Pinned to the oracle's interpolation in tests/test_oracle.py."""
import itertools

import numpy as np
import torch


def torch_interpolate(data, grid, outsize):
    """libcpab/pytorch/interpolation.py:18-171 for any ndim in plain torch
    ops: data [N, C, *size], grid [N, ndim, prod(outsize)] (x fastest)."""
    N = data.shape[0]
    size = data.shape[2:]
    nd = len(size)
    nP = int(np.prod(outsize))
    b = torch.arange(N, device=data.device).repeat_interleave(nP)
    lo, hi, fd = [], [], []
    for a in range(nd):
        x = grid[:, a].reshape(-1) * (size[a] - 1)
        x0 = torch.floor(x).long()
        x1 = torch.clamp(x0 + 1, 0, size[a] - 1)
        x0 = torch.clamp(x0, 0, size[a] - 1)
        lo.append(x0)
        hi.append(x1)
        fd.append((x - x0.float()).reshape(-1, 1))
    # corner values [N * nP, C], keyed by (bit of axis 0, bit of axis 1, ...)
    vals = {}
    for bits in itertools.product((0, 1), repeat=nd):
        idx = [hi[a] if bits[a] else lo[a] for a in range(nd)]
        vals[bits] = data[(b, slice(None), *idx)]
    # blend axis 0 first, then axis 1, then axis 2
    for a in range(nd):
        vals = {k[1:]: vals[(0,) + k[1:]] * (1 - fd[a]) + vals[(1,) + k[1:]] * fd[a]
                for k in vals if k[0] == 0}
    c = vals[()]
    C = c.shape[1]
    c = c.reshape(N, *reversed(list(outsize)), C)
    return c.permute(0, nd + 1, *range(nd, 0, -1)).contiguous()


def touching_mass(grid, size, outsize, gabs):
    """Per data cell, the sum of gabs [N, C, *outsize] over the points whose
    interpolation reads that cell (any of the 2^ndim corners, weight 1): the
    bound on how much dL/ddata of a cell can move when every weight moves by
    at most 1.  Same indexing as torch_interpolate."""
    N, C = gabs.shape[:2]
    nd = len(size)
    nP = int(np.prod(outsize))
    # [N, C, out_0, out_1, ...] -> point order (x fastest), as the grid
    g = gabs.permute(0, 1, *range(nd + 1, 1, -1)).reshape(N, C, nP)
    lo, hi = [], []
    for a in range(nd):
        x = grid[:, a].reshape(N, nP) * (size[a] - 1)
        x0 = torch.floor(x).long()
        hi.append(torch.clamp(x0 + 1, 0, size[a] - 1))
        lo.append(torch.clamp(x0, 0, size[a] - 1))
    strides = [int(np.prod(size[a + 1:])) for a in range(nd)]
    out = torch.zeros(N, C, int(np.prod(size)), dtype=gabs.dtype, device=gabs.device)
    for bits in itertools.product((0, 1), repeat=nd):
        flat = sum((hi[a] if bits[a] else lo[a]) * strides[a] for a in range(nd))  # [N, nP]
        out.scatter_add_(2, flat[:, None, :].expand(N, C, nP), g)
    return out.reshape(N, C, *size)

"""The allocation-free fused forward (cpab_transform_data_fwd_noalloc, A_q in
the library's static scratch) against cpab_transform_data_fwd (A_q in a
workspace): the same kernels, so outputs and saved state must be identical
bit for bit.  Two calls in a row, the second smaller, so the second reads a
scratch array that still holds the first call's larger A_q."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import make_cpab
from libcpab_amd import _native as N
from libcpab_amd.pytorch import hip_ops as H

pytestmark = pytest.mark.gpu


def _inputs(T, n_theta, C, dsize, seed):
    g = torch.Generator().manual_seed(seed)
    theta = (torch.randn(n_theta, T.get_theta_dim(), generator=g) * 0.3).cuda()
    data = torch.rand((n_theta, C) + tuple(dsize), generator=g).cuda()
    return theta, data


def _both(T, theta, data, outsize):
    dev = H.device_tess(T.params, theta.device)
    n_theta, C = data.shape[:2]
    nP = int(np.prod(outsize))
    lin = H.meshgrid_axes(outsize, theta.device)
    stream = H._stream(theta.device)
    res = []
    for noalloc in (False, True):
        out = torch.full((n_theta, C) + tuple(outsize), float("nan"), device=theta.device)
        saved = torch.full((int(N.lib().cpab_saved_floats(dev.ndim, n_theta, nP)),), float("nan"),
                           device=theta.device)
        args = (dev.ref, H._ptr(theta), n_theta, H._ptr(data), C, H._int_arr(data.shape[2:]), H._int_arr(outsize),
                H._ptr(lin), 1.0, H._ptr(out), H._ptr(saved))
        if noalloc:
            N.check(N.lib().cpab_transform_data_fwd_noalloc(*args, stream))
        else:
            ws, nb = dev.workspace(n_theta, nP, outsize)
            N.check(N.lib().cpab_transform_data_fwd(*args, H._ptr(ws), nb, stream))
        torch.cuda.synchronize()
        res.append((out.cpu(), saved.cpu()))
    return res


@pytest.mark.parametrize("name,n1,size1,n2,size2", [
    ("2d_4x4", 64, (128, 128), 3, (40, 24)),
    ("1d_10", 256, (1024,), 5, (100,)),
    ("3d_2x2x2", 8, (24, 24, 24), 2, (9, 10, 11)),
])
def test_noalloc_forward_equals_workspace_forward(cuda, name, n1, size1, n2, size2):
    T, _ = make_cpab(name)
    for k, (n, size) in enumerate(((n1, size1), (n2, size2))):
        theta, data = _inputs(T, n, 3 if len(size) == 2 else 1, size, seed=10 + k)
        (o_ws, s_ws), (o_na, s_na) = _both(T, theta, data, size)
        assert torch.isfinite(o_ws).all()
        # bit patterns: slots of the chunk log a point does not use keep the NaN fill
        assert torch.equal(o_ws.view(torch.int32), o_na.view(torch.int32)), "call %d: outputs differ" % k
        assert torch.equal(s_ws.view(torch.int32), s_na.view(torch.int32)), "call %d: saved state differs" % k


def test_noalloc_rejects_problems_beyond_the_scratch(cuda):
    T, _ = make_cpab("2d_4x4")
    dev = H.device_tess(T.params, cuda)
    per_theta = T.params.nC * 6 * 4
    n_theta = N.CPAB_NOALLOC_SCRATCH_BYTES // per_theta + 1
    outsize = (8, 8)
    theta = torch.zeros(n_theta, T.get_theta_dim(), device=cuda)
    data = torch.zeros(n_theta, 1, 8, 8, device=cuda)
    out = torch.empty(n_theta, 1, 8, 8, device=cuda)
    lin = H.meshgrid_axes(outsize, cuda)
    rc = N.lib().cpab_transform_data_fwd_noalloc(dev.ref, H._ptr(theta), n_theta, H._ptr(data), 1,
                                                 H._int_arr((8, 8)), H._int_arr(outsize), H._ptr(lin), 1.0,
                                                 H._ptr(out), None, H._stream(cuda))
    assert rc == 7
    assert b"scratch" in N.lib().cpab_status_string(rc)

"""The ragged batch (one persistent launch per direction) against the
single-problem fused path, the plain per-problem launches, and graph replays."""
import numpy as np
import pytest
import torch

from libcpab_amd import Cpab
from libcpab_amd.batch import CpabBatch

pytestmark = pytest.mark.gpu

PROBLEMS = [  # tess, zero_boundary, channels, data size, output size
    ([10], True, 1, [300], [257]),
    ([23], True, 2, [64], [100]),
    ([2, 3], True, 3, [40, 33], [37, 29]),
    ([4, 4], True, 1, [64, 64], [64, 64]),
    ([2, 2], False, 2, [17, 21], [19, 23]),
    ([2, 2, 2], True, 1, [9, 8, 7], [9, 8, 7]),
    ([2, 2, 2], True, 2, [10, 9, 8], [11, 9, 8]),  # 3D output tiles
    ([8, 8], True, 3, [48, 40], [48, 40]),  # tables leave no LDS for the dL/ddata box
    ([2, 2, 2], False, 1, [8, 7, 9], [8, 9, 7]),  # regions outside the domain
]


def _make(cuda, seed=0, save=True, problems=PROBLEMS):
    g = torch.Generator().manual_seed(seed)
    cpabs = [Cpab(nc, backend="pytorch", device="gpu", zero_boundary=zb) for nc, zb, _, _, _ in problems]
    probs = [(T, c, s, o) for T, (_, _, c, s, o) in zip(cpabs, problems)]
    b = CpabBatch(probs, device=cuda, save=save)
    for i, (T, c, s, o) in enumerate(probs):
        b.theta_views[i].copy_(torch.randn(T.get_theta_dim(), generator=g) * 0.4)
        b.data_views[i].copy_(torch.rand([c] + s, generator=g))
        b.grad_out_views[i].copy_(torch.randn([c] + o, generator=g))
    return b, probs


def _reference(b, probs):
    outs, gth, gda = [], [], []
    for i, (T, c, s, o) in enumerate(probs):
        th = b.theta_views[i].detach().clone()[None].requires_grad_(True)
        da = b.data_views[i].detach().clone()[None].requires_grad_(True)
        out = T.transform_data(da, th, o)
        (out * b.grad_out_views[i][None]).sum().backward()
        outs.append(out[0].detach())
        gth.append(th.grad[0])
        gda.append(da.grad[0])
    return outs, gth, gda


def _check(b, ref):
    outs, gth, gda = ref
    for i in range(b.n):
        assert torch.allclose(b.out_views[i], outs[i], atol=1e-6, rtol=0), i
        assert torch.allclose(b.grad_data_views[i], gda[i], atol=1e-5, rtol=0), i
        scale = gth[i].abs().max().item()
        assert (b.grad_theta_views[i] - gth[i]).abs().max().item() <= 1e-4 * max(scale, 1e-6), i


@pytest.mark.parametrize("ppt,bpc", [(1, 2), (2, 4), (8, 1)])
@pytest.mark.parametrize("save", [True, False])
def test_batch_matches_single_problem_path(cuda, ppt, bpc, save):
    b, probs = _make(cuda, save=save)
    b.set_launch(ppt, bpc * 256)
    b.forward()
    b.backward()
    torch.cuda.synchronize()
    _check(b, _reference(b, probs))


def test_plain_matches_single_problem_path(cuda):
    b, probs = _make(cuda, seed=1)
    b.set_launch(2, 512)
    b.forward_plain()
    b.backward_plain()
    torch.cuda.synchronize()
    _check(b, _reference(b, probs))


def test_autotune_and_graph_replays(cuda):
    b, probs = _make(cuda, seed=2)
    b.autotune(reps=1)
    assert b.ppt in (1, 2, 4, 8)
    ref = _reference(b, probs)
    b.capture()
    for rep in range(3):
        b.out.zero_()
        b.grad_theta.zero_()
        torch.cuda.synchronize()
        b.replay_forward()
        b.replay_backward()
        torch.cuda.synchronize()
        _check(b, ref)
        # eager work between replays
        x = torch.randn(1000, device=cuda).sum()
        torch.cuda.synchronize()


@pytest.mark.parametrize("save", [True, False])
def test_batch_without_3d_instances_matches_single_problem_path(cuda, save):
    """A batch of 1D/2D instances runs the batch kernels built without the 3D
    path (fewer registers); same results as the single-problem path."""
    b, probs = _make(cuda, save=save, problems=[p for p in PROBLEMS if len(p[0]) < 3])
    b.set_launch(2, 512)
    b.forward()
    b.backward()
    torch.cuda.synchronize()
    _check(b, _reference(b, probs))

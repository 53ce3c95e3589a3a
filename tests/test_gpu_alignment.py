"""CpabAligner on the GPU: the captured Adam loop against the same loop run
eagerly (every replay checked, with other eager work and synchronisations in
between), convergence on a known warp, and the random search."""
import pytest
import torch

from libcpab_amd import Cpab, CpabAligner
from libcpab_amd.alignment import _AdamLoop

pytestmark = pytest.mark.gpu


def _smooth(n, ch, size, seed):
    g = torch.Generator().manual_seed(seed)
    y, x = torch.meshgrid(torch.linspace(0, 1, size[0]), torch.linspace(0, 1, size[1]), indexing="ij")
    fr = torch.rand(n, ch, 3, 2, generator=g) * 2.0 + 0.5
    ph = torch.rand(n, ch, 3, generator=g) * 6.2832
    arg = fr[..., 0, None, None] * x + fr[..., 1, None, None] * y
    return 0.5 + 0.5 * torch.sin(6.2832 * arg + ph[..., None, None]).mean(2)


def _pair(cuda, size=(64, 64), scale=0.4):
    T = Cpab([3, 3], backend="pytorch", device="gpu", zero_boundary=True)
    x1 = _smooth(1, 3, size, 0).to(cuda)
    theta_true = (torch.randn(1, T.get_theta_dim(), generator=torch.Generator().manual_seed(1)) * scale).to(cuda)
    x2 = T.transform_data(x1, theta_true, list(size)).detach()
    return T, x1, x2


def test_captured_iteration_replays_equal_the_eager_loop(cuda):
    T, x1, x2 = _pair(cuda)
    eager = _AdamLoop(T, x1, x2, lr=1e-2, graph_safe=True)
    ref_theta, ref_loss = [], []
    for _ in range(9):
        eager.step()
        ref_theta.append(eager.theta.detach().clone())
        ref_loss.append(eager.loss.clone())
    loop = _AdamLoop(T, x1, x2, lr=1e-2, graph_safe=True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            loop.step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert torch.equal(loop.theta.detach(), ref_theta[2])
    loop.capture()
    other = torch.rand(2, 3, 40, 40, device=cuda)
    th_other = torch.randn(2, T.get_theta_dim(), device=cuda) * 0.3
    for i in range(3, 9):
        loop.replay()
        torch.cuda.synchronize()
        assert torch.equal(loop.theta.detach(), ref_theta[i]), i
        assert torch.equal(loop.loss, ref_loss[i]), i
        # eager work of the program between replays
        assert torch.isfinite(T.transform_data(other, th_other, [40, 40]).sum()).item()


@pytest.mark.parametrize("graph", [False, True])
def test_alignment_by_gradient_reduces_the_misfit(cuda, graph):
    T, x1, x2 = _pair(cuda)
    A = CpabAligner(T)
    theta = A.alignment_by_gradient(x1, x2, maxiter=150, lr=2e-2, graph=graph, progress=False)
    assert theta.shape == (1, T.get_theta_dim())
    start = torch.norm(x1 - x2).item()
    end = torch.norm(T.transform_data(x1, theta.detach(), [64, 64]) - x2).item()
    assert end < 0.5 * start, (start, end)


def test_alignment_by_sampling_never_increases_the_misfit(cuda):
    T, x1, x2 = _pair(cuda)
    torch.manual_seed(0)
    theta = CpabAligner(T).alignment_by_sampling(x1, x2, maxiter=30, progress=False)
    assert theta.shape == (1, T.get_theta_dim())
    start = torch.norm(x1 - x2).item()
    end = torch.norm(T.transform_data(x1, theta, [64, 64]) - x2).item()
    assert end <= start * (1 + 1e-6)


def test_first_transform_data_of_an_output_size_under_graph_capture(cuda):
    """The meshgrid axes of an output size not seen before are built inside
    the capture (no host synchronisation, no cache entry); replays equal the
    eager call, with eager work and synchronisations between them."""
    T = Cpab([3, 3], backend="pytorch", device="gpu", zero_boundary=True)
    g = torch.Generator().manual_seed(3)
    theta = (torch.randn(2, T.get_theta_dim(), generator=g) * 0.4).to(cuda)
    data = torch.rand(2, 3, 45, 39, generator=g).to(cuda)
    size = [37, 41]  # used by no other test
    with torch.no_grad():  # tables on the device, kernels loaded: eager, another size
        T.transform_data(data, theta, [20, 20])
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.no_grad():
        with torch.cuda.graph(graph, stream=s):
            out = T.transform_data(data, theta, size)
        torch.cuda.synchronize()
        eager = T.transform_data(data, theta, size)
        for _ in range(3):
            out.zero_()
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, eager)
            torch.randn(1000, device=cuda).sum().item()

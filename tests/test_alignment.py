"""CpabAligner (libcpab/alignment.py): host-side pieces on the CPU."""
import pytest
import torch

from libcpab_amd import Cpab, CpabAligner
from libcpab_amd.alignment import _norm_two_stage


def test_aligner_needs_a_cpab_instance():
    with pytest.raises(AssertionError):
        CpabAligner(object())
    T = Cpab([2, 2], backend="pytorch", device="gpu")
    assert CpabAligner(T).T is T


@pytest.mark.parametrize("shape", [(1,), (1000,), (1, 3, 256, 256), (1, 1, 1023), (2, 5, 7, 11)])
def test_two_stage_norm_equals_the_norm(shape):
    x = torch.randn(shape, generator=torch.Generator().manual_seed(len(shape)), dtype=torch.float64)
    assert torch.allclose(_norm_two_stage(x), torch.norm(x), rtol=1e-12, atol=0)
    x32 = x.float().requires_grad_(True)
    _norm_two_stage(x32).backward()
    assert torch.allclose(x32.grad, x.float() / torch.norm(x.float()), rtol=1e-5, atol=1e-7)

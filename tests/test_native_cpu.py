"""CPU-side checks of the boundary: the HIP library loads without a GPU and
exports exactly what include/cpab_hip.h declares, argument errors come back
as status codes (no GPU work is started for them), and the Cpab facade keeps
the reference's host-side contract."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from conftest import ROOT, load_basis
from libcpab_amd import _native as N

HEADER = os.path.join(ROOT, "include", "cpab_hip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cpab_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert sorted(N.EXPORTED) == _declared()


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.cpab_abi_version() == N.ABI_VERSION == 7
    assert lib.cpab_status_string(0) == b"ok"
    assert b"workspace" in lib.cpab_status_string(3)


def _tess(name="2d_3x3"):
    b = load_basis(name)
    t = N.CpabTess()
    nc = [int(v) for v in b["nc"]]
    t.ndim = len(nc)
    t.nc = N.int3(nc)
    t.nC = b["nbr"].shape[0]
    t.d = b["B"].shape[1]
    t.B = 1  # never dereferenced by the argument checks below
    t.planes = 1
    t.nbr = 1
    return t


def test_argument_errors_are_status_codes():
    L = N.lib()
    t = _tess()
    # bad ndim
    bad = _tess()
    bad.ndim = 4
    assert L.cpab_transform_fwd(ctypes.byref(bad), 1, 1, 1, 10, 0, None, 1.0, 1, None, 1, 1 << 20, None) == 2
    # inconsistent cell count
    bad = _tess()
    bad.nC = 7
    assert L.cpab_transform_fwd(ctypes.byref(bad), 1, 1, 1, 10, 0, None, 1.0, 1, None, 1, 1 << 20, None) == 1
    # missing output
    assert L.cpab_transform_fwd(ctypes.byref(t), 1, 1, 1, 10, 0, None, 1.0, None, None, 1, 1 << 20, None) == 1
    # workspace too small
    need = L.cpab_workspace_bytes(ctypes.byref(t), 4, 1000)
    assert need > 0
    assert L.cpab_transform_fwd(ctypes.byref(t), 1, 4, 1, 1000, 0, None, 1.0, 1, None, 1, need - 1, None) == 3
    # empty problems are a no-op
    assert L.cpab_transform_fwd(ctypes.byref(t), 1, 0, 1, 1000, 0, None, 1.0, 1, None, None, 0, None) == 0
    # ... and their empty arrays may be NULL
    assert L.cpab_transform_fwd(ctypes.byref(t), 1, 2, None, 0, 0, None, 1.0, None, None, None, 0, None) == 0
    assert L.cpab_transform_fwd(ctypes.byref(t), None, 0, None, 0, 0, None, 1.0, None, None, None, 0, None) == 0
    assert L.cpab_findcell(2, (ctypes.c_int * 2)(3, 3), None, 0, None, None) == 0
    # a negative size is an error before any pointer is looked at
    assert L.cpab_transform_fwd(ctypes.byref(t), None, 2, None, -1, 0, None, 1.0, None, None, None, 0, None) == 1
    assert L.cpab_interpolate_fwd(5, 1, 1, 1, None, 1, None, 1, None) == 2
    dsz = (ctypes.c_int * 3)(4, 4, 1)
    osz = (ctypes.c_int * 3)(0, 4, 1)
    assert L.cpab_interpolate_fwd(2, 1, 1, 1, dsz, 1, osz, 1, None) == 1


def test_tables_beyond_the_lds_budget_are_an_error_before_any_launch():
    """A tessellation whose cell tables do not fit the 160 KB of LDS comes back
    as CPAB_ERR_LDS (5) from the argument checks, before anything is
    launched (the pointers below are never dereferenced)."""
    L = N.lib()
    t = N.CpabTess()
    t.ndim = 2
    t.nc = N.int3([40, 40, 1])
    t.nC = 4 * 40 * 40
    t.d = 10
    t.B = 1
    t.planes = 1
    t.nbr = 1
    need = L.cpab_workspace_bytes(ctypes.byref(t), 1, 100)
    assert L.cpab_transform_fwd(ctypes.byref(t), 1, 1, 1, 100, 0, None, 1.0, 1, None, 1, need, None) == 5
    assert b"LDS" in L.cpab_status_string(5) or b"lds" in L.cpab_status_string(5).lower()


def test_workspace_and_saved_sizes():
    L = N.lib()
    t = _tess("2d_4x4")
    D = 64 * 6
    ws = L.cpab_workspace_bytes(ctypes.byref(t), 256, 65536)
    assert ws >= 256 * D * 4
    assert L.cpab_saved_floats(2, 256, 65536) == 256 * (6 + 3) * 65536
    assert L.cpab_saved_floats(3, 16, 4096) == 16 * (12 + 3) * 4096
    assert L.cpab_saved_floats(1, 8, 1024) == 8 * 2 * 1024
    assert L.cpab_jac_floats(2, 256, 65536) == 256 * (4 + 3) * 65536
    assert L.cpab_jac_floats(1, 8, 1024) == 8 * 1024
    # transform_data: A_q plus one dL/dA partial per 2D tile (32 x 16 points in the backward)
    osz = (ctypes.c_int * 3)(256, 256, 1)
    wd = L.cpab_data_workspace_bytes(ctypes.byref(t), 256, osz)
    assert wd == 256 * D * 4 + 256 * (256 // 32) * (256 // 16) * D * 4
    # ragged grids still get whole tiles
    osz = (ctypes.c_int * 3)(100, 37, 1)
    # (ppt 1: 16 x 16 tiles, ceil(100 / 16) x ceil(37 / 16) of them)
    assert L.cpab_batch_workspace_bytes(ctypes.byref(t), osz, 1) >= 256 + ((100 + 15) // 16) * ((37 + 15) // 16) * D * 4


def test_cpab_facade_contract():
    from libcpab_amd import Cpab, CpabSequential
    T = Cpab([3, 3], backend="pytorch", device="gpu")
    assert T.get_theta_dim() == 34 and T.get_theta_size() == 34
    assert T.get_basis().shape == (216, 34)
    assert T.identity(3, epsilon=0.1).shape == (3, 34) if torch.cuda.is_available() else True
    with pytest.raises(AssertionError):
        Cpab([3, 3], backend="jax")
    with pytest.raises(AssertionError):
        Cpab([0, 3])
    with pytest.raises(NotImplementedError):
        Cpab([3, 3], backend="numpy", device="cpu")
    with pytest.raises(NotImplementedError):  # the reference's defaults: numpy on the CPU
        Cpab([3, 3])
    with pytest.raises(NotImplementedError):
        Cpab([3, 3], backend="pytorch", device="cpu")
    # tensors on the wrong device are rejected like in the reference
    with pytest.raises(AssertionError):
        T.transform_grid(torch.zeros(2, 10), torch.zeros(1, 34))
    T.set_solver_params(nstepsolver=20, method="numeric")
    assert T.params.nstepsolver == 20 and T.params.method == "numeric"
    with pytest.raises(AssertionError):
        T.set_solver_params(nstepsolver=0)
    S = CpabSequential(T, Cpab([4, 4], backend="pytorch", device="gpu"))
    assert S.get_theta_dim() == [34, 62]


def test_batch_plan_groups_instances_by_backward_lds_class():
    """cpab_batch_plan (host code) lays out each instance's work items
    consecutively and the instances class by class: those whose backward
    tables, accumulators and dL/ddata box fit 4 blocks per CU ([2,2] and a
    1D tessellation), then 3 ([6,6]: 144 simplices), then 2 ([8,8]: 256),
    each class in instance order.  The plan reads no device memory, so host
    stand-ins serve for the table pointers."""
    lib = N.lib()
    dummy = (ctypes.c_int * 4)()
    specs = [([8, 8], [100, 90]), ([2, 2], [64, 48]), ([6, 6], [40, 40]), ([10], [700]), ([8, 8], [30, 20]),
             ([2, 2], [33, 7])]
    insts = (N.CpabInstance * len(specs))()
    for I, (nc, osz) in zip(insts, specs):
        nd = len(nc)
        I.tess.ndim = nd
        I.tess.nc = N.int3(nc)
        I.tess.nC = int(np.prod(nc)) * (4 if nd == 2 else 1)
        I.tess.d = 1
        I.tess.planes = ctypes.addressof(dummy)
        I.tess.nbr = ctypes.addressof(dummy)
        I.outsize = N.int3(osz)
        I.nP = int(np.prod(osz))
    ptr = ctypes.cast(insts, ctypes.c_void_p)
    total = lib.cpab_batch_plan(ptr, len(specs), 2, None, 0)
    assert total > 0
    items = (ctypes.c_int * total)()
    assert lib.cpab_batch_plan(ptr, len(specs), 2, ctypes.cast(items, ctypes.c_void_p), total) == total
    items = np.frombuffer(items, dtype=np.int32)
    # every instance's items are one run at item0
    for i, I in enumerate(insts):
        assert I.tiles > 0 and (items[I.item0:I.item0 + I.tiles] == i).all()
    assert sum(I.tiles for I in insts) == total
    # runs in class order, instance order inside a class
    order = [int(items[k]) for k in range(total) if k == 0 or items[k] != items[k - 1]]
    assert order == [1, 3, 5, 2, 0, 4]


def test_meshgrid_order_matches_reference():
    from libcpab_amd.pytorch.functions import uniform_meshgrid
    import oracle as O
    g = uniform_meshgrid(2, [0, 0], [1, 1], [4, 3]).numpy()
    assert np.allclose(g, O.uniform_meshgrid([4, 3]))
    assert np.allclose(g[:, :4], [[0, 1 / 3, 2 / 3, 1], [0, 0, 0, 0]])


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 (or a multi-rank launch)
    exits with status 2 and says so, before touching a GPU."""
    import subprocess
    import sys
    for args, env in [(["--gpus", "2"], {}), (["--gpus", "1"], {"WORLD_SIZE": "4"})]:
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                           env=dict(os.environ, **env), timeout=120)
        assert r.returncode == 2
        assert "one GPU" in r.stderr
        assert r.stdout == ""


@pytest.mark.parametrize("nc,ls,ov", [([3, 3], 0.1, 1.0), ([4], 0.3, 2.0), ([2, 2, 2], 0.2, 0.5)])
def test_prior_covariance_equals_the_block_loop(nc, ls, ov):
    """prior_covariance (vectorised) equals the reference's nC x nC block
    loop (libcpab/cpab.py:211-234) written out here: squared cell-centre
    distances on each block's diagonal, 100 x the largest elsewhere, squared
    exponential kernel, B^T C B; symmetric and positive semi-definite."""
    from libcpab_amd import Cpab
    T = Cpab(nc, backend="pytorch", device="gpu")
    p = T.params
    c = np.asarray(T._centers, dtype=np.float64)
    dist = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    ppc = p.params_pr_cell
    C = np.zeros((p.D, p.D))
    for i in range(p.nC):
        for j in range(p.nC):
            block = np.full((ppc, ppc), 100 * dist.max())
            block[np.arange(ppc), np.arange(ppc)] = dist[i, j]
            C[ppc * i:ppc * (i + 1), ppc * j:ppc * (j + 1)] = block
    K = ov ** 2 * np.exp(-C / (2 * ls ** 2))
    want = p.basis.T @ K @ p.basis
    got = T.prior_covariance(ls, ov).numpy()
    assert got.shape == (p.d, p.d)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert np.allclose(got, got.T, atol=1e-12 * np.abs(got).max())
    assert np.linalg.eigvalsh(got).min() > -1e-9 * np.abs(got).max()
    assert np.allclose(T.prior_covariance(ls, 2 * ov).numpy(), 4 * got, rtol=1e-12)

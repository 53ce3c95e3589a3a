#!/usr/bin/env python3
"""Benchmark of the CPA transform_data path on one MI355X.

Workload (the 2D image configuration): Cpab([4, 4], zero_boundary=True),
256 thetas ~ N(0, 0.3^2 I), transform_data of 256 x 3 x 256 x 256 float32
U(0,1) images to a 256 x 256 output, MSE loss to a U(0,1) target, backward
w.r.t. theta and the images.  One step = forward + loss + backward of the
whole batch; a point is one (theta, output pixel) pair.

Prints one JSON line: the batched path's throughput (value), the plain
one-launch-per-instance path, the forward / gradient error against the FP64
oracle, the HBM roofline of the dominant kernel, and the reference's own CPU
op timed on a bounded sample on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+bwd points/s, 2D 4x4 tess: batched vs plain HIP; max err vs FP64 oracle"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(T, n_theta, C, size, device, seed=0):
    import torch
    g = torch.Generator().manual_seed(seed)
    theta = torch.randn(n_theta, T.get_theta_dim(), generator=g) * 0.3
    data = torch.rand((n_theta, C) + tuple(size), generator=g)
    target = torch.rand((n_theta, C) + tuple(size), generator=g)
    return theta.to(device), data.to(device), target.to(device)


def timed_loop(fn, steps, warmup):
    """Mean seconds per call of fn over `steps` calls bracketed by device
    synchronisation, and the median of the per-call durations from HIP events
    recorded between the calls on the current stream (no synchronisation
    inside the timed region)."""
    import statistics
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i].record()
        fn()
    ev[steps].record()
    torch.cuda.synchronize()
    mean = (time.perf_counter() - t0) / steps
    med = statistics.median(ev[i].elapsed_time(ev[i + 1]) for i in range(steps)) * 1e-3
    return mean, med


def time_steps(T, theta, data, target, size, steps, warmup):
    """(mean, median) seconds per fwd+bwd step of the image workload."""
    th = theta.detach().clone().requires_grad_(True)
    dd = data.detach().clone().requires_grad_(True)

    def step():
        import torch
        th.grad = None
        dd.grad = None
        out = T.transform_data(dd, th, size)
        loss = torch.nn.functional.mse_loss(out, target)
        loss.backward()

    return timed_loop(step, steps, warmup)


def time_bwd_kernel(T, theta, data, target, size, reps):
    """Average duration of the fused backward (k_transform_data_bwd with its
    two small companions: theta->A and the per-theta reduction), HIP events on
    the stream the kernels are launched on."""
    import torch
    from libcpab_amd import _native as N
    from libcpab_amd.pytorch import hip_ops as H
    dev = H.device_tess(T.params, theta.device)
    n_theta, C = data.shape[:2]
    nP = int(size[0] * size[1])
    lin = H.meshgrid_axes(size, theta.device)
    out = torch.empty_like(data)
    saved = torch.empty(int(N.lib().cpab_saved_floats(2, n_theta, nP)), device=theta.device)
    ws, nb = dev.workspace(n_theta, nP, size)
    stream = H._stream(theta.device)
    dsz, osz = H._int_arr(data.shape[2:]), H._int_arr(size)
    N.check(N.lib().cpab_transform_data_fwd(dev.ref, H._ptr(theta), n_theta, H._ptr(data), C, dsz, osz, H._ptr(lin),
                                            1.0, H._ptr(out), H._ptr(saved), H._ptr(ws), nb, stream))
    gout = (2.0 / out.numel()) * (out - target)
    gdata = torch.zeros_like(data)
    gth = torch.empty(n_theta, dev.d, device=theta.device)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    fwd_ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def bwd():
        N.check(N.lib().cpab_transform_data_bwd(dev.ref, H._ptr(theta), n_theta, H._ptr(data), C, dsz, osz,
                                                H._ptr(lin), 1.0, H._ptr(gout), H._ptr(saved), H._ptr(gdata),
                                                H._ptr(gth), H._ptr(ws), nb, stream))

    def fwd():
        N.check(N.lib().cpab_transform_data_fwd(dev.ref, H._ptr(theta), n_theta, H._ptr(data), C, dsz, osz,
                                                H._ptr(lin), 1.0, H._ptr(out), H._ptr(saved), H._ptr(ws), nb, stream))

    for _ in range(3):
        bwd()
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(reps):
        bwd()
    ev[1].record()
    fwd_ev[0].record()
    for _ in range(reps):
        fwd()
    fwd_ev[1].record()
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps, fwd_ev[0].elapsed_time(fwd_ev[1]) / reps


def cpu_baseline(T, theta, data, target, size, npts, nerr, seed=1):
    """The reference's own CPU op (libcpab/core/cpab_ops.cpp, compiled from its
    sources into oracle/_ref) on a bounded sample of the workload: theta[0],
    npts random output pixels; forward (nstepsolver = 50), interpolation
    forward/backward, the reference's RK2 backward and its reduction
    (libcpab/pytorch/transformer.py:181-202).  Timed single-threaded like the
    reference, and split over the host's threads.  Also measures the GPU path's error against the FP64 oracle on
    the same sample."""
    import numpy as np
    import torch
    import oracle as O
    p = T.get_params()
    rng = np.random.default_rng(seed)
    mesh = O.uniform_meshgrid(size)  # [2, nP], x fastest
    th0 = theta[0:1].detach().cpu().double().numpy()
    res = {}
    # --- GPU error against the FP64 oracle on nerr pixels (theta[0]) ---
    sel = np.sort(rng.choice(mesh.shape[1], size=nerr, replace=False))
    pts = mesh[:, sel]
    A64 = O.expand_A(p.basis, th0, 2)
    exact = O.cf_forward(2, p.nc, A64, p.facet_planes, p.facet_nbr, pts)
    g = torch.tensor(pts, dtype=torch.float32, device=theta.device)
    th = theta[0:1].detach().clone().requires_grad_(True)
    gpu = T.transform_grid(g, th)
    w = torch.tensor(rng.standard_normal((1, 2, nerr)), dtype=torch.float32, device=theta.device)
    (gpu * w).sum().backward()
    res["max_err_fwd"] = float(np.abs(gpu.detach().cpu().double().numpy() - exact).max())
    gA, _ = O.cf_backward(2, p.nc, A64, p.facet_planes, p.facet_nbr, pts, w.cpu().double().numpy())
    gth = gA.reshape(1, -1) @ p.basis
    res["max_rel_err_grad"] = float(np.abs(th.grad.cpu().double().numpy() - gth).max() / np.abs(gth).max())
    res["max_rel_err_grad_fd"] = fd_gradient_error(T, theta, pts)
    # --- reference CPU op on npts pixels ---
    if O.ref_lib() is None:
        return res, None
    sel = np.sort(rng.choice(mesh.shape[1], size=npts, replace=False))
    B32 = p.basis.astype(np.float32)
    th32 = theta[0:1].detach().cpu().numpy().astype(np.float32)
    d0 = data[0:1].detach().cpu().double().numpy()
    # target value of point p = ix + W*iy lives at target[0, c, ix, iy]
    tflat = target[0:1].detach().cpu().numpy().transpose(0, 1, 3, 2).reshape(1, 3, -1)
    A32 = (B32 @ th32.T).T.reshape(1, p.nC, 2, 3).astype(np.float32)
    trels = O.trels(A32.astype(np.float64), 50).astype(np.float32)
    Bs = O.bases_Bs(B32, p.nC, 2).astype(np.float32)

    def run(s_):
        pts32 = mesh[:, s_].astype(np.float32)
        n = pts32.shape[1]
        fwd = O.ref_forward(2, p.nc, 50, pts32, trels)
        # interpolation at the warped sample points (values and d/dgrid)
        outv = O.interpolate(2, d0, fwd.astype(np.float64), [n, 1])
        go = 2.0 * (outv - tflat[:, :, s_].reshape(outv.shape)) / (data.numel())
        _, ggrid = O.interpolate_backward(2, d0, fwd.astype(np.float64), [n, 1], go)
        jac = O.ref_backward(2, p.nc, 50, pts32, A32, Bs)  # [d, 1, 2, n]
        return (jac * ggrid[None].astype(np.float32)).sum(axis=(2, 3)).T

    t0 = time.perf_counter()
    run(sel)
    dt = time.perf_counter() - t0
    # the same sample split over the host's threads (the compiled op releases
    # the GIL); at most 16, the host share of one GPU
    from concurrent.futures import ThreadPoolExecutor
    affinity = len(os.sched_getaffinity(0))
    nthr = max(1, min(16, affinity))
    chunks = np.array_split(sel, nthr * 4)
    with ThreadPoolExecutor(nthr) as ex:
        t0 = time.perf_counter()
        list(ex.map(run, chunks))
        dt_all = time.perf_counter() - t0
    return res, {"value": npts / dt, "unit": "points/s", "cores": 1, "kind": "reference",
                 "sample": "theta[0] x %d random output pixels of the 256x256 image workload: the reference's "
                           "compiled cpab_forward_op (nstepsolver=50) + interpolation fwd/bwd + cpab_backward_op "
                           "(RK2, d=%d) + reduction; %.1f s on one thread" % (npts, p.d, dt),
                 "all_cores": {"value": npts / dt_all, "cores": nthr, "seconds": dt_all},
                 "affinity_cpus": affinity, "torch_threads": torch.get_num_threads()}


def numpy_baseline(T, theta, data, size, npts, seed=1, reps=5):
    """The reference's numpy backend -- FP64 fixed-step integration
    (nstepsolver = 50, scipy expm per cell) and NWHC linear interpolation,
    libcpab/numpy/transformer.py:22-61 and libcpab/numpy/interpolation.py --
    restated in vectorised numpy (oracle/numpy_path.py), timed on one core:
    (a) the plain [3,3] workload whole (one theta ~ N(0, 0.5^2 I), seed 0,
    100x100 meshgrid, one 100x100x3 image), median of `reps` runs; (b)
    theta[0] x npts random pixels of the image workload.  Forward only: the
    numpy backend has no gradient."""
    import statistics
    import numpy as np
    import torch
    import oracle as O
    from oracle import numpy_path as NP
    from libcpab_amd import Cpab
    T3 = Cpab([3, 3], backend="pytorch", device="gpu", zero_boundary=True)
    g = torch.Generator().manual_seed(0)
    th3 = (torch.randn(1, T3.get_theta_dim(), generator=g) * 0.5).double().numpy()
    img3 = np.moveaxis(torch.rand(1, 3, 100, 100, generator=g).double().numpy(), 1, -1)  # NWHC
    mesh3 = O.uniform_meshgrid([100, 100])
    B3 = T3.get_params().basis

    def plain():
        grid = NP.transformer(mesh3, th3, B3, [3, 3], 50)
        return NP.interpolate(2, img3, grid, [100, 100])

    plain()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plain()
        times.append(time.perf_counter() - t0)
    t_plain = statistics.median(times)
    p = T.get_params()
    rng = np.random.default_rng(seed)
    mesh = O.uniform_meshgrid(size)
    sel = np.sort(rng.choice(mesh.shape[1], size=npts, replace=False))
    th0 = theta[0:1].detach().cpu().double().numpy()
    d0 = np.moveaxis(data[0:1].detach().cpu().double().numpy(), 1, -1)
    t0 = time.perf_counter()
    grid = NP.transformer(mesh[:, sel], th0, p.basis, p.nc, 50)
    NP.interpolate(2, d0, grid, [npts, 1])
    t_img = time.perf_counter() - t0
    return {"value": npts / t_img, "unit": "points/s (forward only)", "cores": 1, "kind": "port",
            "sample": "the reference's numpy backend restated in numpy (FP64, nstepsolver=50, scipy expm per cell, "
                      "NWHC interpolation; forward only -- that backend has no gradient): theta[0] x %d random "
                      "output pixels of the 256x256 image workload, %.2f s on one thread" % (npts, t_img),
            "plain_3x3_100x100": {"value": 1e4 / t_plain, "unit": "points/s", "seconds": t_plain,
                                  "workload": "Cpab([3,3]), 1 theta ~ N(0, 0.5^2 I) seed 0, 100x100 meshgrid, "
                                              "1x100x100x3 U(0,1) image, transform_grid + interpolate"}}


def oracle_error(T, theta, pts, seed=2):
    """Max |GPU - FP64 oracle| of transform_grid at pts [ndim, n] for theta[0:1],
    and the max gradient error relative to the gradient's magnitude."""
    import numpy as np
    import torch
    import oracle as O
    p = T.get_params()
    nd = p.ndim
    rng = np.random.default_rng(seed)
    A64 = O.expand_A(p.basis, theta[0:1].detach().cpu().double().numpy(), nd)
    exact = O.cf_forward(nd, p.nc, A64, p.facet_planes, p.facet_nbr, pts)
    th = theta[0:1].detach().clone().requires_grad_(True)
    gpu = T.transform_grid(torch.tensor(pts, dtype=torch.float32, device=theta.device), th)
    w = rng.standard_normal((1, nd, pts.shape[1]))
    (gpu * torch.tensor(w, dtype=torch.float32, device=theta.device)).sum().backward()
    gA, _ = O.cf_backward(nd, p.nc, A64, p.facet_planes, p.facet_nbr, pts, w)
    gth = gA.reshape(1, -1) @ p.basis
    return (float(np.abs(gpu.detach().cpu().double().numpy() - exact).max()),
            float(np.abs(th.grad.cpu().double().numpy() - gth).max() / np.abs(gth).max()))


def fd_gradient_error(T, theta, pts, n=1024, eps=1e-5, seed=3):
    """Max |dL/dtheta (GPU) - FP64 central differences| / max |FD| for
    L = <w, T_theta(pts[:, :n])>, random w, theta[0]: the differences are of
    the oracle's exact FP64 flow, one pair per theta direction."""
    import numpy as np
    import torch
    import oracle as O
    p = T.get_params()
    nd = p.ndim
    pts = pts[:, :n]
    w = np.random.default_rng(seed).standard_normal((1, nd, pts.shape[1]))
    th = theta[0:1].detach().clone().requires_grad_(True)
    gpu = T.transform_grid(torch.tensor(pts, dtype=torch.float32, device=theta.device), th)
    (gpu * torch.tensor(w, dtype=torch.float32, device=theta.device)).sum().backward()
    th0 = theta[0:1].detach().cpu().double().numpy()
    fd = np.empty(th0.shape[1])
    for k in range(th0.shape[1]):
        e = np.zeros_like(th0)
        e[0, k] = eps
        fp = O.cf_forward(nd, p.nc, O.expand_A(p.basis, th0 + e, nd), p.facet_planes, p.facet_nbr, pts)
        fm = O.cf_forward(nd, p.nc, O.expand_A(p.basis, th0 - e, nd), p.facet_planes, p.facet_nbr, pts)
        fd[k] = ((fp - fm) * w).sum() / (2 * eps)
    return float(np.abs(th.grad.cpu().double().numpy()[0] - fd).max() / np.abs(fd).max())


def config_inputs(name, device, seed=0):
    """The 1D time-series, 3D volume and 4-stage sequential configurations:
    (transforms, thetas, data, target, outsize, data needs grad, description)."""
    import math
    import torch
    from libcpab_amd import Cpab
    g = torch.Generator().manual_seed(seed)
    if name == "ts1d":
        T = Cpab([50], backend="pytorch", device="gpu", zero_boundary=True)
        n, L = 4096, 1024
        theta = torch.randn(n, T.get_theta_dim(), generator=g) * 0.5
        x = torch.linspace(0, 1, L)
        f = torch.rand(n, 3, 1, generator=g) * 10 + 1
        ph = torch.rand(n, 3, 1, generator=g) * 2 * math.pi
        amp = torch.rand(n, 3, 1, generator=g)
        data = (amp * torch.sin(2 * math.pi * f * x + ph)).sum(1, keepdim=True)
        data = data + 0.05 * torch.randn(n, 1, L, generator=g)
        target = torch.rand(n, 1, L, generator=g)
        desc = ("ts1d: Cpab([50], zero_boundary=True), 4096 thetas ~ N(0, 0.5^2 I), 1024-sample signals "
                "(3 random sinusoids + N(0, 0.05) noise), MSE to a U(0,1) target, fwd+bwd w.r.t. theta")
        return [T], [theta.to(device)], data.to(device), target.to(device), [L], False, desc
    if name == "vol3d":
        T = Cpab([3, 3, 3], backend="pytorch", device="gpu", zero_boundary=True)
        n, S = 16, 64
        theta = torch.randn(n, T.get_theta_dim(), generator=g) * 0.2
        ax = torch.linspace(0, 1, S)
        zz, yy, xx = torch.meshgrid(ax, ax, ax, indexing="ij")
        data = torch.zeros(n, 1, S, S, S)
        for i in range(n):
            c = torch.rand(8, 3, generator=g)
            s = torch.rand(8, generator=g) * 0.15 + 0.05
            a = torch.rand(8, generator=g) * 0.5 + 0.5
            for k in range(8):
                r2 = (xx - c[k, 0]) ** 2 + (yy - c[k, 1]) ** 2 + (zz - c[k, 2]) ** 2
                data[i, 0] += a[k] * torch.exp(-r2 / (2 * s[k] ** 2))
        target = torch.rand(n, 1, S, S, S, generator=g)
        desc = ("vol3d: Cpab([3,3,3], zero_boundary=True), 16 thetas ~ N(0, 0.2^2 I), 16x1x64x64x64 volumes "
                "(8 random Gaussian blobs each), MSE to a U(0,1) target, fwd+bwd w.r.t. theta")
        return [T], [theta.to(device)], data.to(device), target.to(device), [S, S, S], False, desc
    if name == "seq2d":
        Ts = [Cpab([4, 4], backend="pytorch", device="gpu", zero_boundary=True) for _ in range(4)]
        n = 64
        thetas = [(torch.randn(n, Ts[0].get_theta_dim(), generator=g) * 0.3).to(device) for _ in range(4)]
        data = torch.rand(n, 3, 256, 256, generator=g)
        target = torch.rand(n, 3, 256, 256, generator=g)
        desc = ("seq2d: CpabSequential of 4 Cpab([4,4], zero_boundary=True), 64 thetas each ~ N(0, 0.3^2 I), "
                "64x3x256x256 U(0,1) images, MSE to a U(0,1) target, fwd+bwd w.r.t. all thetas with "
                "trajectory recomputation")
        return Ts, thetas, data.to(device), target.to(device), [256, 256], False, desc
    raise ValueError(name)


def config_measure(args, workload):
    """One of the secondary configurations: batched (one launch over all
    thetas) and plain (one launch per theta) fwd+bwd points/s, and the error
    against the FP64 oracle on a sample of the workload's grid.  Returns the
    JSON line as a dict."""
    import numpy as np
    import torch
    import oracle as O
    from libcpab_amd import CpabSequential
    dev = torch.device("cuda:0")
    Ts, thetas, data, target, size, data_grad, desc = config_inputs(workload, dev)
    n_theta = thetas[0].shape[0]
    points = n_theta * int(np.prod(size))
    seq = len(Ts) > 1
    S = CpabSequential(*Ts) if seq else None
    log("workload ready: %s" % desc)

    def step():
        ths = [t.detach().clone().requires_grad_(True) for t in thetas]
        dd = data.detach().clone().requires_grad_(data_grad)
        out = S.transform_data(dd, ths, size) if seq else Ts[0].transform_data(dd, ths[0], size)
        loss = torch.nn.functional.mse_loss(out, target)
        loss.backward()
        return ths

    def timed(steps, warmup):
        return timed_loop(step, steps, warmup)

    sec, med = timed(args.steps, args.warmup)
    log("batched: %.3f ms/step, %.3e points/s" % (sec * 1e3, points / sec))
    plain = None
    if not args.no_plain and not seq:
        Ts[0].set_solver_params(launch="plain")
        psec, _ = timed(args.plain_steps, 1)
        Ts[0].set_solver_params(launch="batched")
        plain = {"value": points / psec, "ms_per_step": psec * 1e3}
        log("plain: %.3f ms/step" % (psec * 1e3))
    mesh = O.uniform_meshgrid(size)
    rng = np.random.default_rng(1)
    pts = mesh[:, np.sort(rng.choice(mesh.shape[1], size=min(args.err_points, mesh.shape[1]), replace=False))]
    errs = [oracle_error(T, th, pts) for T, th in zip(Ts, thetas)]
    fd_errs = [fd_gradient_error(T, th, pts) for T, th in zip(Ts, thetas)]
    return {
        "metric": "fwd+bwd points/s (%s); max err vs FP64 oracle" % workload,
        "value": points / sec, "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "ms_per_step_median": med * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "n_theta": n_theta, "outsize": size, "points_per_step": points,
                   "stages": len(Ts)},
        "plain": plain,
        "batched_over_plain": (points / sec / plain["value"]) if plain else None,
        "max_err_fwd": max(e[0] for e in errs),
        "max_rel_err_grad": max(e[1] for e in errs),
        "max_rel_err_grad_fd": max(fd_errs),
        "err_sample": "theta[0] of each stage, %d grid points (finite differences: the first %d)" % (
            pts.shape[1], min(1024, pts.shape[1])),
    }


def run_config(args):
    print(json.dumps(config_measure(args, args.workload)), flush=True)


def run_copy(args):
    """Device-to-device copy of a 1 GiB buffer: the reference kernel of known
    traffic (1 GiB read + 1 GiB written per launch) that scales the FETCH_SIZE /
    WRITE_SIZE counters into bytes."""
    import torch
    dev = torch.device("cuda:0")
    a = torch.ones(256 * 1024 * 1024, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"metric": "1 GiB device copy", "value": 2 * a.numel() * 4 / sec / 1e9, "unit": "GB/s",
                      "ms_per_step": sec * 1e3, "steps": args.steps}), flush=True)


def run_align(args):
    """CpabAligner.alignment_by_gradient on one 3x256x256 image pair ([4,4]
    tessellation, a known warp of smooth synthetic content): Adam iterations
    per second run eagerly (torch.norm loss, no per-iteration host read) and as
    one replayed HIP graph per iteration."""
    import torch
    from libcpab_amd import Cpab, CpabAligner

    dev = torch.device("cuda:0")
    T = Cpab([4, 4], backend="pytorch", device="gpu", zero_boundary=True)
    g = torch.Generator().manual_seed(0)
    y, x = torch.meshgrid(torch.linspace(0, 1, 256), torch.linspace(0, 1, 256), indexing="ij")
    fr = torch.rand(3, 3, 2, generator=g) * 2.0 + 0.5
    x1 = (0.5 + 0.5 * torch.sin(6.2832 * (fr[..., 0, None, None] * x + fr[..., 1, None, None] * y)).mean(1))[None]
    x1 = x1.to(dev)
    theta_true = (torch.randn(1, T.get_theta_dim(), generator=g) * 0.3).to(dev)
    x2 = T.transform_data(x1, theta_true, [256, 256]).detach()
    A = CpabAligner(T)
    iters = max(args.steps, 20)
    res = {}
    for graph in (False, True):
        A.alignment_by_gradient(x1, x2, maxiter=5, graph=graph, progress=False)  # set-up, caches
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        theta = A.alignment_by_gradient(x1, x2, maxiter=iters, graph=graph, progress=False)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / iters
        loss = torch.norm(T.transform_data(x1, theta.detach(), [256, 256]) - x2).item()
        res["graph" if graph else "eager"] = {"ms_per_iter": sec * 1e3, "final_misfit": loss}
        log("%s: %.3f ms/iteration, misfit %.4f" % ("graph" if graph else "eager", sec * 1e3, loss))
    line = {"metric": "alignment_by_gradient iterations/s (1x3x256x256, [4,4])",
            "value": 1e3 / res["graph"]["ms_per_iter"], "unit": "iterations/s", "n_gpus": 1, "steps": iters,
            "higher_is_better": True, "data": "synthetic", "config": {"workload": "align", "tess": [4, 4],
                                                                      "image": [1, 3, 256, 256]},
            "eager": res["eager"], "graph": res["graph"],
            "graph_over_eager": res["eager"]["ms_per_iter"] / res["graph"]["ms_per_iter"],
            "initial_misfit": torch.norm(x1 - x2).item()}
    print(json.dumps(line), flush=True)


def ragged_problems(n, seed=0):
    """The ragged mix: n instances, half 1D (10-100 cells, 256-8192 points,
    1 channel), half 2D ([2..8] x [2..8] cells, 64^2-512^2 pixels drawn
    log-uniformly, 3 channels); output size = data size."""
    import numpy as np
    from libcpab_amd import Cpab
    rng = np.random.default_rng(seed)
    cache = {}
    probs = []
    for _ in range(n):
        if rng.random() < 0.5:
            nc = [int(rng.integers(10, 101))]
            size = [int(round(np.exp(rng.uniform(np.log(256), np.log(8192)))))]
            C = 1
        else:
            nc = [int(rng.integers(2, 9)), int(rng.integers(2, 9))]
            side = int(round(np.exp(rng.uniform(np.log(64), np.log(512)))))
            size = [side, side]
            C = 3
        key = tuple(nc)
        if key not in cache:
            cache[key] = Cpab(nc, backend="pytorch", device="gpu", zero_boundary=True)
        probs.append((cache[key], C, size, size))
    return probs


def ragged_oracle_errors(b, probs):
    """Forward / gradient error of the graph-replayed batch against the FP64
    oracle on two whole instances: the smallest [8,8] one (tables that fill
    the LDS) and the largest 1D one.  The forward is read from the saved
    warped points; dL/dtheta is compared with the oracle's exact adjoint of
    the integration applied to dL/dgrid of the same grad_out, taken by the
    FP64 interpolation backward at the GPU's warped points with the sample
    coordinates formed in FP32 like the GPU's (the interpolant of a noise
    image changes slope at every pixel edge)."""
    import numpy as np
    import oracle as O
    npts = [int(np.prod(o)) for _, _, _, o in probs]
    big = [i for i, (T, _, _, _) in enumerate(probs) if T.params.nc == [8, 8]]
    one = [i for i, (T, _, _, _) in enumerate(probs) if T.params.ndim == 1]
    pick = [min(big, key=lambda i: npts[i])] if big else []
    pick += [max(one, key=lambda i: npts[i])] if one else []
    ef, eg = 0.0, 0.0
    for i in pick:
        T, C, size, osz = probs[i]
        p = T.get_params()
        nd = p.ndim
        A = O.expand_A(p.basis, b.theta_views[i].detach().double().cpu().numpy()[None], nd)
        grid = O.uniform_meshgrid(osz)
        exact = O.cf_forward(nd, p.nc, A, p.facet_planes, p.facet_nbr, grid)[0]
        exact_t = exact.reshape((nd,) + tuple(osz[::-1])).transpose((0,) + tuple(range(nd, 0, -1))).reshape(nd, -1)
        warped = b.saved_views[i][:nd].detach().double().cpu().numpy()
        ef = max(ef, float(np.abs(warped - exact_t).max()))
        data = b.data_views[i].detach().double().cpu().numpy()[None]
        gout = b.grad_out_views[i].detach().double().cpu().numpy()[None]
        w_pts = warped.reshape((nd,) + tuple(osz)).transpose((0,) + tuple(range(nd, 0, -1))).reshape(1, nd, -1)
        _, ggrid = O.interpolate_backward(nd, data, w_pts, osz, gout, index32=True)
        gA, _ = O.cf_backward(nd, p.nc, A, p.facet_planes, p.facet_nbr, grid, ggrid)
        gth = (gA.reshape(1, -1) @ p.basis)[0]
        got = b.grad_theta_views[i].detach().double().cpu().numpy()
        eg = max(eg, float(np.abs(got - gth).max() / np.abs(gth).max()))
    return ef, eg, ["%d: %s %s" % (i, probs[i][0].params.nc, probs[i][3]) for i in pick]


def ragged_measure(args):
    import torch
    from libcpab_amd.batch import CpabBatch
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    probs = ragged_problems(args.instances)
    b = CpabBatch(probs, device=dev, save=True)
    g = torch.Generator().manual_seed(0)
    target = torch.zeros_like(b.out)
    for i, (T, C, size, out) in enumerate(probs):
        b.theta_views[i].copy_(torch.randn(T.get_theta_dim(), generator=g) * 0.3)
        b.data_views[i].copy_(torch.rand([C] + size, generator=g))
        off = (b.out_views[i].data_ptr() - b.out.data_ptr()) // 4
        target[off:off + b.out_views[i].numel()].copy_(torch.rand(b.out_views[i].numel(), generator=g))
    n_used = float(sum(b.used["out"]))
    log("ragged batch ready: %d instances, %d points (%.1f s)" % (b.n, b.points, time.perf_counter() - t0))
    best = b.autotune(reps=1)
    log("autotune: ppt=%d blocks=%d (%.3f ms)" % (best[1], best[2], best[0] * 1e3))
    b.capture()

    def step_batched():
        b.replay_forward()
        loss = ((b.out - target) ** 2).sum() / n_used
        b.grad_out.copy_((b.out - target) * (2.0 / n_used))
        b.replay_backward()
        return loss

    def step_plain():
        b.forward_plain()
        loss = ((b.out - target) ** 2).sum() / n_used
        b.grad_out.copy_((b.out - target) * (2.0 / n_used))
        b.backward_plain()
        return loss

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    sec = timed(step_batched, args.steps, args.warmup)
    psec = timed(step_plain, max(1, args.plain_steps), 1)
    value = b.points / sec
    log("ragged batched %.3f ms/step, plain %.3f ms/step" % (sec * 1e3, psec * 1e3))
    # errors of the graph-replayed batch: one more replayed step, then the oracle
    step_batched()
    torch.cuda.synchronize()
    err_fwd, err_grad, err_inst = ragged_oracle_errors(b, probs)
    log("ragged errors vs FP64 oracle: fwd %.3g, grad %.3g (%s)" % (err_fwd, err_grad, err_inst))
    return {
        "metric": "fwd+bwd points/s, ragged 1D/2D mix: batched (persistent, graph-captured) vs plain HIP",
        "value": value, "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "ragged: %d instances, 50/50 1D (10-100 cells, 256-8192 points, C=1) / 2D "
                               "([2..8]^2 cells, 64^2-512^2 log-uniform, C=3), thetas ~ N(0, 0.3^2 I), MSE, "
                               "fwd+bwd w.r.t. theta and data" % b.n,
                   "instances": b.n, "points_per_step": b.points, "ppt": b.ppt, "blocks": b.blocks},
        "plain": {"value": b.points / psec, "ms_per_step": psec * 1e3},
        "batched_over_plain": psec / sec,
        "max_err_fwd": err_fwd,
        "max_rel_err_grad": err_grad,
        "err_sample": "whole instances %s: warped points of the replayed forward and dL/dtheta of the replayed "
                      "backward against the FP64 oracle" % ", ".join(err_inst),
    }


def run_ragged(args):
    print(json.dumps(ragged_measure(args)), flush=True)


BWD_KERNEL = "k_transform_data_bwd<2, true>"
FWD_KERNEL = "k_transform_data<2, true>"
TRAFFIC_FILE = "profiles/images2d_k_transform_data.traffic.json"


VALU_FILE = "profiles/images2d_k_transform_data.valu.json"
# wave64 VALU instructions per second: 256 CUs x 4 SIMDs x 2.4 GHz, 4 cycles per instruction
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def _profiled(path, kernel):
    try:
        with open(os.path.join(ROOT, path)) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the
    1 GiB copy calibration) from the committed rocprofv3 counter summary of
    this workload, or None."""
    return _profiled(TRAFFIC_FILE, kernel)


def valu_roofline(kernel, launch_ms):
    """Issue-rate roofline of `kernel`, the bound the fused kernels actually
    meet: its VALU wave-instructions per launch (SQ_INSTS_VALU from the
    committed counter summary of this workload) over the launch time measured
    in this run, against the nominal issue peak; with the profiled fraction of
    lanes active per VALU instruction (SQ_THREAD_CYCLES_VALU / (64
    SQ_ACTIVE_INST_VALU)), the lane fraction (issue x lanes), and SALU per VALU
    instruction."""
    prof = _profiled(VALU_FILE, kernel)
    if prof is None:
        return None
    n = prof["valu_instr"]
    rate = n / (launch_ms * 1e-3)
    frac = rate / VALU_ISSUE_PEAK
    return {"kernel": kernel.split("<")[0], "bound": "valu-issue", "achieved": rate / 1e9,
            "peak": VALU_ISSUE_PEAK / 1e9, "unit": "G wave-instr/s", "frac": frac,
            "lanes_active": prof["lanes_active"], "lane_frac": frac * prof["lanes_active"],
            "salu_per_valu": prof["salu_per_valu"], "instr_per_launch": n, "instr_source": VALU_FILE,
            "launch_ms": launch_ms}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--cpu-points", type=int, default=40960)
    ap.add_argument("--err-points", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--workload", default="images2d", choices=["images2d", "ragged", "ts1d", "vol3d", "seq2d", "copy",
                                                                  "align"])
    ap.add_argument("--instances", type=int, default=1000)
    ap.add_argument("--no-extra", action="store_true",
                    help="images2d only: skip the ragged and seq2d sub-objects of the line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this path has no multi-GPU mode (the theta batch would be split "
              "across GPUs with no communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ragged":
        return run_ragged(args)
    if args.workload == "copy":
        return run_copy(args)
    if args.workload == "align":
        return run_align(args)
    if args.workload != "images2d":
        return run_config(args)

    import torch
    from libcpab_amd import Cpab

    dev = torch.device("cuda:0")
    n_theta, C, size = 256, 3, [256, 256]
    T = Cpab([4, 4], backend="pytorch", device="gpu", zero_boundary=True, volume_perservation=False)
    theta, data, target = make_inputs(T, n_theta, C, size, dev)
    points = n_theta * size[0] * size[1]
    log("workload ready: %d thetas x %d points" % (n_theta, size[0] * size[1]))

    sec, med = time_steps(T, theta, data, target, size, args.steps, args.warmup)
    value = points / sec
    log("batched: %.3f ms/step, %.3e points/s" % (sec * 1e3, value))

    plain = None
    if not args.no_plain:
        T.set_solver_params(launch="plain")
        psec, _ = time_steps(T, theta, data, target, size, args.plain_steps, 1)
        T.set_solver_params(launch="batched")
        plain = {"value": points / psec, "ms_per_step": psec * 1e3}
        log("plain: %.3f ms/step, %.3e points/s" % (psec * 1e3, points / psec))

    bwd_ms, fwd_ms = time_bwd_kernel(T, theta, data, target, size, args.kernel_reps)
    algo_bytes = 12 * C * points  # grad_out read + data read + grad_data write, 4 B each, per point
    achieved = algo_bytes / (bwd_ms * 1e-3) / 1e9
    log("k_transform_data_bwd call: %.3f ms; forward call %.3f ms" % (bwd_ms, fwd_ms))

    res, cpu, cpu_np = {}, None, None
    if not args.no_cpu_baseline:
        res, cpu = cpu_baseline(T, theta, data, target, size, args.cpu_points, args.err_points)
        log("cpu baseline:", cpu, res)
        cpu_np = numpy_baseline(T, theta, data, size, args.cpu_points)
        log("numpy baseline:", cpu_np)

    extra = {}
    if not args.no_extra:
        keys = ("ms_per_step", "value", "unit", "config", "plain", "batched_over_plain", "max_err_fwd",
                "max_rel_err_grad", "max_rel_err_grad_fd", "err_sample")
        for name in ("ragged", "seq2d"):
            r = ragged_measure(args) if name == "ragged" else config_measure(args, name)
            extra[name] = {k: r[k] for k in keys if k in r}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "ms_per_step_median": med * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "images2d: Cpab([4,4], zero_boundary=True), 256 thetas ~ N(0, 0.3^2 I), "
                               "transform_data of 256x3x256x256 U(0,1) images, MSE to a U(0,1) target, "
                               "fwd+bwd w.r.t. theta and images",
                   "tess": [4, 4], "n_theta": n_theta, "channels": C, "outsize": size, "points_per_step": points},
        "plain": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "max_err_fwd": res.get("max_err_fwd"),
        "max_rel_err_grad": res.get("max_rel_err_grad"),
        "max_rel_err_grad_fd": res.get("max_rel_err_grad_fd"),
        "roofline": {"kernel": "k_transform_data_bwd", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(BWD_KERNEL),
                     "traffic_source": TRAFFIC_FILE,
                     "algorithmic_bytes_per_launch": algo_bytes, "launch_ms": bwd_ms},
        "valu": valu_roofline(BWD_KERNEL, bwd_ms),
        "valu_fwd": valu_roofline(FWD_KERNEL, fwd_ms),
        "cpu_baseline": cpu,
        "cpu_baseline_numpy": cpu_np,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

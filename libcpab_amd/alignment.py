"""Alignment of two data tensors by the theta of a CPA transformation
(libcpab/alignment.py:12-106): a greedy random search over theta and Adam on
the gradient of |T_theta(x1) - x2|, for the pytorch backend on the GPU.

``alignment_by_gradient(..., graph=True)`` records one iteration -- the fused
transform_data forward and backward, the loss, Adam's update -- as a HIP graph
and replays it, so an iteration costs one graph launch instead of ~20 kernel
launches with their host work.  For the graph the loss is reduced in two
stages ([rows, 1024] -> [rows] -> scalar), each a single-block reduction per
output: a multi-block reduction of the whole tensor may record a memset of
its semaphores, a node that does not replay reliably.  Adam then runs with
capturable=True (its step count stays on the device).
"""
import math

import torch

from .cpab import Cpab
from .pytorch import functions as backend

_ROW = 1024


def _norm_two_stage(x):
    """|x|_2 as sqrt(sum(sum(rows^2))) over zero-padded rows of 1024."""
    flat = x.reshape(-1)
    rows = max(1, math.ceil(flat.numel() / _ROW))
    pad = rows * _ROW - flat.numel()
    if pad:
        flat = torch.nn.functional.pad(flat, (0, pad))
    return torch.sqrt((flat.view(rows, _ROW) ** 2).sum(1).sum())


class _AdamLoop:
    """Adam on theta [1, d] for the loss |transform_data(x1, theta) - x2|;
    iterations run eagerly (step) or as a captured graph (capture, replay)."""

    def __init__(self, T, x1, x2, lr, graph_safe):
        self.T, self.x1, self.x2 = T, x1, x2
        self.outsize = tuple(x1.shape[2:])
        self.norm = _norm_two_stage if graph_safe else backend.norm
        self.theta = T.identity(1, epsilon=1e-6).requires_grad_(True)
        self.opt = torch.optim.Adam([self.theta], lr=lr, capturable=graph_safe)
        self.loss = None
        self.graph = None

    def _iteration(self):
        out = self.T.transform_data(self.x1, self.theta, self.outsize)
        loss = self.norm(out - self.x2)
        loss.backward()
        self.opt.step()
        return loss.detach()

    def step(self):
        self.opt.zero_grad(set_to_none=True)
        self.loss = self._iteration()

    def capture(self):
        """Record one iteration.  The gradient is unset before the capture, so
        the recorded backward writes (not accumulates) theta.grad."""
        self.opt.zero_grad(set_to_none=True)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.loss = self._iteration()

    def replay(self):
        self.graph.replay()


class CpabAligner(object):
    """Aligns x1 to x2 by a CPA transformation of the given Cpab instance
    (libcpab/alignment.py:12-30)."""

    def __init__(self, cpab_class):
        assert isinstance(cpab_class, Cpab), "The input class needs to be an instance of the core cpab class"
        self.T = cpab_class
        self.backend = backend

    def _check_pair(self, x1, x2):
        self.T._check_type(x1)
        self.T._check_type(x2)
        self.T._check_device(x1)
        self.T._check_device(x2)
        assert x1.shape == x2.shape, " Two data points does not have the same shape "

    def alignment_by_sampling(self, x1, x2, maxiter=100, progress=True):
        """Greedy random search (libcpab/alignment.py:31-56): propose
        theta = 0.1 * N(current, I), keep it when |T_theta(x1) - x2| drops."""
        self._check_pair(x1, x2)
        outsize = tuple(x2.shape[2:])
        current_sample = self.T.identity(1)
        current_error = self.backend.norm(x1 - x2)
        accept = 0
        it = range(maxiter)
        if progress:
            from tqdm import tqdm
            it = tqdm(it, desc="Alignment of samples", unit="samples")
        for _ in it:
            theta = 1e-1 * self.T.sample_transformation(1, mean=current_sample.flatten())
            x1_trans = self.T.transform_data(x1, theta, outsize=outsize)
            new_error = self.backend.norm(x1_trans - x2)
            if new_error < current_error:
                current_sample = theta
                current_error = new_error
                accept += 1
        if progress:
            print("Acceptence ratio: ", accept / maxiter * 100, "%")
        return current_sample

    def alignment_by_gradient(self, x1, x2, maxiter=100, lr=1e-2, graph=False, progress=True, warmup=3):
        """Adam from the identity (libcpab/alignment.py:59-85).  graph=True
        runs the first `warmup` iterations eagerly on a side stream (lazy
        set-up stays out of the capture), records one iteration and replays
        it for the rest; the loss is then read once, at the end."""
        self._check_pair(x1, x2)
        loop = _AdamLoop(self.T, x1, x2, lr, graph_safe=graph)
        pb = None
        if progress:
            from tqdm import tqdm
            pb = tqdm(desc="Alignment of samples", unit="iters", total=maxiter)
        if not graph:
            for _ in range(maxiter):
                loop.step()
                if pb is not None:
                    pb.update()
                    pb.set_postfix({"loss": loop.loss.item()})
        else:
            n_eager = min(warmup, maxiter)
            side = torch.cuda.Stream(device=x1.device)
            side.wait_stream(torch.cuda.current_stream(x1.device))
            with torch.cuda.stream(side):
                for _ in range(n_eager):
                    loop.step()
            torch.cuda.current_stream(x1.device).wait_stream(side)
            if maxiter > n_eager:
                loop.capture()
                for _ in range(maxiter - n_eager):
                    loop.replay()
            if pb is not None:
                pb.update(maxiter)
                pb.set_postfix({"loss": loop.loss.item()})
        if pb is not None:
            pb.close()
        return loop.theta

"""Ragged batches of transform_data problems in one persistent launch.

A batch holds problems that may differ in dimension, tessellation, theta
size, channels, data size and output size (e.g. 1D signals with 10-100
cells next to 2D images on [2,2]..[8,8] grids).  All inputs and outputs
live in flat device buffers owned by the batch, every problem is split into
tiles of 256*ppt points, and one persistent kernel drains the (problem, tile)
work queue; a second one does the backward and a per-problem kernel reduces
its tile partials (fixed order) into dL/dtheta.  Points-per-thread and the
persistent grid size come from a small autotuner that times the candidates
on the batch itself; the forward and backward launch sequences can be
captured into HIP graphs and replayed.

No counterpart in the reference, whose pytorch op handles one tessellation
and one grid size per call (libcpab/pytorch/transformer.py:136-202).
"""
import ctypes
import time

import numpy as np
import torch

from . import _native as N
from .pytorch import hip_ops as H

PPT_CHOICES = (1, 2, 4, 8)
BLOCKS_PER_CU = (2, 4, 8)
_TUNE_CACHE = {}


def _align(n, a=64):
    return (n + a - 1) // a * a


class CpabBatch:
    """problems: list of (cpab, channels, data_size, outsize)."""

    def __init__(self, problems, device="cuda", save=True, ppt=None, blocks=None):
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.problems = problems
        self.save = save
        self.n = len(problems)
        L = N.lib()
        self._tess = [H.device_tess(T.params, self.device) for T, _, _, _ in problems]
        nd = [T.params.ndim for T, _, _, _ in problems]
        d = [t.d for t in self._tess]
        C = [c for _, c, _, _ in problems]
        dsz = [list(s) for _, _, s, _ in problems]
        osz = [list(o) for _, _, _, o in problems]
        nP = [int(np.prod(o)) for o in osz]
        n_data = [c * int(np.prod(s)) for c, s in zip(C, dsz)]
        n_out = [c * p for c, p in zip(C, nP)]
        n_sv = [int(N.lib().cpab_saved_floats(k, 1, p)) for k, p in zip(nd, nP)]
        n_lin = [sum(o) for o in osz]
        ws_b = [int(L.cpab_batch_workspace_bytes(t.ref, H._int_arr(o), 1)) for t, o in zip(self._tess, osz)]

        def offsets(sizes, align=64):
            off, o = [], 0
            for s in sizes:
                off.append(o)
                o += _align(s, align)
            return off, max(o, align)

        o_th, n_th = offsets(d)
        o_da, n_da = offsets(n_data)
        o_out, n_o = offsets(n_out)
        o_sv, n_s = offsets(n_sv)
        o_lin, n_l = offsets(n_lin)
        o_ws, n_w = offsets(ws_b, 256)
        f32 = dict(device=self.device, dtype=torch.float32)
        self.theta = torch.zeros(n_th, **f32)
        self.data = torch.zeros(n_da, **f32)
        self.out = torch.zeros(n_o, **f32)
        self.grad_out = torch.zeros(n_o, **f32)
        self.grad_data = torch.zeros(n_da, **f32)
        self.grad_theta = torch.zeros(n_th, **f32)
        self.saved = torch.empty(n_s if save else 64, **f32)
        self.lin = torch.zeros(n_l, **f32)
        self.ws = torch.empty(n_w, device=self.device, dtype=torch.uint8)
        self.counter = torch.zeros(64, device=self.device, dtype=torch.int32)
        for i in range(self.n):
            self.lin[o_lin[i]:o_lin[i] + n_lin[i]] = H.meshgrid_axes(osz[i], self.device)
        # per-problem views into the flat buffers
        self.theta_views = [self.theta[o:o + k] for o, k in zip(o_th, d)]
        self.data_views = [self.data[o:o + k].view([c] + s) for o, k, c, s in zip(o_da, n_data, C, dsz)]
        self.out_views = [self.out[o:o + k].view([c] + s) for o, k, c, s in zip(o_out, n_out, C, osz)]
        self.grad_out_views = [self.grad_out[o:o + k].view([c] + s) for o, k, c, s in zip(o_out, n_out, C, osz)]
        self.grad_data_views = [self.grad_data[o:o + k].view([c] + s) for o, k, c, s in zip(o_da, n_data, C, dsz)]
        self.grad_theta_views = [self.grad_theta[o:o + k] for o, k in zip(o_th, d)]
        # saved state of the forward per problem, [rows, nP] in output order
        # (rows 0..ndim-1: the warped points in x-space); None without save
        self.saved_views = [self.saved[o:o + k].view(-1, p) if save else None
                            for o, k, p in zip(o_sv, n_sv, nP)]
        self.points = sum(nP)
        self.used = {"data": n_data, "out": n_out}

        inst = (N.CpabInstance * max(self.n, 1))()
        fp = lambda t, o: t.data_ptr() + 4 * o
        for i, t in enumerate(self._tess):
            I = inst[i]
            I.tess = t.struct
            I.C = C[i]
            I.dsize = N.int3(dsz[i])
            I.outsize = N.int3(osz[i])
            I.nP = nP[i]
            I.theta = fp(self.theta, o_th[i])
            I.data = fp(self.data, o_da[i])
            I.lin = fp(self.lin, o_lin[i])
            I.out = fp(self.out, o_out[i])
            I.saved = fp(self.saved, o_sv[i]) if save else None
            I.grad_out = fp(self.grad_out, o_out[i])
            I.grad_data = fp(self.grad_data, o_da[i])
            I.grad_theta = fp(self.grad_theta, o_th[i])
            I.ws = self.ws.data_ptr() + o_ws[i]
            I.ws_bytes = ws_b[i]
        self._inst = inst
        self._dev_inst = None
        self._items = None
        self.ppt = None
        self.blocks = None
        self._graphs = None
        key = tuple((T.params.ndim, tuple(T.params.nc), c, tuple(s), tuple(o)) for T, c, s, o in problems)
        self._key = (hash(key), save)
        if ppt is not None and blocks is not None:
            self.set_launch(ppt, blocks)
        elif self._key in _TUNE_CACHE:
            self.set_launch(*_TUNE_CACHE[self._key])

    # -- launch configuration ------------------------------------------------
    def set_launch(self, ppt, blocks):
        L = N.lib()
        total = L.cpab_batch_plan(ctypes.cast(self._inst, ctypes.c_void_p), self.n, ppt, None, 0)
        if total < 0:
            N.check(-total)
        items = (ctypes.c_int * max(total, 1))()
        n_items = L.cpab_batch_plan(ctypes.cast(self._inst, ctypes.c_void_p), self.n, ppt,
                                    ctypes.cast(items, ctypes.c_void_p), total)
        if n_items < 0:
            N.check(-n_items)
        raw = bytes(memoryview(self._inst).cast("B"))
        self._dev_inst = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.device)
        self._items = torch.tensor(np.frombuffer(items, dtype=np.int32)[:n_items].copy(), device=self.device)
        self.n_items = n_items
        self.ppt = ppt
        self.blocks = blocks
        self._graphs = None

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    # -- launches ------------------------------------------------------------
    def forward(self):
        """out <- transform_data(data, theta) for every problem (one persistent launch)."""
        if self.ppt is None:
            self.autotune()
        N.check(N.lib().cpab_batch_fwd(ctypes.cast(self._inst, ctypes.c_void_p), self._dev_inst.data_ptr(), self.n,
                                       self._items.data_ptr(), self.n_items, self.ppt, self.blocks, 1.0,
                                       self.counter.data_ptr(), self._stream()))

    def backward(self):
        """grad_data, grad_theta <- backward of forward() for grad_out."""
        if self.ppt is None:
            self.autotune()
        N.check(N.lib().cpab_fill_zero(self.grad_data.data_ptr(), self.grad_data.numel(), self._stream()))
        N.check(N.lib().cpab_batch_bwd(ctypes.cast(self._inst, ctypes.c_void_p), self._dev_inst.data_ptr(), self.n,
                                       self._items.data_ptr(), self.n_items, self.ppt, self.blocks, 1.0,
                                       self.counter.data_ptr() + 4, self._stream()))

    def forward_plain(self):
        """The same forward as one launch sequence per problem."""
        N.check(N.lib().cpab_batch_fwd_plain(ctypes.cast(self._inst, ctypes.c_void_p), self.n, 1.0, self._stream()))

    def backward_plain(self):
        N.check(N.lib().cpab_fill_zero(self.grad_data.data_ptr(), self.grad_data.numel(), self._stream()))
        N.check(N.lib().cpab_batch_bwd_plain(ctypes.cast(self._inst, ctypes.c_void_p), self.n, 1.0, self._stream()))

    # -- graphs ------------------------------------------------------------------
    def capture(self):
        """Capture forward() and backward() into HIP graphs; replay with
        replay_forward() / replay_backward()."""
        if self.ppt is None:
            self.autotune()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):  # warm up outside capture
            self.forward()
            self.backward()
        torch.cuda.current_stream(self.device).wait_stream(s)
        gf, gb = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
        with torch.cuda.graph(gf):
            self.forward()
        with torch.cuda.graph(gb):
            self.backward()
        self._graphs = (gf, gb)

    def replay_forward(self):
        self._graphs[0].replay()

    def replay_backward(self):
        self._graphs[1].replay()

    # -- autotuner ---------------------------------------------------------------
    def autotune(self, reps=2, verbose=False):
        """Pick points-per-thread and the persistent grid size by timing
        forward + backward for each candidate on this batch."""
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        best = None
        results = []
        for ppt in PPT_CHOICES:
            for bpc in BLOCKS_PER_CU:
                self.set_launch(ppt, bpc * cus)
                try:
                    self.forward()
                    self.backward()
                except N.CpabError:  # this ppt does not fit the LDS budget of the batch
                    continue
                torch.cuda.synchronize(self.device)
                t0 = time.perf_counter()
                for _ in range(reps):
                    self.forward()
                    self.backward()
                torch.cuda.synchronize(self.device)
                dt = (time.perf_counter() - t0) / reps
                results.append((dt, ppt, bpc * cus))
                if best is None or dt < best[0]:
                    best = (dt, ppt, bpc * cus)
        if best is None:
            raise N.CpabError("no launch configuration of this batch fits the LDS budget")
        if verbose:
            for r in sorted(results):
                print("autotune: %.3f ms ppt=%d blocks=%d" % (r[0] * 1e3, r[1], r[2]))
        self.set_launch(best[1], best[2])
        _TUNE_CACHE[self._key] = (best[1], best[2])
        self.tuning = sorted(results)
        return best

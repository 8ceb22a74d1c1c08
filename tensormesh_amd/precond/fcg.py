"""Flexible CG with the learned preconditioner, Jacobi-PCG as the fallback.

``LearnedFCG`` owns persistent device buffers (matrix values, right-hand
side, iterate, flexible-CG work) so that ``chunk`` iterations - the network's
~14 kernels plus the three flexible-CG kernels each - are captured once in a
hipGraph and replayed; the host reads the device status after every replay.

Fallback rules (the true residual decides convergence):

* breakdown (``p.Aq <= 0`` or non-finite), the iteration cap, or stagnation
  (the recursive residual shrinks by less than ``stall_factor`` over
  ``stall_chunks`` consecutive replays) hand the current iterate to
  Jacobi-PCG;
* on convergence the true residual ``||b - A x|| / ||b||`` is recomputed in
  FP64; if it misses ``rtol``, Jacobi-PCG continues from ``x``.
"""
from __future__ import annotations

import ctypes
import time
import warnings
from typing import Optional

import numpy as np
import torch

from .. import _lib
from .._lib import check, ptr
from ..sparse.ops import CSR, _spmv_raw, pcg_raw
from .hierarchy import Hierarchy
from .network import GNNPreconditioner


class LearnedFCG:
    def __init__(self, net: GNNPreconditioner, hier: Hierarchy, csr: CSR, chunk: int = 8,
                 use_graph: bool = True, fused: bool = True, m: int = 4):
        self.net, self.hier, self.csr = net, hier, csr
        self.dev = hier.device
        self.n = csr.n_rows
        assert hier.levels[0].n == self.n, "hierarchy level 0 must be the matrix graph"
        self.prog = net.plan(hier)
        L = _lib.lib()
        self.m = int(m)
        ws = int(L.tm_fcg_work_size(self.n, self.m))
        if ws < 0:
            raise ValueError(f"flexible CG keeps 1, 2, 4 or 8 directions, got m={m}")
        self.work = torch.zeros(ws // 8 + 8, dtype=torch.float64, device=self.dev)
        self.view = _lib.FcgView()
        check(L.tm_fcg_layout(self.n, self.m, ptr(self.work), ctypes.byref(self.view)))
        base = self.work.data_ptr()
        off = lambda a: (a - base)
        self._status = self.work.view(torch.int32)[off(self.view.status) // 4: off(self.view.status) // 4 + 2]
        self._relres = self.work[off(self.view.relres) // 8: off(self.view.relres) // 8 + 1]
        self.values = torch.zeros(csr.nnz, dtype=torch.float32, device=self.dev)
        self.b = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        self.x = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        self.lk = torch.zeros(self.n, dtype=torch.float32, device=self.dev)
        self.chunk = int(chunk)
        self.use_graph = use_graph
        self.fused = fused
        self.graph = None

    # ---- set-up per system -----------------------------------------------------
    def set_system(self, values: torch.Tensor, lk: torch.Tensor):
        """Matrix values (canonical CSR order of ``csr``) and the log-coefficient feature."""
        self.values.copy_(values.detach().to(torch.float32))
        self.lk.copy_(lk.detach().to(torch.float32))
        # the fused symmetric finest level reads A and d only
        self.hier.set_matrix(self.values, level0_normalized=not (self.fused and self.prog.fusable_sym0))

    def _iterate_once(self):
        L = _lib.lib()
        v = self.view
        st = _lib.stream_of(self.dev)
        # the flexible-CG kernels keep s = rms(d r) at v.scale; the fused encoder reads it (nparts 0)
        fused_scale = self.fused and (self.prog.fusable_plain or self.prog.fusable_sym0)
        self.prog.forward(v.r, self.lk, ss_part=v.ss_part, nparts=0 if fused_scale else v.nparts, z=v.z,
                          scale=v.scale, stop=v.status, fused=self.fused, dr=v.dr)
        check(L.tm_fcg_iterate(self.n, self.m, ptr(self.csr.row_ptr), ptr(self.csr.col_idx), ptr(self.values),
                               ptr(self.x), ptr(self.hier.levels[0].d), ptr(self.work), st))

    def _capture(self):
        s = torch.cuda.Stream(self.dev)
        s.wait_stream(torch.cuda.current_stream(self.dev))
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=s):
            for _ in range(self.chunk):
                self._iterate_once()
        torch.cuda.current_stream(self.dev).wait_stream(s)

    # ---- solve ---------------------------------------------------------------
    def solve(self, b: torch.Tensor, rtol: float = 1e-8, maxiter: int = 2000, x0: Optional[torch.Tensor] = None,
              stall_factor: float = 0.9, stall_chunks: int = 3, fallback_maxiter: int = 100000):
        """Returns ``(x, info)``; ``x`` is a fresh FP64 tensor."""
        L = _lib.lib()
        st = _lib.stream_of(self.dev)
        t0 = time.perf_counter()
        if self.use_graph and self.graph is None:
            # first launches load the code objects outside the capture; the
            # state they leave is overwritten by tm_fcg_start
            self._iterate_once()
        self.b.copy_(b.detach().to(torch.float64))
        if x0 is None:
            self.x.zero_()
        else:
            self.x.copy_(x0.detach().to(torch.float64))
        check(L.tm_fcg_start(self.n, self.m, ptr(self.csr.row_ptr), ptr(self.csr.col_idx), ptr(self.values),
                             ptr(self.b), ptr(self.x), ptr(self.hier.levels[0].d), ptr(self.work), float(rtol), 0.0,
                             int(maxiter), st))
        if self.use_graph and self.graph is None:
            self._capture()
        hist = []
        stall = 0
        status, it = 0, 0
        while True:
            if self.use_graph:
                self.graph.replay()
            else:
                for _ in range(self.chunk):
                    self._iterate_once()
            sv = self._status.cpu()
            status, it = int(sv[0]), int(sv[1])
            rel = float(self._relres.item())
            hist.append(rel)
            if status != 0:
                break
            if len(hist) >= 2 and rel > stall_factor * hist[-2]:
                stall += 1
                if stall >= stall_chunks:
                    break
            else:
                stall = 0
        fcg_iters = it
        reason = {0: "stagnation", 1: "converged", 2: "breakdown", 3: "max iterations"}[status]
        x = self.x.clone()
        bn = float(torch.linalg.vector_norm(self.b))
        res = self.b - _spmv_raw(self.csr, self.values, x)
        true_rel = float(torch.linalg.vector_norm(res)) / bn if bn > 0 else 0.0
        fb_iters = 0
        if not (status == 1 and true_rel <= rtol):
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                x, fb_iters, _, ok = pcg_raw(self.csr, self.values, self.b, x0=x, rtol=rtol,
                                             maxiter=fallback_maxiter)
            res = self.b - _spmv_raw(self.csr, self.values, x)
            true_rel = float(torch.linalg.vector_norm(res)) / bn if bn > 0 else 0.0
        torch.cuda.synchronize(self.dev)
        info = {"fcg_iterations": fcg_iters, "fcg_exit": reason, "fallback_iterations": fb_iters,
                "relres": true_rel, "converged": true_rel <= rtol, "time_s": time.perf_counter() - t0,
                "relres_history": hist}
        return x, info

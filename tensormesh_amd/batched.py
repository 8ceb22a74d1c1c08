"""Many independent P1 Poisson problems assembled and solved together.

The reference has no batched path for meshes that differ: it loops over them
in Python (``docs/source/user_guide/batched_workflows.rst:247-270``).  Here the
meshes are stacked into one block-diagonal problem:

* one owner-computes assembly launch over the concatenated connectivity (the
  plan partitions each mesh separately: partition coordinates are shifted per
  mesh, the kernel uses the true ones),
* one load-vector launch, two gathers for the condensed values / rhs (zero
  Dirichlet on every mesh boundary),
* one persistent Jacobi-PCG launch in which every mesh keeps its own scalars
  and stops on its own relative residual.

All launches are asynchronous and allocation-free once the solver is built,
so the whole pipeline is captured into one hipGraph (:meth:`capture`).
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .assemble.topology import P1Topology
from .element import ELEMENT_DIM
from .sparse.ops import gather_raw


BatchedPCGArgs = _lib.BatchedArgs


def _L():
    return _lib.lib()


def make_tiles(mesh_rows: np.ndarray, rows_per_tile: int):
    """Tiles of at most ``rows_per_tile`` rows that never straddle two systems."""
    tile_row0, tile_mesh, mesh_tile0 = [0], [], [0]
    start = 0
    for k, n in enumerate(mesh_rows):
        n = int(n)
        nt = max(1, -(-n // rows_per_tile))
        for t in range(nt):
            tile_row0.append(start + min(n, (t + 1) * rows_per_tile))
            tile_mesh.append(k)
        mesh_tile0.append(mesh_tile0[-1] + nt)
        start += n
    return (np.asarray(tile_row0, np.int32), np.asarray(tile_mesh, np.int32), np.asarray(mesh_tile0, np.int32))


class BatchedPoisson:
    """``-div(kappa grad u) = c`` with zero Dirichlet data on a list of P1 meshes.

    ``meshes`` is a list of ``(points [N_k, d], cells [E_k, d+1], boundary [N_k])``
    numpy arrays (one element type for all).
    """

    def __init__(self, meshes: Sequence[Tuple[np.ndarray, np.ndarray, np.ndarray]], device="cuda",
                 rows_per_tile: int = 1024, lanes: int = 4, blocks: int = 0, value_dtype=torch.float32,
                 plan_rows: Optional[int] = None, plan_budget: Optional[int] = None):
        self.device = torch.device(device)
        dim = meshes[0][0].shape[1]
        nb = meshes[0][1].shape[1]
        self.element_type = {3: "triangle", 4: "tetra"}[nb]
        assert ELEMENT_DIM[self.element_type] == dim
        self.n_mesh = len(meshes)
        npts = np.array([m[0].shape[0] for m in meshes])
        nel = np.array([m[1].shape[0] for m in meshes])
        self.node_off = np.concatenate([[0], np.cumsum(npts)])
        self.elem_off = np.concatenate([[0], np.cumsum(nel)])
        pts = np.concatenate([m[0] for m in meshes]).astype(np.float64)
        conn = np.concatenate([m[1] + self.node_off[k] for k, m in enumerate(meshes)]).astype(np.int64)
        bnd = np.concatenate([m[2] for m in meshes]).astype(bool)
        # partition coordinates: each mesh moved to its own place so no block mixes meshes
        shift = np.zeros_like(pts)
        shift[:, 0] = 4.0 * np.repeat(np.arange(self.n_mesh), npts)
        self.topology = P1Topology(pts + shift, conn, len(pts), self.element_type)
        self.n_points, self.n_elements = len(pts), len(conn)
        dev = self.device
        self.coords = torch.from_numpy(pts).to(dev)
        self.plan = self.topology.device_plan(dev, "scalar", plan_rows, plan_budget)
        self.quad = self.topology.quad(2)
        self.value_dtype = value_dtype

        # condensation layout (zero Dirichlet on every boundary node)
        rp, ci = self.topology.row_ptr, self.topology.col_idx
        row = np.repeat(np.arange(self.n_points), np.diff(rp))
        inner = ~bnd
        keep = inner[row] & inner[ci]
        local = np.full(self.n_points, -1, np.int64)
        n_in = int(inner.sum())
        local[inner] = np.arange(n_in)
        r_in, c_in = local[row[keep]], local[ci[keep]]
        rp_in = np.zeros(n_in + 1, np.int32)
        np.cumsum(np.bincount(r_in, minlength=n_in), out=rp_in[1:])
        self.n_inner = n_in
        self.nnz_inner = int(keep.sum())
        self.mesh_inner = np.add.reduceat(inner.astype(np.int64), self.node_off[:-1]) if self.n_mesh else []
        self.inner_off = np.concatenate([[0], np.cumsum(self.mesh_inner)])
        self.rp_in = torch.from_numpy(rp_in).to(dev)
        self.ci_in = torch.from_numpy(c_in.astype(np.int32)).to(dev)
        self.edge_in = torch.from_numpy(np.nonzero(keep)[0].astype(np.int32)).to(dev)
        self.dof_in = torch.from_numpy(np.nonzero(inner)[0].astype(np.int32)).to(dev)

        # static buffers (the captured graph reads and writes only these)
        self.values = torch.empty(self.topology.nnz, dtype=value_dtype, device=dev)
        self.b_full = torch.empty(self.n_points, dtype=torch.float64, device=dev)
        self.vals_in = torch.empty(self.nnz_inner, dtype=value_dtype, device=dev)
        self.b_in = torch.empty(n_in, dtype=torch.float64, device=dev)
        self.x = torch.empty(n_in, dtype=torch.float64, device=dev)
        self.status = torch.zeros(self.n_mesh, dtype=torch.int32, device=dev)
        self.iters = torch.zeros(self.n_mesh, dtype=torch.int32, device=dev)
        self.relres = torch.zeros(self.n_mesh, dtype=torch.float64, device=dev)
        self.timeout = torch.zeros(1, dtype=torch.int32, device=dev)
        self.kappa = torch.ones(self.n_points, dtype=torch.float32, device=dev)
        self.graph = None
        self.configure(rows_per_tile, lanes, blocks)

    def configure(self, rows_per_tile: int, lanes: int, blocks: int = 0):
        t0, tm, mt0 = make_tiles(np.asarray(self.mesh_inner), rows_per_tile)
        dev = self.device
        self.rows_per_tile, self.lanes, self.blocks = int(rows_per_tile), int(lanes), int(blocks)
        self.tile_row0 = torch.from_numpy(t0).to(dev)
        self.tile_mesh = torch.from_numpy(tm).to(dev)
        self.mesh_tile0 = torch.from_numpy(mt0).to(dev)
        self.n_tiles = len(tm)
        L = _L()
        nw = int(L.tm_batched_pcg_work_size(self.n_inner, self.n_tiles))
        self.work = torch.empty(nw // 8 + 1, dtype=torch.float64, device=dev)
        self.graph = None

    def _args(self, rtol, maxiter) -> BatchedPCGArgs:
        p = _lib.ptr
        return BatchedPCGArgs(self.n_inner, self.n_tiles, self.n_mesh, p(self.rp_in).value, p(self.ci_in).value,
                              p(self.vals_in).value, _lib.dtype_code(self.vals_in), 0, p(self.b_in).value,
                              p(self.x).value, p(self.tile_row0).value, p(self.tile_mesh).value,
                              p(self.mesh_tile0).value, float(rtol), int(maxiter), p(self.work).value,
                              p(self.status).value, p(self.iters).value, p(self.relres).value)

    # ---- pipeline stages (asynchronous, allocation-free) ----------------------
    def _assemble(self, stream):
        L = _lib.lib()
        _lib.check(L.tm_asm_p1_matrix(_lib.FORM_KAPPA_LAPLACE, self.topology.nb, self.topology.dim,
                                      self.plan.struct, _lib.ptr(self.coords), _lib.ptr(self.kappa),
                                      _lib.dtype_code(self.kappa), 1.0, self.quad, _lib.ptr(self.values),
                                      _lib.dtype_code(self.values), stream))
        _lib.check(L.tm_asm_p1_load(_lib.LOAD_CONST, self.topology.nb, self.topology.dim, self.plan.struct,
                                    _lib.ptr(self.coords), None, 0, self.load_c, self.quad, _lib.ptr(self.b_full),
                                    stream))
        _lib.check(L.tm_gather(self.nnz_inner, _lib.ptr(self.edge_in), _lib.ptr(self.values),
                               _lib.ptr(self.vals_in), self.values.element_size(), stream))
        _lib.check(L.tm_gather(self.n_inner, _lib.ptr(self.dof_in), _lib.ptr(self.b_full), _lib.ptr(self.b_in),
                               8, stream))

    def _solve(self, stream, rtol, maxiter):
        L = _L()
        a = self._args(rtol, maxiter)
        _lib.check(L.tm_batched_pcg(ctypes.byref(a), self.lanes, self.blocks, stream))
        _lib.check(L.tm_batched_timeout_flag(ctypes.byref(a), _lib.ptr(self.timeout), stream))

    def run(self, kappa: Optional[torch.Tensor] = None, c: float = 1.0, rtol: float = 1e-8, maxiter: int = 20000):
        """Assemble and solve every system (stream-ordered, returns without syncing)."""
        if kappa is not None:
            self.kappa.copy_(kappa.to(self.kappa.dtype))
        self.load_c = float(c)
        st = _lib.stream_of(self.device)
        self._assemble(st)
        self._solve(st, rtol, maxiter)

    def capture(self, c: float = 1.0, rtol: float = 1e-8, maxiter: int = 20000) -> torch.cuda.CUDAGraph:
        """Capture assembly + condensation + batched PCG into one graph; replay with ``graph.replay()``
        after writing new coefficients into ``self.kappa``."""
        self.load_c = float(c)
        self.run(None, c, rtol, maxiter)          # warm-up outside the capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            st = _lib.stream_of(self.device)
            self._assemble(st)
            self._solve(st, rtol, maxiter)
        self.graph = g
        return g

    def results(self):
        """``(x per mesh [inner nodes], status, iterations, relres)`` after a synchronised run."""
        if int(self.timeout.item()) != 0:
            raise _lib.TMError(_lib.TM_ERR_HIP, "batched PCG grid barrier timed out (grid not resident)")
        xs = [self.x[self.inner_off[k]:self.inner_off[k + 1]] for k in range(self.n_mesh)]
        return xs, self.status.cpu().numpy(), self.iters.cpu().numpy(), self.relres.cpu().numpy()

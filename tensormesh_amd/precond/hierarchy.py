"""Nested node aggregates and Galerkin coarse graphs for the learned preconditioner.

Level 0 is the graph of the (condensed) system matrix.  Level ``l >= 1``
groups level-0 nodes by the cell of a regular grid of spacing ``2^l h0``
containing them (``h0`` = median shortest edge, so a structured lattice
coarsens 2:1 per axis); cells are nested, so every level-``l`` aggregate is a
union of level-``(l-1)`` aggregates.  The coarse graph has an edge (I, J)
whenever some fine edge joins a child of I to a child of J, and its values are
the Galerkin sums ``A_c = P^T A P`` with piecewise-constant ``P``, reduced on
the GPU in a fixed order through ``perm``/``seg`` (coarse entry ``s`` sums
fine entries ``perm[seg[s]:seg[s+1]]``).

Everything here depends on the mesh only; :meth:`Hierarchy.set_matrix`
refreshes the per-matrix values (diagonal scaling, normalised values,
pooling weights) with HIP kernels.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from .. import _lib
from .._lib import check, ptr


@dataclass
class HostLevel:
    n: int
    row_ptr: np.ndarray            # int32 [n+1]
    col: np.ndarray                # int32 [nnz]
    agg: Optional[np.ndarray] = None        # int32 [n]: parent on the next level
    child_ptr: Optional[np.ndarray] = None  # int32 [n+1]: children on the previous level
    child: Optional[np.ndarray] = None      # int32
    perm: Optional[np.ndarray] = None       # int32 [nnz_prev]: previous-level entries by coarse slot
    seg: Optional[np.ndarray] = None        # int32 [nnz+1]
    # level 0 only: pattern of B = A P (rows of this level, columns on level 1)
    b_row_ptr: Optional[np.ndarray] = None
    b_col: Optional[np.ndarray] = None
    b_perm: Optional[np.ndarray] = None     # int32 [nnz]: entries of this level by B slot
    b_seg: Optional[np.ndarray] = None


def _csr_rows(row_ptr: np.ndarray) -> np.ndarray:
    return np.repeat(np.arange(len(row_ptr) - 1, dtype=np.int64), np.diff(row_ptr))


def shortest_edge(coords: np.ndarray, row_ptr: np.ndarray, col: np.ndarray) -> float:
    rows = _csr_rows(row_ptr)
    off = rows != col
    d = np.linalg.norm(coords[rows[off]] - coords[col[off]], axis=1)
    # per-row minimum, then the median over rows
    big = np.full(len(row_ptr) - 1, np.inf)
    np.minimum.at(big, rows[off], d)
    big = big[np.isfinite(big)]
    return float(np.median(big)) if len(big) else 1.0


def build_host_levels(coords: np.ndarray, row_ptr: np.ndarray, col: np.ndarray, n_levels: int) -> List[HostLevel]:
    """Levels 0..n_levels-1 (coarsening stops early once a level has one node)."""
    coords = np.asarray(coords, dtype=np.float64)
    row_ptr = np.asarray(row_ptr, dtype=np.int32)
    col = np.asarray(col, dtype=np.int32)
    n0 = len(row_ptr) - 1
    levels = [HostLevel(n0, row_ptr, col)]
    if n_levels <= 1 or n0 <= 1:
        return levels
    h0 = shortest_edge(coords, row_ptr, col)
    cells = np.floor((coords - coords.min(0)) / (2.0 * h0) + 1e-6).astype(np.int64)  # level-1 cells
    member = np.arange(n0)          # level-0 node -> its node on the current level
    for lev in range(1, n_levels):
        if lev > 1:
            cells = cells // 2
        ext = cells.max(0) + 1
        key = np.zeros(n0, dtype=np.int64)
        for ax in range(cells.shape[1] - 1, -1, -1):
            key = key * int(ext[ax]) + cells[:, ax]
        uniq, node_of0 = np.unique(key, return_inverse=True)
        prev = levels[-1]
        # parent of each previous-level node (all its level-0 members share the cell)
        agg = np.empty(prev.n, dtype=np.int64)
        agg[member] = node_of0
        nc = len(uniq)
        order = np.argsort(agg, kind="stable")
        counts = np.bincount(agg, minlength=nc)
        child_ptr = np.zeros(nc + 1, dtype=np.int64)
        np.cumsum(counts, out=child_ptr[1:])
        rows = _csr_rows(prev.row_ptr)
        ckey = agg[rows] * nc + agg[prev.col.astype(np.int64)]
        ukey, inv = np.unique(ckey, return_inverse=True)
        perm = np.argsort(inv, kind="stable")
        seg = np.zeros(len(ukey) + 1, dtype=np.int64)
        np.cumsum(np.bincount(inv, minlength=len(ukey)), out=seg[1:])
        crow, ccol = ukey // nc, ukey % nc
        crp = np.zeros(nc + 1, dtype=np.int64)
        np.cumsum(np.bincount(crow, minlength=nc), out=crp[1:])
        prev.agg = agg.astype(np.int32)
        if lev == 1:
            bkey = rows * nc + agg[prev.col.astype(np.int64)]
            bu, binv = np.unique(bkey, return_inverse=True)
            bseg = np.zeros(len(bu) + 1, dtype=np.int64)
            np.cumsum(np.bincount(binv, minlength=len(bu)), out=bseg[1:])
            brp = np.zeros(prev.n + 1, dtype=np.int64)
            np.cumsum(np.bincount(bu // nc, minlength=prev.n), out=brp[1:])
            prev.b_row_ptr, prev.b_col = brp.astype(np.int32), (bu % nc).astype(np.int32)
            prev.b_perm, prev.b_seg = np.argsort(binv, kind="stable").astype(np.int32), bseg.astype(np.int32)
        levels.append(HostLevel(nc, crp.astype(np.int32), ccol.astype(np.int32), child_ptr=child_ptr.astype(np.int32),
                                child=order.astype(np.int32), perm=perm.astype(np.int32), seg=seg.astype(np.int32)))
        member = node_of0
        if nc <= 1:
            break
    return levels


@dataclass
class DeviceLevel:
    n: int
    nnz: int
    row_ptr: torch.Tensor
    col: torch.Tensor
    agg: Optional[torch.Tensor]
    child_ptr: Optional[torch.Tensor]
    child: Optional[torch.Tensor]
    perm: Optional[torch.Tensor]
    seg: Optional[torch.Tensor]
    values: Optional[torch.Tensor] = None   # Galerkin values (FP32), level 0: the matrix itself
    d: Optional[torch.Tensor] = None        # diag^-1/2
    vn: Optional[torch.Tensor] = None       # normalised values D A D
    w: Optional[torch.Tensor] = None        # pooling weight of each node towards its parent
    b_row_ptr: Optional[torch.Tensor] = None  # level 0: B = A^ P^ (CSR onto level 1)
    b_col: Optional[torch.Tensor] = None
    b_perm: Optional[torch.Tensor] = None
    b_seg: Optional[torch.Tensor] = None
    b_values: Optional[torch.Tensor] = None
    row_sums: Optional[torch.Tensor] = None   # level 0: A^ 1


class Hierarchy:
    """Device copy of the aggregation levels of one mesh graph."""

    def __init__(self, host_levels: List[HostLevel], device: torch.device, level0_csr=None):
        """``level0_csr``: optional device ``(row_ptr, col)`` of the level-0
        matrix (int32, the same pattern as ``host_levels[0]``) to share with
        the solver instead of a second copy, so one iteration streams one
        column array for A and A^ (at 96^3 that is 50 MB less per iteration)."""
        self.device = torch.device(device)
        self.level0_ready = False
        t = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
        self.levels = []
        for k, h in enumerate(host_levels):
            rp, col = (level0_csr if (k == 0 and level0_csr is not None) else (t(h.row_ptr), t(h.col)))
            if k == 0 and level0_csr is not None:
                assert rp.dtype == torch.int32 and col.dtype == torch.int32 and rp.numel() == h.n + 1 \
                    and col.numel() == h.col.size, "level-0 CSR does not match the hierarchy's pattern"
            self.levels.append(DeviceLevel(h.n, int(h.col.size), rp, col, t(h.agg), t(h.child_ptr), t(h.child),
                                           t(h.perm), t(h.seg), b_row_ptr=t(h.b_row_ptr), b_col=t(h.b_col),
                                           b_perm=t(h.b_perm), b_seg=t(h.b_seg)))

    @classmethod
    def from_csr(cls, coords: np.ndarray, row_ptr: np.ndarray, col: np.ndarray, n_levels: int,
                 device: torch.device, level0_csr=None) -> "Hierarchy":
        return cls(build_host_levels(coords, row_ptr, col, n_levels), device, level0_csr)

    @property
    def n_levels(self) -> int:
        return len(self.levels)

    def set_matrix(self, values: torch.Tensor, level0_normalized: bool = True):
        """Per-matrix set-up from the level-0 FP32 values (canonical CSR order).

        ``level0_normalized=False`` skips the level-0 products that only the
        unfused network paths and training read (the normalised values
        ``A^`` at level 0, ``B = A^ P^`` and ``A^ 1``); the fused symmetric
        finest level works from ``A`` and ``d``.  ``level0_ready`` records
        which set-up ran."""
        _lib.require_gpu(values, "matrix values")
        L = _lib.lib()
        st = _lib.stream_of(self.device)
        v = values.detach()
        if v.dtype != torch.float32:
            v = v.to(torch.float32)
        v = v.contiguous()
        for k, lev in enumerate(self.levels):
            if k == 0:
                lev.values = v
            else:
                if lev.values is None:
                    lev.values = torch.empty(lev.nnz, dtype=torch.float32, device=self.device)
                prev = self.levels[k - 1]
                check(L.tm_gnn_level_coarsen(lev.nnz, ptr(lev.seg), ptr(lev.perm), ptr(prev.values),
                                             ptr(lev.values), st))
            if lev.d is None:
                lev.d = torch.empty(lev.n, dtype=torch.float32, device=self.device)
            if lev.vn is None or lev.vn.numel() != lev.nnz:
                lev.vn = torch.empty(lev.nnz, dtype=torch.float32, device=self.device)
            skip_vn = k == 0 and not level0_normalized
            check(L.tm_gnn_level_setup(lev.n, ptr(lev.row_ptr), ptr(lev.col), ptr(lev.values), ptr(lev.d),
                                       ptr(None) if skip_vn else ptr(lev.vn), st))
        for k in range(len(self.levels) - 1):
            lev, nxt = self.levels[k], self.levels[k + 1]
            if lev.w is None:
                lev.w = torch.empty(lev.n, dtype=torch.float32, device=self.device)
            check(L.tm_gnn_pool_weights(lev.n, ptr(lev.agg), ptr(lev.d), ptr(nxt.d), ptr(lev.w), st))
        l0 = self.levels[0]
        self.level0_ready = bool(level0_normalized)
        if l0.b_row_ptr is not None and level0_normalized:
            if l0.b_values is None:
                l0.b_values = torch.empty(l0.b_col.numel(), dtype=torch.float32, device=self.device)
                l0.row_sums = torch.empty(l0.n, dtype=torch.float32, device=self.device)
            check(L.tm_gnn_level_prolong(l0.b_col.numel(), ptr(l0.b_seg), ptr(l0.b_perm), ptr(l0.col), ptr(l0.vn),
                                         ptr(l0.w), ptr(l0.b_values), st))
            check(L.tm_gnn_row_sums(l0.n, ptr(l0.row_ptr), ptr(l0.vn), ptr(l0.row_sums), st))

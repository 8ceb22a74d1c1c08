"""Dirichlet static condensation.

Drop-in for ``tensormesh.Condenser`` (``tensormesh/operator/condense.py:144-528``):
``Condenser(dirichlet_mask, dirichlet_value=None)``, ``__call__(matrix,
rhs=None) -> (K_ii, f_i - K_io u_o)``, ``condense_rhs``, ``recover``,
``restrict``, ``prolong``, ``update_dirichlet``.  The layout (inner / outer
DOF masks, monotone local numbering, inner-inner and inner-outer edge
masks, ``_compute_layout`` ``:182-220``) is computed once on the matrix's
device and cached; it is bit-identical to the reference's.  For a scalar
GPU matrix in canonical CSR order the layout comes from the HIP
condensation kernels (``tm_condense_count`` / ``tm_condense_fill``: prefix
sums over the rows, one host read of the four sizes), and the per-call work
is two value gathers plus one kernel for ``f_i - K_io u_o``; ``recover`` and
``prolong`` are one kernel each.  ``K_ii`` keeps the canonical CSR order, so
it carries its CSR view into ``solve`` without conversion.  Calls that need
gradients with respect to the right-hand side, the values or ``u`` take the
differentiable path (value gathers with autograd, torch indexing).

Deviation: the condensed right-hand side keeps the wider of the rhs and
matrix dtypes (an FP64 load vector stays FP64 next to FP32 values).
"""
from __future__ import annotations

from typing import Optional

import torch
import torch.nn as nn

from .. import _lib
from ..sparse.matrix import SparseMatrix
from ..sparse.ops import CSR, Gather


class Condenser(nn.Module):
    _LAZY_BUFFERS = ("inner_row", "inner_col", "ou2in_row", "ou2in_col", "is_inner_edge", "is_ou2in_edge",
                     "is_inner_dof", "is_outer_dof")

    def __init__(self, dirichlet_mask: torch.Tensor, dirichlet_value: Optional[torch.Tensor] = None):
        super().__init__()
        assert dirichlet_mask.dtype == torch.bool, f"dirichlet_mask must be a bool tensor, got {dirichlet_mask.dtype}"
        assert dirichlet_mask.ndim == 1, f"dirichlet_mask must be 1D, got shape {tuple(dirichlet_mask.shape)}"
        self.register_buffer("dirichlet_mask", dirichlet_mask)
        self.register_buffer("dirichlet_value", self._normalize_value(dirichlet_value))
        for name in self._LAZY_BUFFERS:
            self.register_buffer(name, None, persistent=False)
        self.inner_shape = None
        self.ou2in_shape = None
        self.n_inner_dof = None
        self.n_outer_dof = None
        self.n_dof = None
        self.layout_signature = None
        self.K_ou2in = None
        self._dev = None

    def _normalize_value(self, value):
        mask = self.dirichlet_mask
        n_outer = int(mask.sum())
        if value is None:
            return torch.zeros(n_outer, device=mask.device)
        assert value.ndim == 1, f"dirichlet_value must be 1D, got shape {tuple(value.shape)}"
        if value.shape[0] == mask.shape[0]:
            return value[mask.to(value.device)]
        assert value.shape[0] == n_outer, \
            f"dirichlet_value must have length n_dof ({mask.shape[0]}) or n_outer_dof ({n_outer}), got {value.shape[0]}"
        return value

    def update_dirichlet(self, dirichlet_value: torch.Tensor):
        self.dirichlet_value = self._normalize_value(dirichlet_value)

    def _device_layout(self, matrix: SparseMatrix) -> bool:
        """Layout from the HIP condensation kernels; False when the matrix is
        not a scalar canonical-CSR GPU matrix."""
        dev = matrix.device
        if dev.type != "cuda" or matrix.block_size != 1:
            return False
        csr = matrix.csr()   # built once for a COO matrix; the solve reuses it
        if matrix._perm is not None:
            return False
        L = _lib.lib()
        st = _lib.stream_of(dev)
        n, nnz = matrix.shape[0], int(csr.col_idx.numel())
        i32, i64, u8 = torch.int32, torch.int64, torch.uint8
        dirm = self.dirichlet_mask.to(dev).to(u8).contiguous()
        work = torch.empty(int(L.tm_condense_work_size(n)), dtype=u8, device=dev)
        local = torch.empty(n, dtype=i32, device=dev)
        sizes = torch.empty(4, dtype=i64, device=dev)
        _lib.check(L.tm_condense_count(n, _lib.ptr(csr.row_ptr), _lib.ptr(csr.col_idx), _lib.ptr(dirm),
                                       _lib.ptr(work), _lib.ptr(local), _lib.ptr(sizes), st))
        n_inner, n_outer, nnz_ii, nnz_io = (int(v) for v in sizes.cpu())

        # at least one element each, so every pointer is valid (an empty
        # tensor's data_ptr() is 0); the attributes are views of length k
        sz = dict(rp_ii=(n_inner + 1, i32), ci_ii=(nnz_ii, i32), e_ii=(nnz_ii, i32), rp_io=(n_inner + 1, i32),
                  ci_io=(nnz_io, i32), e_io=(nnz_io, i32), inner_dof=(n_inner, i32), outer_dof=(n_outer, i32),
                  coo_ii=(2 * nnz_ii, i64), coo_io=(2 * nnz_io, i64), m_ii=(nnz, u8), m_io=(nnz, u8))
        full = {k: torch.empty(max(1, c), dtype=dt, device=dev) for k, (c, dt) in sz.items()}
        _lib.check(L.tm_condense_fill(n, _lib.ptr(csr.row_ptr), _lib.ptr(csr.col_idx), _lib.ptr(dirm),
                                      _lib.ptr(work), _lib.ptr(local),
                                      *(_lib.ptr(full[k]) for k in ("rp_ii", "ci_ii", "e_ii", "rp_io", "ci_io",
                                                                    "e_io", "inner_dof", "outer_dof", "coo_ii",
                                                                    "coo_io", "m_ii", "m_io")), st))
        v = {k: full[k][:c] for k, (c, _) in sz.items()}
        rp_ii, ci_ii, e_ii, rp_io, ci_io, e_io = (v[k] for k in ("rp_ii", "ci_ii", "e_ii", "rp_io", "ci_io", "e_io"))
        inner_dof, outer_dof, coo_ii, coo_io, m_ii, m_io = (v[k] for k in ("inner_dof", "outer_dof", "coo_ii",
                                                                            "coo_io", "m_ii", "m_io"))
        self.inner_row, self.inner_col = coo_ii[:nnz_ii], coo_ii[nnz_ii:]
        self.ou2in_row, self.ou2in_col = coo_io[:nnz_io], coo_io[nnz_io:]
        self.is_inner_edge, self.is_ou2in_edge = m_ii.view(torch.bool), m_io.view(torch.bool)
        self.is_inner_dof = dirm == 0
        self.is_outer_dof = dirm != 0
        self.inner_shape = (n_inner, n_inner)
        self.ou2in_shape = (n_inner, n_outer)
        self.n_inner_dof, self.n_outer_dof, self.n_dof = n_inner, n_outer, n
        self.layout_signature = matrix.layout_signature
        self._dev = dict(
            ii=e_ii, ii_long=e_ii.to(i64), io=e_io, io_long=e_io.to(i64),
            csr_ii=CSR(rp_ii, ci_ii, n_inner, n_inner), csr_io=CSR(rp_io, ci_io, n_inner, n_outer),
            dirm=dirm, local=local, inner_dof=inner_dof, outer_dof=outer_dof,
            inner_dof_buf=full["inner_dof"], ci_io_buf=full["ci_io"], rp_io_buf=full["rp_io"],
        )
        return True

    def _compute_layout(self, matrix: SparseMatrix):
        if self._device_layout(matrix):
            return
        dev = matrix.device
        edge_u, edge_v = matrix.row, matrix.col
        n_dof = matrix.shape[0]
        is_inner_dof = ~self.dirichlet_mask.to(dev)
        is_outer_dof = ~is_inner_dof
        is_inner_edge = is_inner_dof[edge_u] & is_inner_dof[edge_v]
        is_ou2in_edge = is_inner_dof[edge_u] & is_outer_dof[edge_v]
        n_inner, n_outer = int(is_inner_dof.sum()), int(is_outer_dof.sum())
        local = torch.full((n_dof,), -1, dtype=torch.long, device=dev)
        local[is_inner_dof] = torch.arange(n_inner, device=dev)
        local[is_outer_dof] = torch.arange(n_outer, device=dev)
        self.inner_row = local[edge_u[is_inner_edge]]
        self.inner_col = local[edge_v[is_inner_edge]]
        self.ou2in_row = local[edge_u[is_ou2in_edge]]
        self.ou2in_col = local[edge_v[is_ou2in_edge]]
        self.is_inner_edge = is_inner_edge
        self.is_ou2in_edge = is_ou2in_edge
        self.is_inner_dof = is_inner_dof
        self.is_outer_dof = is_outer_dof
        self.inner_shape = (n_inner, n_inner)
        self.ou2in_shape = (n_inner, n_outer)
        self.n_inner_dof, self.n_outer_dof, self.n_dof = n_inner, n_outer, n_dof
        self.layout_signature = matrix.layout_signature
        self._dev = None
        if dev.type == "cuda":
            ii = torch.nonzero(is_inner_edge).flatten()
            io = torch.nonzero(is_ou2in_edge).flatten()
            canonical = matrix._csr is not None and matrix._perm is None

            def csr_of(rows, cols, n_rows, n_cols):
                rp = torch.zeros(n_rows + 1, dtype=torch.int32, device=dev)
                rp[1:] = torch.cumsum(torch.bincount(rows, minlength=n_rows), 0).to(torch.int32)
                return CSR(rp, cols.to(torch.int32), n_rows, n_cols)

            self._dev = dict(
                ii=ii.to(torch.int32), ii_long=ii, io=io.to(torch.int32), io_long=io,
                csr_ii=csr_of(self.inner_row, self.inner_col, n_inner, n_inner) if canonical else None,
                csr_io=csr_of(self.ou2in_row, self.ou2in_col, n_inner, n_outer) if canonical else None,
            )

    def __call__(self, matrix: SparseMatrix, rhs: Optional[torch.Tensor] = None):
        if rhs is None:
            rhs = torch.zeros(matrix.shape[0], device=matrix.device)
        if self.inner_row is None:
            self._compute_layout(matrix)
        assert matrix.shape[0] == self.n_dof, f"the shape of matrix must be [{self.n_dof}, {self.n_dof}], but got {matrix.shape}"
        assert matrix.shape[1] == self.n_dof, f"the shape of matrix must be [{self.n_dof}, {self.n_dof}], but got {matrix.shape}"
        assert matrix.has_same_layout(self.layout_signature), \
            "the layout of the matrix is changed, please recompute the condensed matrix"
        assert rhs.shape[0] == self.n_dof, f"the shape of rhs must be [{self.n_dof}, ...], but got {rhs.shape}"
        d = self._dev
        if d is not None and matrix.values.is_cuda:
            v_ii = Gather.apply(matrix.edata, d["ii"], d["ii_long"])
            v_io = Gather.apply(matrix.edata, d["io"], d["io_long"])
            if d["csr_ii"] is not None:
                K_inner = SparseMatrix._from_csr(v_ii, self.inner_row, self.inner_col, self.inner_shape,
                                                 d["csr_ii"], symmetric=matrix.symmetric)
                K_ou2in = SparseMatrix._from_csr(v_io, self.ou2in_row, self.ou2in_col, self.ou2in_shape,
                                                 d["csr_io"])
            else:
                K_inner = SparseMatrix(v_ii, self.inner_row, self.inner_col, self.inner_shape)
                K_ou2in = SparseMatrix(v_io, self.ou2in_row, self.ou2in_col, self.ou2in_shape)
        else:
            K_inner = SparseMatrix(matrix.edata[self.is_inner_edge], self.inner_row, self.inner_col, self.inner_shape)
            K_ou2in = SparseMatrix(matrix.edata[self.is_ou2in_edge], self.ou2in_row, self.ou2in_col, self.ou2in_shape)
        self.K_ou2in = K_ou2in
        return K_inner, self._rhs(rhs, K_inner.edata)

    def _kernel_ok(self, *ts) -> bool:
        d = self._dev
        return (d is not None and "inner_dof" in d and all(t.is_cuda for t in ts)
                and not (torch.is_grad_enabled() and any(t.requires_grad for t in ts)))

    def _rhs(self, rhs: torch.Tensor, values: torch.Tensor) -> torch.Tensor:
        dt = torch.promote_types(rhs.dtype, values.dtype)
        if rhs.dim() == 1 and self._kernel_ok(rhs, values, self.dirichlet_value.to(values.device)):
            d = self._dev
            u_o = self.dirichlet_value.to(device=values.device, dtype=torch.float64).contiguous()
            self.dirichlet_value = self.dirichlet_value.to(device=values.device, dtype=dt)
            r64 = rhs.to(device=values.device, dtype=torch.float64).contiguous()
            out = torch.empty(max(1, self.n_inner_dof), dtype=torch.float64, device=values.device)
            csr_io = d["csr_io"]
            has_io = csr_io.col_idx.numel() > 0
            _lib.check(_lib.lib().tm_condense_rhs(self.n_inner_dof, _lib.ptr(d["inner_dof_buf"]), _lib.ptr(r64),
                                                  _lib.ptr(d["rp_io_buf"]), _lib.ptr(d["ci_io_buf"]),
                                                  _lib.ptr(self.K_ou2in.values) if has_io else _lib.ptr(None),
                                                  _lib.dtype_code(self.K_ou2in.values),
                                                  _lib.ptr(u_o) if has_io else _lib.ptr(None),
                                                  _lib.ptr(out), _lib.stream_of(values.device)))
            return out[:self.n_inner_dof].to(dt)
        rhs = rhs.to(device=values.device, dtype=dt)
        self.dirichlet_value = self.dirichlet_value.to(device=values.device, dtype=dt)
        minus = (self.K_ou2in @ self.dirichlet_value).to(dt)
        for _ in range(rhs.dim() - 1):
            minus = minus.unsqueeze(-1)
        return rhs[self.is_inner_dof] - minus

    def condense_rhs(self, rhs: torch.Tensor) -> torch.Tensor:
        assert self.K_ou2in is not None, \
            "call the Condenser on a matrix first; condense_rhs reuses the cached K_io block"
        return self._rhs(rhs, self.K_ou2in.edata)

    def _recover_kernel(self, u: torch.Tensor, u_outer: Optional[torch.Tensor]) -> torch.Tensor:
        d = self._dev
        u64 = u.detach().to(torch.float64).contiguous()
        uo = None if u_outer is None else u_outer.to(device=u.device, dtype=torch.float64).contiguous()
        out = torch.empty(self.n_dof, dtype=torch.float64, device=u.device)
        _lib.check(_lib.lib().tm_condense_recover(self.n_dof, _lib.ptr(d["dirm"]), _lib.ptr(d["local"]),
                                                  _lib.ptr(u64), _lib.ptr(uo) if self.n_outer_dof else
                                                  _lib.ptr(None), _lib.ptr(out), _lib.stream_of(u.device)))
        return out.to(u.dtype)

    def recover(self, u: torch.Tensor) -> torch.Tensor:
        assert u.shape[0] == self.n_inner_dof, f"the shape of u must be [{self.n_inner_dof}, ...], but got {u.shape}"
        if u.dim() == 1 and u.is_floating_point() and self._kernel_ok(u) and self.n_inner_dof > 0:
            return self._recover_kernel(u, self.dirichlet_value)
        shape = list(u.shape)
        shape[0] = self.n_dof
        out = torch.zeros(shape, dtype=u.dtype, device=u.device)
        out[self.is_inner_dof.to(u.device)] = u
        bv = self.dirichlet_value.to(device=u.device, dtype=u.dtype)
        for _ in range(u.dim() - 1):
            bv = bv.unsqueeze(-1)
        out[self.is_outer_dof.to(u.device)] = bv.expand((self.n_outer_dof,) + tuple(u.shape[1:]))
        return out

    def restrict(self, f: torch.Tensor) -> torch.Tensor:
        assert f.shape[0] == self.dirichlet_mask.shape[0], \
            f"the shape of f must be [{self.dirichlet_mask.shape[0]}, ...], but got {f.shape}"
        return f[~self.dirichlet_mask.to(f.device)]

    def prolong(self, f_inner: torch.Tensor) -> torch.Tensor:
        if (f_inner.dim() == 1 and f_inner.is_floating_point() and self._kernel_ok(f_inner)
                and f_inner.shape[0] == self.n_inner_dof and self.n_inner_dof > 0):
            return self._recover_kernel(f_inner, None)
        mask = self.dirichlet_mask.to(f_inner.device)
        n_inner = int((~mask).sum())
        assert f_inner.shape[0] == n_inner, f"the shape of f_inner must be [{n_inner}, ...], but got {f_inner.shape}"
        shape = list(f_inner.shape)
        shape[0] = mask.shape[0]
        out = torch.zeros(shape, dtype=f_inner.dtype, device=f_inner.device)
        out[~mask] = f_inner
        return out

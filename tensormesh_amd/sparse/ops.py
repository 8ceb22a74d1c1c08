"""Autograd-aware device operators over CSR matrices (HIP kernels).

* :class:`SpMV` - ``y = A x`` with ``dx = A^T g`` and ``dvalues = g[row] x[col]``
  (the in-tree reference ``SparseMVTorch``, ``tensormesh/sparse/mm/torch_mm.py:32-48``).
* :class:`PCGSolve` - Jacobi-PCG solve with the adjoint backward of
  ``tensormesh/sparse/solve/torch_sla_solve.py:102-136``:
  ``lambda = A^-T g``, ``db = lambda``, ``dvalues = -lambda[row] x[col]``.
* :class:`Gather` - value gather used by condensation, scatter in backward.
"""
from __future__ import annotations

import ctypes
import warnings
from dataclasses import dataclass, field
from typing import Optional

import torch

from .. import _lib


@dataclass
class CSR:
    """Device CSR view: int32 row pointers / columns, values in canonical order."""
    row_ptr: torch.Tensor
    col_idx: torch.Tensor
    n_rows: int
    n_cols: int
    # int16 copy of col - row (built on first use by the scalar solver; False when
    # some entry lies more than 32767 from the diagonal)
    col_delta: Optional[object] = field(default=None, repr=False, compare=False)

    @property
    def nnz(self) -> int:
        return int(self.col_idx.numel())


#: Scalar Jacobi-PCG reads column indices as int16 offsets from the diagonal when
#: every entry of the matrix fits (2 bytes per entry less to stream per SpMV).
COL_DELTA16 = True


def col_delta16(csr: CSR) -> Optional[torch.Tensor]:
    """The int16 ``col - row`` array of ``csr`` (cached), or None if an entry does not fit."""
    if csr.col_delta is None:
        dev = csr.col_idx.device
        cd = torch.empty(max(csr.nnz, 1), dtype=torch.int16, device=dev)
        fits = torch.empty(1, dtype=torch.int32, device=dev)
        _lib.check(_lib.lib().tm_csr_delta16(csr.n_rows, _lib.ptr(csr.row_ptr), _lib.ptr(csr.col_idx), _lib.ptr(cd),
                                             _lib.ptr(fits), _lib.stream_of(dev)))
        csr.col_delta = cd if int(fits.item()) == 1 else False
    return csr.col_delta if csr.col_delta is not False else None


def _spmv_raw(csr: CSR, values: torch.Tensor, x: torch.Tensor, block_size: int = 1) -> torch.Tensor:
    """``y = A x``; with ``block_size`` 3 the CSR is over block rows and values hold
    one row-major 3x3 block per entry."""
    _lib.require_gpu(values, "matrix values")
    x64 = x.detach().to(torch.float64).contiguous()
    y = torch.empty(csr.n_rows * block_size, dtype=torch.float64, device=values.device)
    fn = _lib.lib().tm_spmv if block_size == 1 else _lib.lib().tm_spmv_bsr3
    assert block_size in (1, 3), f"block size {block_size} not supported"
    _lib.check(fn(csr.n_rows, csr.nnz, _lib.ptr(csr.row_ptr), _lib.ptr(csr.col_idx),
                  _lib.ptr(values.detach().contiguous()), _lib.dtype_code(values),
                  _lib.ptr(x64), _lib.ptr(y), _lib.stream_of(values.device)))
    return y


def _transpose_csr(csr: CSR, values: torch.Tensor):
    """Row-major CSR of A^T (torch sort; used only for non-symmetric adjoints)."""
    rows = torch.repeat_interleave(torch.arange(csr.n_rows, device=values.device),
                                   (csr.row_ptr[1:] - csr.row_ptr[:-1]).long())
    cols = csr.col_idx.long()
    key = cols * csr.n_rows + rows
    order = torch.argsort(key)
    counts = torch.bincount(cols, minlength=csr.n_cols)
    rp = torch.zeros(csr.n_cols + 1, dtype=torch.int32, device=values.device)
    rp[1:] = torch.cumsum(counts, 0).to(torch.int32)
    return CSR(rp, rows[order].to(torch.int32), csr.n_cols, csr.n_rows), values.detach()[order]


class SpMV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, x, csr: CSR, row: torch.Tensor, col: torch.Tensor, symmetric: bool,
                block_size: int = 1):
        ctx.csr, ctx.symmetric, ctx.bs = csr, symmetric, block_size
        ctx.save_for_backward(values, x, row, col)
        y = _spmv_raw(csr, values, x, block_size)
        return y.to(torch.promote_types(values.dtype, x.dtype))

    @staticmethod
    def backward(ctx, g):
        values, x, row, col = ctx.saved_tensors
        gv = gx = None
        if ctx.needs_input_grad[0]:
            gv = (g[row] * x[col]).to(values.dtype)
        if ctx.needs_input_grad[1]:
            if ctx.symmetric:
                gx = _spmv_raw(ctx.csr, values, g, ctx.bs)
            else:
                assert ctx.bs == 1, "transpose products of non-symmetric block matrices are not supported"
                csr_t, vt = _transpose_csr(ctx.csr, values)
                gx = _spmv_raw(csr_t, vt, g)
            gx = gx.to(x.dtype)
        return gv, gx, None, None, None, None, None


def pcg_raw(csr: CSR, values: torch.Tensor, b: torch.Tensor, x0: Optional[torch.Tensor] = None,
            rtol: float = 1e-8, atol: float = 0.0, maxiter: int = 10000, check_every: int = 0,
            raise_on_fail: bool = False, block_size: int = 1, status: Optional[dict] = None):
    """Jacobi-PCG on the device; returns ``(x [f64], iterations, relres, converged)``.
    ``status`` (optional) receives ``breakdown``: True when ``p.Ap <= 0`` ended
    the solve (the matrix is not positive definite)."""
    _lib.require_gpu(values, "matrix values")
    dev = values.device
    n = csr.n_rows * block_size
    b64 = b.detach().to(torch.float64).contiguous()
    x = (torch.zeros(n, dtype=torch.float64, device=dev) if x0 is None
         else x0.detach().to(torch.float64).clone().contiguous())
    L = _lib.lib()
    work = torch.empty(int(L.tm_pcg_work_size(n)) // 8 + 1, dtype=torch.float64, device=dev)
    it = ctypes.c_int64(0)
    rr = ctypes.c_double(0.0)
    fn, cols = (L.tm_pcg_jacobi, csr.col_idx) if block_size == 1 else (L.tm_pcg_jacobi_bsr3, csr.col_idx)
    if block_size == 1 and COL_DELTA16 and csr.n_rows > 0:
        cd = col_delta16(csr)
        if cd is not None:
            fn, cols = L.tm_pcg_jacobi_d16, cd
    st = fn(csr.n_rows, _lib.ptr(csr.row_ptr), _lib.ptr(cols), _lib.ptr(values.detach().contiguous()),
                         _lib.dtype_code(values), _lib.ptr(b64), _lib.ptr(x), _lib.ptr(work), float(rtol),
                         float(atol), int(maxiter), int(check_every), ctypes.byref(it), ctypes.byref(rr),
                         _lib.stream_of(dev))
    converged = st == _lib.TM_OK
    breakdown = False
    if not converged:
        if st != _lib.TM_ERR_NOTCONV or raise_on_fail:
            _lib.check(st)
        msg = L.tm_last_error().decode()
        breakdown = msg.startswith("PCG breakdown")
        warnings.warn(f"PCG did not converge: {msg}")
    if status is not None:
        status["breakdown"] = breakdown
    return x, int(it.value), float(rr.value), converged


class PCGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, csr: CSR, row, col, rtol, atol, maxiter, info: dict, block_size: int = 1):
        st: dict = {}
        x, it, rel, ok = pcg_raw(csr, values, b, None, rtol, atol, maxiter, block_size=block_size, status=st)
        info.update(iterations=it, relres=rel, converged=ok, breakdown=st["breakdown"], method="jacobi-pcg")
        ctx.csr, ctx.tol, ctx.bs = csr, (rtol, atol, maxiter), block_size
        ctx.save_for_backward(values, x, row, col)
        return x.to(torch.promote_types(values.dtype, b.dtype))

    @staticmethod
    def backward(ctx, g):
        values, x, row, col = ctx.saved_tensors
        rtol, atol, maxiter = ctx.tol
        lam, _, _, _ = pcg_raw(ctx.csr, values, g, None, rtol, atol, maxiter,
                               block_size=ctx.bs)  # A symmetric: A^-T = A^-1
        gv = (-lam[row] * x[col]).to(values.dtype) if ctx.needs_input_grad[0] else None
        gb = lam.to(g.dtype) if ctx.needs_input_grad[1] else None
        return gv, gb, None, None, None, None, None, None, None, None


def bicgstab_raw(csr: CSR, values: torch.Tensor, b: torch.Tensor, x0: Optional[torch.Tensor] = None,
                 rtol: float = 1e-8, atol: float = 0.0, maxiter: int = 10000, check_every: int = 0,
                 raise_on_fail: bool = False):
    """Jacobi-preconditioned BiCGSTAB on the device (square CSR, any symmetry);
    returns ``(x [f64], iterations, relres, converged)``."""
    _lib.require_gpu(values, "matrix values")
    dev = values.device
    n = csr.n_rows
    b64 = b.detach().to(torch.float64).contiguous()
    x = (torch.zeros(n, dtype=torch.float64, device=dev) if x0 is None
         else x0.detach().to(torch.float64).clone().contiguous())
    L = _lib.lib()
    work = torch.empty(int(L.tm_bicgstab_work_size(n)) // 8 + 1, dtype=torch.float64, device=dev)
    it = ctypes.c_int64(0)
    rr = ctypes.c_double(0.0)
    st = L.tm_bicgstab_jacobi(n, csr.nnz, _lib.ptr(csr.row_ptr), _lib.ptr(csr.col_idx),
                              _lib.ptr(values.detach().contiguous()), _lib.dtype_code(values), _lib.ptr(b64),
                              _lib.ptr(x), _lib.ptr(work), float(rtol), float(atol), int(maxiter), int(check_every),
                              ctypes.byref(it), ctypes.byref(rr), _lib.stream_of(dev))
    converged = st == _lib.TM_OK
    if not converged:
        if st != _lib.TM_ERR_NOTCONV or raise_on_fail:
            _lib.check(st)
        warnings.warn(f"BiCGSTAB did not converge: {L.tm_last_error().decode()}")
    return x, int(it.value), float(rr.value), converged


class BiCGSTABSolve(torch.autograd.Function):
    """``x = A^-1 b`` by BiCGSTAB; backward ``lam = A^-T g`` by BiCGSTAB on the
    transposed CSR, ``dL/dvalues = -lam[row] x[col]`` (``torch_sla_solve.py:102-136``)."""

    @staticmethod
    def forward(ctx, values, b, csr: CSR, row, col, rtol, atol, maxiter, info: dict):
        x, it, rel, ok = bicgstab_raw(csr, values, b, None, rtol, atol, maxiter)
        info.update(iterations=it, relres=rel, converged=ok, method="bicgstab")
        ctx.csr, ctx.tol = csr, (rtol, atol, maxiter)
        ctx.save_for_backward(values, x, row, col)
        return x.to(torch.promote_types(values.dtype, b.dtype))

    @staticmethod
    def backward(ctx, g):
        values, x, row, col = ctx.saved_tensors
        rtol, atol, maxiter = ctx.tol
        csr_t, vt = _transpose_csr(ctx.csr, values)
        lam, _, _, _ = bicgstab_raw(csr_t, vt, g, None, rtol, atol, maxiter)
        gv = (-lam[row] * x[col]).to(values.dtype) if ctx.needs_input_grad[0] else None
        gb = lam.to(g.dtype) if ctx.needs_input_grad[1] else None
        return gv, gb, None, None, None, None, None, None, None


def gather_raw(src: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    _lib.require_gpu(src, "gather source")
    src = src.contiguous()
    out = torch.empty(idx.numel(), dtype=src.dtype, device=src.device)
    _lib.check(_lib.lib().tm_gather(idx.numel(), _lib.ptr(idx), _lib.ptr(src), _lib.ptr(out),
                                    src.element_size(), _lib.stream_of(src.device)))
    return out


class Gather(torch.autograd.Function):
    """``out = src[idx]`` for a duplicate-free int32 index list."""

    @staticmethod
    def forward(ctx, src, idx, idx_long):
        ctx.n = src.numel()
        ctx.save_for_backward(idx_long)
        return gather_raw(src, idx)

    @staticmethod
    def backward(ctx, g):
        (idx_long,) = ctx.saved_tensors
        out = torch.zeros(ctx.n, dtype=g.dtype, device=g.device)
        out.index_copy_(0, idx_long, g)
        return out, None, None

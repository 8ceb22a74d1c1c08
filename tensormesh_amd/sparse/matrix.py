"""COO sparse matrix with the reference's FEM surface, backed by device CSR kernels.

Mirrors ``tensormesh/sparse/matrix.py`` (``SparseMatrix(edata, row, col,
shape)``, ``:78-84``) and the members callers use (``values``/``edata``,
``row``/``col``, ``layout_signature`` / ``has_same_layout``
(``tensormesh/sparse/mixin.py:48-86``), ``transpose``/``T`` (``:231-241``),
``to_scipy_coo`` (``:245-251``), ``from_block_coo`` (``:325-375``),
``to_dense``, ``degree``, ``@`` and ``solve``).  ``@`` and ``solve`` came
from the third-party ``torch_sla.SparseTensor`` in the reference; here they
run the HIP SpMV / Jacobi-PCG on GPU tensors (CPU tensors use SciPy SuperLU,
the CPU backend the reference reports, ``README.md:132``).

Matrices produced by the assemblers and the Condenser carry their CSR view
(rows ascending, columns ascending): no conversion happens on ``@`` /
``solve``.  Other COO inputs are converted once (sorted, duplicates summed).
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import scipy.sparse
import scipy.sparse.linalg
import torch

from .ops import CSR, SpMV, PCGSolve, BiCGSTABSolve


class SparseMatrix:
    def __init__(self, edata: torch.Tensor, row: torch.Tensor, col: torch.Tensor, shape: Tuple[int, int]):
        assert edata.dim() == 1, f"edata must be 1D, got {tuple(edata.shape)}"
        assert row.shape == edata.shape and col.shape == edata.shape, \
            f"row/col/edata must have the same shape, got {tuple(row.shape)}, {tuple(col.shape)}, {tuple(edata.shape)}"
        self.values = edata
        self.row_indices = row if row.dtype == torch.long else row.long()
        self.col_indices = col if col.dtype == torch.long else col.long()
        self.shape = (int(shape[0]), int(shape[1]))
        self._csr: Optional[CSR] = None
        self._perm = None          # canonical order of a non-canonical COO
        self.symmetric: Optional[bool] = None
        # 3 when the COO is the block expansion of 3x3 blocks on a node pattern
        # (from_block_coo order); then _csr is the node-level pattern
        self.block_size: int = 1
        self.last_solve_info: dict = {}

    @classmethod
    def _from_csr(cls, values, row, col, shape, csr: CSR, symmetric: Optional[bool] = None):
        m = cls(values, row, col, shape)
        m._csr = csr
        m.symmetric = symmetric
        return m

    # ---- reference aliases ------------------------------------------------------
    @property
    def edata(self):
        return self.values

    @property
    def row(self):
        return self.row_indices

    @property
    def col(self):
        return self.col_indices

    @property
    def edges(self):
        return torch.stack([self.row_indices, self.col_indices], 0)

    @property
    def nnz(self) -> int:
        return int(self.values.numel())

    @property
    def sparse_shape(self):
        return self.shape

    @property
    def dtype(self):
        return self.values.dtype

    @property
    def device(self):
        return self.values.device

    @property
    def grad(self):
        return self.values.grad

    @property
    def requires_grad(self):
        return self.values.requires_grad

    @property
    def layout_signature(self):
        r, c = self.row_indices, self.col_indices
        return (r.data_ptr(), r._version, c.data_ptr(), c._version, r.numel(), self.shape)

    def has_same_layout(self, other) -> bool:
        if isinstance(other, tuple):
            return self.layout_signature == other
        return self.layout_signature == other.layout_signature

    def layout_mask(self) -> torch.Tensor:
        m = torch.zeros(self.shape, dtype=torch.bool, device=self.device)
        m[self.row_indices, self.col_indices] = True
        return m

    def degree(self, axis: int = 0) -> torch.Tensor:
        idx = self.row_indices if axis == 0 else self.col_indices
        return torch.bincount(idx, minlength=self.shape[axis])

    # ---- conversions ------------------------------------------------------------
    def _wrap(self, values, row=None, col=None):
        m = SparseMatrix(values, self.row_indices if row is None else row,
                         self.col_indices if col is None else col, self.shape)
        if row is None and col is None and values.device == self.values.device:
            m._csr, m._perm, m.symmetric, m.block_size = self._csr, self._perm, self.symmetric, self.block_size
        return m

    def to(self, *args, **kwargs) -> "SparseMatrix":
        v = self.values.to(*args, **kwargs)
        if v.device != self.values.device:
            return SparseMatrix(v, self.row_indices.to(v.device), self.col_indices.to(v.device), self.shape)
        return self._wrap(v)

    def cuda(self, device=None):
        return self.to(torch.device("cuda") if device is None else device)

    def cpu(self):
        return self.to("cpu")

    def double(self):
        return self._wrap(self.values.double())

    def float(self):
        return self._wrap(self.values.float())

    def detach(self):
        return self._wrap(self.values.detach())

    def requires_grad_(self, flag: bool = True):
        self.values.requires_grad_(flag)
        return self

    def transpose(self) -> "SparseMatrix":
        t = SparseMatrix(self.values, self.col_indices, self.row_indices, (self.shape[1], self.shape[0]))
        t.symmetric = self.symmetric
        return t

    @property
    def T(self):
        return self.transpose()

    def to_dense(self) -> torch.Tensor:
        d = torch.zeros(self.shape, dtype=self.values.dtype, device=self.values.device)
        return d.index_put((self.row_indices, self.col_indices), self.values, accumulate=True)

    def to_scipy_coo(self) -> scipy.sparse.coo_matrix:
        return scipy.sparse.coo_matrix((self.values.detach().cpu().numpy(),
                                        (self.row_indices.detach().cpu().numpy(),
                                         self.col_indices.detach().cpu().numpy())), shape=self.shape)

    def to_sparse_coo(self) -> torch.Tensor:
        return torch.sparse_coo_tensor(self.edges, self.values, self.shape)

    def sum(self, dim=None):
        if dim is None:
            return self.values.sum()
        idx = self.row_indices if dim == 1 else self.col_indices
        out = torch.zeros(self.shape[0 if dim == 1 else 1], dtype=self.values.dtype, device=self.device)
        return out.index_add(0, idx, self.values)

    @staticmethod
    def _sum_symmetry(a: Optional[bool], b: Optional[bool]) -> Optional[bool]:
        """Symmetry flag of a sum on one layout: symmetric + symmetric is
        symmetric, symmetric + non-symmetric is not, anything else is checked
        when needed."""
        if a is True and b is True:
            return True
        if (a is True and b is False) or (a is False and b is True):
            return False
        return None

    def __add__(self, other):
        if isinstance(other, SparseMatrix) and self.has_same_layout(other):
            m = self._wrap(self.values + other.values)
            m.symmetric = self._sum_symmetry(self.symmetric, other.symmetric)
            return m
        if isinstance(other, SparseMatrix):
            return SparseMatrix(torch.cat([self.values, other.values]),
                                torch.cat([self.row_indices, other.row_indices]),
                                torch.cat([self.col_indices, other.col_indices]), self.shape)
        return NotImplemented

    def __mul__(self, s):
        if isinstance(s, (int, float)) or (isinstance(s, torch.Tensor) and s.dim() == 0):
            return self._wrap(self.values * s)
        return NotImplemented

    __rmul__ = __mul__

    def __neg__(self):
        return self._wrap(-self.values)

    def __sub__(self, other):
        return self + (-other)

    # ---- CSR view -----------------------------------------------------------------
    def csr(self) -> CSR:
        """Device CSR of this matrix (built once from the COO when not attached)."""
        if self._csr is None:
            n, m = self.shape
            r, c = self.row_indices, self.col_indices
            key = r * m + c
            canonical = bool((key[1:] > key[:-1]).all()) if key.numel() > 1 else True
            if canonical:
                counts = torch.bincount(r, minlength=n)
                rp = torch.zeros(n + 1, dtype=torch.int32, device=r.device)
                rp[1:] = torch.cumsum(counts, 0).to(torch.int32)
                self._csr = CSR(rp, c.to(torch.int32), n, m)
            else:
                uk, inv = torch.unique(key, sorted=True, return_inverse=True)
                self._perm = (inv, uk.numel())
                ur, uc = uk // m, uk % m
                counts = torch.bincount(ur, minlength=n)
                rp = torch.zeros(n + 1, dtype=torch.int32, device=r.device)
                rp[1:] = torch.cumsum(counts, 0).to(torch.int32)
                self._csr = CSR(rp, uc.to(torch.int32), n, m)
                self._canon_rc = (ur, uc)
        return self._csr

    def _canonical(self):
        """(values, row, col) in the CSR order of :meth:`csr`."""
        csr = self.csr()
        if self._perm is None:
            return self.values, self.row_indices, self.col_indices
        inv, nu = self._perm
        v = torch.zeros(nu, dtype=self.values.dtype, device=self.values.device).index_add(0, inv, self.values)
        return v, self._canon_rc[0], self._canon_rc[1]

    def is_symmetric(self) -> bool:
        """Pattern symmetric and ``max |v - v^T| <= tol * max |v|`` with
        ``tol = 32 eps`` of the value type: duplicate COO entries are summed
        with atomics in no fixed order, so the two triangles of a symmetric
        input can differ in the last bits."""
        if self.symmetric is None:
            v, r, c = self._canonical()
            t = SparseMatrix(v, c, r, self.shape)
            tv, tr, tc = t._canonical()
            sym = bool(torch.equal(tr, r) and torch.equal(tc, c))
            if sym and v.numel() > 0:
                vd, tvd = v.detach(), tv.detach()
                tol = 32.0 * torch.finfo(vd.dtype).eps if vd.is_floating_point() else 0.0
                scale = float(vd.abs().max())
                sym = float((vd - tvd).abs().max()) <= tol * scale
            self.symmetric = sym
        return self.symmetric

    # ---- products and solves -------------------------------------------------------
    def __matmul__(self, x: torch.Tensor) -> torch.Tensor:
        if isinstance(x, SparseMatrix):
            raise NotImplementedError("sparse @ sparse is not on the assembly/solve path")
        assert x.shape[0] == self.shape[1], f"shape mismatch: {self.shape} @ {tuple(x.shape)}"
        if not self.values.is_cuda:
            return (self.to_sparse_coo() @ x.to(self.values.dtype))
        if x.dim() == 2:
            return torch.stack([self @ x[:, k] for k in range(x.shape[1])], 1)
        if self.block_size != 1:
            return SpMV.apply(self.values, x, self._csr, self.row_indices, self.col_indices,
                              bool(self.symmetric), self.block_size)
        v, r, c = self._canonical()
        sym = bool(self.symmetric) if self.symmetric is not None else False
        return SpMV.apply(v, x, self.csr(), r, c, sym)

    def solve(self, b: torch.Tensor, backend: Optional[str] = None, method: Optional[str] = None,
              atol: float = 1e-10, tol: float = 1e-12, maxiter: int = 10000, verbose: bool = False,
              **kwargs) -> torch.Tensor:
        """Solve ``A x = b``.

        GPU tensors: Jacobi-preconditioned CG for symmetric (SPD) matrices,
        Jacobi-preconditioned BiCGSTAB (the reference's ``lse_solver``) for
        non-symmetric ones (assembled matrices carry the flag; others are
        checked once), with ``method="bicgstab"``, and for symmetric matrices
        on which CG breaks down (``p.Ap <= 0``: not positive definite); stops
        when ``||b - A x|| <= max(tol ||b||, atol)``.  ``backend`` may be
        ``None``/``"hip"``; ``method`` ``None``/``"cg"``/``"pcg"``/``"bicgstab"``.
        CPU tensors: SciPy SuperLU.  Differentiable w.r.t. values and b.
        """
        assert self.shape[0] == self.shape[1], f"solve needs a square matrix, got {self.shape}"
        assert b.shape[0] == self.shape[0], f"rhs must have {self.shape[0]} rows, got {tuple(b.shape)}"
        if b.dim() == 2:
            return torch.stack([self.solve(b[:, k], backend, method, atol, tol, maxiter, verbose)
                                for k in range(b.shape[1])], 1)
        if not self.values.is_cuda:
            A = self.to_scipy_coo().tocsc()
            x = scipy.sparse.linalg.spsolve(A, b.detach().cpu().numpy().astype(np.float64))
            if verbose:
                print(f"[tensormesh_amd] solve: n={self.shape[0]}, nnz={self.nnz}, device=cpu, "
                      f"backend=scipy, method=lu")
            return torch.from_numpy(np.asarray(x)).to(torch.promote_types(self.dtype, b.dtype))
        if backend not in (None, "hip", "mi355x", "auto"):
            raise ValueError(f"backend {backend!r} is not available for GPU tensors (use 'hip')")
        if method not in (None, "cg", "pcg", "jacobi-pcg", "bicgstab", "auto"):
            raise ValueError(f"method {method!r} is not available on the GPU (Jacobi-PCG or BiCGSTAB)")
        info: dict = {}
        if method == "bicgstab" or (method in (None, "auto") and self.block_size == 1 and not self.is_symmetric()):
            if self.block_size != 1:
                raise ValueError("BiCGSTAB covers scalar CSR matrices")
            v, r, c = self._canonical()
            x = BiCGSTABSolve.apply(v, b.to(self.device), self.csr(), r, c, float(tol), float(atol), int(maxiter),
                                    info)
            self.last_solve_info = info
            if verbose:
                print(f"[tensormesh_amd] solve: n={self.shape[0]}, nnz={self.nnz}, dtype={self.dtype}, "
                      f"device={self.device}, backend=hip, method=bicgstab, "
                      f"iterations={info.get('iterations')}, relres={info.get('relres'):.3e}")
            return x
        if self.block_size != 1:
            x = PCGSolve.apply(self.values, b.to(self.device), self._csr, self.row_indices, self.col_indices,
                               float(tol), float(atol), int(maxiter), info, self.block_size)
        else:
            v, r, c = self._canonical()
            x = PCGSolve.apply(v, b.to(self.device), self.csr(), r, c, float(tol), float(atol), int(maxiter), info)
            if info.get("breakdown") and method in (None, "auto"):
                # symmetric but not positive definite: the reference picks a
                # general solver for such systems (spd=False, README.md:132);
                # here that is BiCGSTAB
                pcg_info = dict(info)
                info = {}
                x = BiCGSTABSolve.apply(v, b.to(self.device), self.csr(), r, c, float(tol), float(atol),
                                        int(maxiter), info)
                info["after_pcg_breakdown"] = pcg_info
        self.last_solve_info = info
        if verbose:
            print(f"[tensormesh_amd] solve: n={self.shape[0]}, nnz={self.nnz}, dtype={self.dtype}, "
                  f"device={self.device}, backend=hip, method={info.get('method')}, "
                  f"iterations={info.get('iterations')}, relres={info.get('relres'):.3e}")
        return x

    # ---- constructors ---------------------------------------------------------------
    @staticmethod
    def from_block_coo(edata: torch.Tensor, row: torch.Tensor, col: torch.Tensor,
                       shape: Tuple[int, int]) -> "SparseMatrix":
        """Expand ``[n, h, h]`` blocks at coarse ``(row, col)`` into a flat COO matrix:
        entry ``(row*h + i, col*h + j)`` for each block, row-major inside blocks."""
        h = edata.shape[1]
        assert edata.shape[2] == h, f"from_block_coo expects square blocks, got {tuple(edata.shape)}"
        i = torch.arange(h, device=row.device)
        rr = (row.long()[:, None, None] * h + i[None, :, None]).expand(-1, h, h).reshape(-1)
        cc = (col.long()[:, None, None] * h + i[None, None, :]).expand(-1, h, h).reshape(-1)
        return SparseMatrix(edata.reshape(-1), rr, cc, (shape[0] * h, shape[1] * h))

    @staticmethod
    def from_dense(t: torch.Tensor) -> "SparseMatrix":
        rows, cols = torch.where(t != 0)
        return SparseMatrix(t[rows, cols], rows, cols, tuple(t.shape))

    def __repr__(self):
        return f"SparseMatrix(shape={self.shape}, nnz={self.nnz}, dtype={self.dtype}, device={self.device})"

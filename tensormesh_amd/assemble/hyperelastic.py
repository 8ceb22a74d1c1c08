"""Compressible neo-Hookean hyperelasticity on P1 tetrahedra.

Drop-in for ``tensormesh.assemble.NeoHookeanModel`` (``tensormesh/assemble/builtin.py:220-334``):
``from_mesh(mesh, E=1.0, nu=0.3)``, ``energy(u)``, ``element_energy``.
``E`` may be a scalar or a per-element tensor (the reference passes
per-element data through ``element_data``, ``element_assembler.py:604-618``).
The reference obtains the residual by autograd of the energy and has no
tangent assembly; here all three come from HIP kernels:

* :meth:`residual` - ``R = dW/du`` ``[N, 3]``, differentiable: the backward is the
  hand-written adjoint ``dL/du = K(u) g`` (matrix-free tangent action) and
  ``dL/dE_e = g . dR/dE_e``;
* :meth:`tangent` - ``K(u)`` as 3x3 blocks on the scalar pattern, returned as
  the reference's block COO (``SparseMatrix.from_block_coo`` order) with a
  block-CSR view for ``@`` / ``solve``;
* :meth:`newton` - Newton's method with Jacobi-PCG inner solves on the
  condensed tangent (``sparse/nonlinear_solve.py:21-77`` update rule).
"""
from __future__ import annotations

from typing import Optional

import torch

from .. import _lib
from ..sparse.matrix import SparseMatrix
from ..sparse.ops import CSR, Gather, pcg_raw
from .element_assembler import ElementAssembler


def _e_arg(E):
    if torch.is_tensor(E):
        return E.contiguous(), _lib.dtype_code(E), 0.0
    return None, _lib.F64, float(E)


class _Residual(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, E_t, model, E0):
        ctx.model, ctx.E0 = model, E0
        ctx.save_for_backward(u, E_t if E_t is not None else torch.empty(0))
        ctx.has_E = E_t is not None
        return model._launch(u, E_t if E_t is not None else E0, want_r=True)[1]

    @staticmethod
    def backward(ctx, g):
        u, E_t = ctx.saved_tensors
        E = E_t if ctx.has_E else ctx.E0
        need_E = ctx.has_E and ctx.needs_input_grad[1]
        y, dE = ctx.model._tangent_action(u, E, g.contiguous().to(torch.float64), want_dE=need_E)
        gu = y.to(u.dtype) if ctx.needs_input_grad[0] else None
        gE = dE.to(E_t.dtype) if need_E else None
        return gu, gE, None, None


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, model, E):
        ctx.model, ctx.E = model, E
        ctx.save_for_backward(u)
        return model._launch(u, E, want_energy=True)[2]

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        R = ctx.model._launch(u, ctx.E, want_r=True)[1]
        return (g * R).to(u.dtype), None, None


def block_matrix(asm: ElementAssembler, kv: torch.Tensor, dev) -> SparseMatrix:
    """``[3N, 3N]`` matrix from 3x3 blocks ``kv[nnz*9]`` on ``asm``'s scalar pattern:
    COO in the ``SparseMatrix.from_block_coo`` order (edge-major, then the block
    row-major: row = node*3 + i, col = node*3 + k) with a block-CSR view for
    ``@`` / ``solve``.  The expanded indices are cached on ``asm`` per device."""
    key = ("_bcoo", dev)
    c = asm.__dict__.get("_bcoo_cache")
    if c is None or c[0] != key:
        e0, e1 = asm.edges[0].to(dev), asm.edges[1].to(dev)
        i = torch.arange(3, device=dev)
        rr = (e0[:, None, None] * 3 + i[None, :, None]).expand(-1, 3, 3).reshape(-1)
        cc = (e1[:, None, None] * 3 + i[None, None, :]).expand(-1, 3, 3).reshape(-1)
        c = (key, rr, cc)
        asm.__dict__["_bcoo_cache"] = c
    rp, ci = asm.topology.csr(dev)
    m = SparseMatrix._from_csr(kv, c[1], c[2], (3 * asm.n_points, 3 * asm.n_points),
                               CSR(rp, ci, asm.n_points, asm.n_points), symmetric=True)
    m.block_size = 3
    return m


class NeoHookeanModel(ElementAssembler):
    value_dtype = torch.float32

    def __post_init__(self, E=1.0, nu=0.3):
        self.E = E
        self.nu = float(nu)
        if not torch.is_tensor(E):
            self.mu = E / (2 * (1 + nu))
            self.lam = E * nu / ((1 + nu) * (1 - 2 * nu))

    def element_energy(self, graddisplacement):
        grad_u = graddisplacement
        dim = grad_u.shape[-1]
        F = torch.eye(dim, device=grad_u.device, dtype=grad_u.dtype) + grad_u
        J = torch.linalg.det(F)
        log_J = torch.log(J)
        return (self.mu / 2) * ((F * F).sum() - dim) - self.mu * log_J + (self.lam / 2) * (log_J ** 2)

    # ---- kernels ----------------------------------------------------------------
    def _coords(self, dev):
        key = ("_x64", dev)
        c = self.__dict__.get("_coords_cache")
        if c is None or c[0] != key:
            c = (key, self.points.detach().to(dev, torch.float64).contiguous())
            self.__dict__["_coords_cache"] = c
        return c[1]

    def _vplan(self, dev, assemble: bool = False):
        """``assemble=True``: the 3x3-block plan of the tangent kernel, fastest with bisection
        leaves of at most 56 rows in 40 KiB (four blocks per CU; 64^3 tangent + residual 172.8 us
        against 206.0 / 185.9 / 181.8 / 184.0 / 235.0 us for 40 / 54 / 58 / 60 / 62 rows; leaves
        that overflow the budget are halved, so 56-row leaves give 28-row blocks there).
        Otherwise the accumulator-free
        plan (80 rows in 24 KiB) of the calls that keep no CSR accumulators: the matrix-free
        tangent action, and residual or energy evaluations without the tangent."""
        assert self.topology.nb == 4, "neo-Hookean kernels cover P1 tetrahedra"
        if assemble:
            return self.topology.device_plan(dev, "vector3", max_rows=56, lds_budget=40 * 1024 - 64)
        return self.topology.device_plan(dev, "vector3a")

    def _resolve_E(self, E):
        E = self.E if E is None else E
        if torch.is_tensor(E):
            assert E.shape == (self.topology.n_elements,), \
                f"E must be a scalar or [{self.topology.n_elements}] per element, got {tuple(E.shape)}"
        return E

    def _launch(self, u, E, want_k=False, want_r=False, want_energy=False, kdtype=torch.float32):
        dev = u.device
        _lib.require_gpu(u, "displacement")
        u64 = u.detach().to(torch.float64).contiguous()
        assert u64.shape == (self.n_points, 3), f"u must be [{self.n_points}, 3], got {tuple(u.shape)}"
        Et, Edt, E0 = _e_arg(E.detach().to(dev) if torch.is_tensor(E) else E)
        kv = torch.empty(self.topology.nnz * 9, dtype=kdtype, device=dev) if want_k else None
        R = torch.empty((self.n_points, 3), dtype=torch.float64, device=dev) if want_r else None
        en = torch.zeros(1, dtype=torch.float64, device=dev) if want_energy else None
        # residual / energy without the tangent need no CSR accumulators
        plan = self._vplan(dev, assemble=want_k)
        _lib.check(_lib.lib().tm_asm_nh(plan.struct, _lib.ptr(self._coords(dev)), _lib.ptr(u64), _lib.ptr(Et), Edt,
                                        E0, self.nu, _lib.ptr(kv), _lib.dtype_code(kv) if kv is not None else 0,
                                        _lib.ptr(R), _lib.ptr(en), _lib.stream_of(dev)))
        return kv, R, (en[0] if en is not None else None)

    def _tangent_action(self, u, E, lam, want_dE=False):
        dev = u.device
        u64 = u.detach().to(torch.float64).contiguous()
        Et, Edt, E0 = _e_arg(E.detach().to(dev) if torch.is_tensor(E) else E)
        y = torch.empty((self.n_points, 3), dtype=torch.float64, device=dev)
        dE = torch.zeros(self.topology.n_elements, dtype=torch.float64, device=dev) if want_dE else None
        plan = self._vplan(dev)
        _lib.check(_lib.lib().tm_nh_apply(plan.struct, _lib.ptr(self._coords(dev)), _lib.ptr(u64), _lib.ptr(Et), Edt,
                                          E0, self.nu, _lib.ptr(lam.reshape(-1, 3)), _lib.ptr(y), _lib.ptr(dE),
                                          _lib.stream_of(dev)))
        return y, dE

    # ---- public -----------------------------------------------------------------
    def energy(self, u, E=None):
        """Total strain energy ``sum_e vol_e psi(F_e)`` (differentiable w.r.t. u)."""
        E = self._resolve_E(E)
        if not u.is_cuda:
            if torch.is_tensor(E):
                raise NotImplementedError("per-element E on CPU tensors: use the GPU path")
            return super().energy(point_data={"displacement": u})
        if torch.is_tensor(E) and E.requires_grad:
            raise NotImplementedError("the energy is differentiable w.r.t. u; use residual() for dE")
        return _Energy.apply(u, self, E)

    def residual(self, u, E=None) -> torch.Tensor:
        """``R = dW/du`` as ``[N, 3]`` float64; backward = hand-written adjoint."""
        E = self._resolve_E(E)
        if torch.is_tensor(E):
            return _Residual.apply(u, E, self, 0.0)
        return _Residual.apply(u, None, self, float(E))

    def apply_tangent(self, u, lam, E=None) -> torch.Tensor:
        return self._tangent_action(u, self._resolve_E(E), lam.to(torch.float64).contiguous())[0]

    def tangent(self, u, E=None, value_dtype: Optional[torch.dtype] = None) -> SparseMatrix:
        """Tangent ``dR/du`` as a ``[3N, 3N]`` block matrix (3x3 blocks per scalar edge)."""
        kv, _, _ = self._launch(u, self._resolve_E(E), want_k=True, kdtype=value_dtype or self.value_dtype)
        return self._block_matrix(kv, u.device)

    def _block_matrix(self, kv, dev) -> SparseMatrix:
        return block_matrix(self, kv, dev)

    def newton(self, u0: torch.Tensor, fixed: torch.Tensor, E=None, rtol: float = 1e-8, max_iter: int = 30,
               lin_rtol: float = 1e-10, lin_maxiter: int = 20000, verbose: bool = False):
        """Solve ``R(u) = 0`` on the free nodes; ``u0`` must already carry the
        prescribed displacements on the ``fixed`` nodes (bool ``[N]``).

        Stops when ``||R_free|| <= rtol * ||R_free(u0)||``.  Returns ``(u, info)``.
        """
        dev = u0.device
        E = self._resolve_E(E)
        u = u0.detach().to(torch.float64).clone()
        free = ~fixed.to(dev)
        cond = BlockCondenser(self, free)
        _, R, _ = self._launch(u, E, want_r=True)
        r0 = R[free].norm().item()
        hist, lin_its = [1.0], []
        if r0 == 0.0:
            return u, dict(iterations=0, residual_history=hist, linear_iterations=lin_its, converged=True)
        for it in range(max_iter):
            kv, _, _ = self._launch(u, E, want_k=True, kdtype=self.value_dtype)
            vals, csr = cond.condense(kv)
            rhs = -R[free].reshape(-1)
            du, nit, rel, ok = pcg_raw(csr, vals, rhs, rtol=lin_rtol, maxiter=lin_maxiter, block_size=3)
            lin_its.append(nit)
            u[free] += du.view(-1, 3)
            _, R, _ = self._launch(u, E, want_r=True)
            hist.append(R[free].norm().item() / r0)
            if verbose:
                print(f"[newton] it {it + 1}: relres {hist[-1]:.3e}, pcg its {nit}")
            if hist[-1] <= rtol:
                break
        return u, dict(iterations=len(hist) - 1, residual_history=hist, linear_iterations=lin_its,
                       converged=hist[-1] <= rtol)


class BlockCondenser:
    """Node-level condensation of a 3x3 block tangent: keeps the blocks of
    free-free scalar edges in canonical order (the Condenser layout of
    ``operator/condense.py:182-220`` applied to nodes)."""

    def __init__(self, model: NeoHookeanModel, free: torch.Tensor):
        dev = free.device
        e0, e1 = model.edges[0].to(dev), model.edges[1].to(dev)
        keep = free[e0] & free[e1]
        nf = int(free.sum())
        local = torch.full((model.n_points,), -1, dtype=torch.long, device=dev)
        local[free] = torch.arange(nf, device=dev)
        rows, cols = local[e0[keep]], local[e1[keep]]
        rp = torch.zeros(nf + 1, dtype=torch.int32, device=dev)
        rp[1:] = torch.cumsum(torch.bincount(rows, minlength=nf), 0).to(torch.int32)
        self.csr = CSR(rp, cols.to(torch.int32), nf, nf)
        blocks = torch.nonzero(keep).flatten()
        idx = (blocks[:, None] * 9 + torch.arange(9, device=dev)[None]).reshape(-1)
        self.idx, self.idx_long = idx.to(torch.int32), idx
        self.n_free = nf

    def condense(self, kv):
        return Gather.apply(kv, self.idx, self.idx_long), self.csr

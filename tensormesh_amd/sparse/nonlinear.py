"""Newton solve of ``F(u, *params) = 0`` with the adjoint on the backward pass.

Drop-in for ``tensormesh.sparse.nonlinear_solve`` (``sparse/nonlinear_solve.py:21-175``):
``nonlinear_solve(f, j, u0, params, max_iter=100, tol=1e-6, verbose=False)`` with
``f(u, *params) -> residual`` and ``j(u, *params) -> SparseMatrix`` (``dF/du``).
The update is ``u <- u - J^-1 F(u)`` with the linear tolerance
``max(0.01 tol, 1e-12)``, stopping when ``||F(u)|| < tol`` (``:56-74``); the
solve's absolute floor is lowered to that tolerance as well (its default,
1e-10, would stop Newton short of tolerances below it).  The
linear solves are :meth:`SparseMatrix.solve`: Jacobi-PCG or BiCGSTAB on the
device for GPU tensors.  Backward (implicit function theorem, ``:95-124``):
``J^T lam = g``, then ``dL/dparams = -lam^T dF/dparams`` by one autograd VJP.
"""
from __future__ import annotations

from typing import Callable, Tuple

import torch

from .matrix import SparseMatrix


def _newton(f, j, u0, params, max_iter, tol, verbose):
    u = u0.clone()
    res_norm = None
    for i in range(max_iter):
        res = f(u, *params)
        res_norm = torch.linalg.vector_norm(res)
        if verbose:
            print(f"Iter {i}: |F(u)| = {float(res_norm):.6e}")
        if res_norm < tol:
            return u
        J = j(u, *params)
        lin = max(tol * 0.01, 1e-12)
        du = J.solve(res, tol=lin, atol=min(1e-10, lin))
        u = u - du.to(u.dtype)
    if verbose:
        print(f"Newton solver reached max_iter ({max_iter}) with residual {float(res_norm):.6e}")
    return u


class NonLinearSolveFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, j, u0, config, *params):
        with torch.no_grad():
            u = _newton(f, j, u0, params, **config)
        ctx.save_for_backward(u, *params)
        ctx.f, ctx.j = f, j
        return u

    @staticmethod
    def backward(ctx, g):
        u, *params = ctx.saved_tensors
        with torch.no_grad():
            J = ctx.j(u, *params)
        if not isinstance(J, SparseMatrix):
            raise TypeError(f"Jacobian function j must return a SparseMatrix, got {type(J)}")
        lam = J.T.solve(g)
        with torch.enable_grad():
            res = ctx.f(u.detach(), *params)
            grads = torch.autograd.grad(res, params, grad_outputs=-lam.to(res.dtype), allow_unused=True)
        return (None, None, None, None) + tuple(grads)


def nonlinear_solve(f: Callable[..., torch.Tensor], j: Callable[..., SparseMatrix], u0: torch.Tensor,
                    params: Tuple[torch.Tensor, ...], max_iter: int = 100, tol: float = 1e-6,
                    verbose: bool = False) -> torch.Tensor:
    """Solve ``F(u, params) = 0``; gradients reach ``params`` through one adjoint solve."""
    config = dict(max_iter=max_iter, tol=tol, verbose=verbose)
    return NonLinearSolveFunction.apply(f, j, u0, config, *params)

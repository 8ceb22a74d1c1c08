"""Float64 torch-CPU restatement of the reference pipeline, for the CPU
baseline only (timed beside the GPU path; never part of the product).

The reference runs this path as torch tensor ops, so the baseline is torch
too and uses the host's threads (``torch.set_num_threads``):

* pattern and ``P_E`` - SciPy canonical CSR (``element_assembler.py:838-884``),
  :func:`oracle.pattern.pattern_and_slots`;
* Map - ``X = points[conn]``, ``J = einsum(X, dphi)``, ``inv``/``det``,
  ``grad = einsum(dphi, J^-1)``, kappa interpolated to the quadrature points,
  ``K_e = sum_q jxw kappa_q grad_a . grad_b`` (``transformation.py:1236-1293``,
  ``element_assembler.py:402-466``), in element chunks so the ``[E, Q, nb, nb]``
  intermediate stays bounded (the reference's ``batch_size``);
* Reduce - ``values.index_add_(0, P_E, K_e)`` (``projector.py:95-130``);
* condensation - boolean inner / outer edge masks and ranks
  (``condense.py:182-309``);
* solve - Jacobi-PCG with a torch CSR product and FP64 dot products
  (``sparse/solve/torch_solve.py:24-128`` with the Jacobi diagonal).
"""
from __future__ import annotations

import numpy as np
import torch

from .assemble import rule


def assemble_kappa_laplace(points: torch.Tensor, conn: torch.Tensor, kappa: torch.Tensor, pe: torch.Tensor,
                           nnz: int, chunk: int = 1 << 18, order: int = 2) -> torch.Tensor:
    w, phi, G = rule("tetra" if conn.shape[1] == 4 else "triangle", order)
    w = torch.as_tensor(w, dtype=torch.float64)
    phi = torch.as_tensor(phi, dtype=torch.float64)
    G = torch.as_tensor(np.ascontiguousarray(G), dtype=torch.float64)        # [Q, d, nb]
    E, nb = conn.shape
    values = torch.zeros(nnz, dtype=torch.float64)
    for s in range(0, E, chunk):
        c = conn[s:s + chunk]
        X = points[c]                                                          # [e, nb, d]
        J = torch.einsum("ebj,qib->eqij", X, G)
        grad = torch.einsum("qib,eqji->eqbj", G, torch.linalg.inv(J))
        jxw = w[None] * torch.linalg.det(J).abs()
        kq = torch.einsum("eb,qb->eq", kappa[c], phi)
        ke = torch.einsum("eq,eqad,eqbd->eab", kq * jxw, grad, grad)
        values.index_add_(0, pe[s * nb * nb:(s + c.shape[0]) * nb * nb], ke.reshape(-1))
    return values


def condense(values: torch.Tensor, row: torch.Tensor, col: torch.Tensor, mask: torch.Tensor, rhs: torch.Tensor):
    inner = ~mask
    is_inner_edge = inner[row] & inner[col]
    is_ou2in = inner[row] & mask[col]
    local = torch.full((mask.shape[0],), -1, dtype=torch.int64)
    local[inner] = torch.arange(int(inner.sum()))
    local[mask] = torch.arange(int(mask.sum()))
    ni = int(inner.sum())
    kv, kr, kc = values[is_inner_edge], local[row[is_inner_edge]], local[col[is_inner_edge]]
    uo = torch.zeros(int(mask.sum()), dtype=values.dtype)      # zero Dirichlet data
    fi = rhs[inner].clone()
    fi.index_add_(0, local[row[is_ou2in]], -values[is_ou2in] * uo[local[col[is_ou2in]]])
    return kv, kr, kc, ni, fi


def jacobi_pcg(A: torch.Tensor, b: torch.Tensor, dinv: torch.Tensor, rtol: float = 1e-8, maxiter: int = 100000):
    """Returns ``(x, iterations)``."""
    x = torch.zeros_like(b)
    r = b.clone()
    z = dinv * r
    p = z.clone()
    rz = torch.dot(r, z)
    tol2 = (rtol * float(torch.linalg.vector_norm(b))) ** 2
    it = 0
    while float(torch.dot(r, r)) > tol2 and it < maxiter:
        q = A @ p
        alpha = rz / torch.dot(p, q)
        x += alpha * p
        r -= alpha * q
        z = dinv * r
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    return x, it

"""Isotropic linear elasticity on P1 simplices (float64), following the reference's route.

Test infrastructure only (the package never imports it).

``LinearElasticityElementAssembler.forward`` (``tensormesh/assemble/builtin.py:173-179``)
returns ``B_u^T C B_v`` per quadrature point, with the Voigt strain-displacement
matrix ``B = [diag(g); skew(g, sign=False)]`` of one basis gradient
(``functional/elasticity.py:132-186``, ``functional/ops.py:39-99``: 3D shear
rows (yz, xz, xy) = [0, g_z, g_y], [g_z, 0, g_x], [g_y, g_x, 0]; 2D row xy =
[g_y, g_x]) and Hooke's matrix ``C`` (``elasticity.py:188-255``: lam + 2 mu on
the normal diagonal, lam off it, mu on the shear diagonal).  The Map
(``element_assembler.py:371-398``) vmaps the u basis outside the v basis, the
pair layout (``:838-851``) puts the u basis on the row, and ``_integrate``
(``:196-221``) sums jxw; blocks are reduced per scalar edge in the
``from_block_coo`` order (``sparse/matrix.py:325-375``).
"""
from __future__ import annotations

import numpy as np

from .assemble import geometry


def lame(E, nu):
    E = np.asarray(E, np.float64)
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def voigt_B(g: np.ndarray) -> np.ndarray:
    """``[..., d]`` basis gradients -> ``[..., 3 | 6, d]`` strain-displacement matrices."""
    d = g.shape[-1]
    z = np.zeros(g.shape[:-1])
    if d == 2:
        gx, gy = g[..., 0], g[..., 1]
        rows = [[gx, z], [z, gy], [gy, gx]]
    else:
        gx, gy, gz = g[..., 0], g[..., 1], g[..., 2]
        rows = [[gx, z, z], [z, gy, z], [z, z, gz], [z, gz, gy], [gz, z, gx], [gy, gx, z]]
    return np.stack([np.stack(r, -1) for r in rows], -2)


def voigt_C(E: float, nu: float, d: int) -> np.ndarray:
    mu, lam = lame(E, nu)
    nv = 3 if d == 2 else 6
    C = np.zeros((nv, nv))
    C[:d, :d] = lam
    C[np.arange(nv), np.arange(nv)] += mu
    C[np.arange(d), np.arange(d)] += mu
    return C


def element_stiffness(points, conn, element_type: str, E, nu=0.3, order: int = 2) -> np.ndarray:
    """``[E, nb, nb, d, d]``: block (a, b) = sum_q jxw B_a^T C B_b; ``E`` scalar or per element."""
    _, grad, jxw = geometry(points, conn, element_type, order)      # [E, Q, nb, d], [E, Q]
    d = grad.shape[-1]
    B = voigt_B(grad)                                                # [E, Q, nb, nv, d]
    Es = np.broadcast_to(np.asarray(E, np.float64), (grad.shape[0],))
    C1 = voigt_C(1.0, nu, d)                                         # C is linear in E
    K = np.einsum("eqapi,pr,eqbrk,eq->eabik", B, C1, B, jxw)
    return K * Es[:, None, None, None, None]


def stiffness_blocks(points, conn, element_type, E, pe, nnz, nu=0.3) -> np.ndarray:
    """Assembled stiffness ``[nnz, d, d]`` on the scalar pattern (``pe`` = element slots)."""
    ke = element_stiffness(points, conn, element_type, E, nu)
    d = ke.shape[-1]
    ke = ke.reshape(-1, d * d)
    out = np.zeros((nnz, d * d))
    for c in range(d * d):
        out[:, c] = np.bincount(pe, weights=ke[:, c], minlength=nnz)
    return out.reshape(nnz, d, d)


def dense(blocks, eu, ev, n_points) -> np.ndarray:
    """Dense ``[N d, N d]`` matrix of the blocks (row = node d + i, col = node d + k)."""
    d = blocks.shape[-1]
    A = np.zeros((n_points * d, n_points * d))
    for i in range(d):
        for k in range(d):
            np.add.at(A, (eu * d + i, ev * d + k), blocks[:, i, k])
    return A

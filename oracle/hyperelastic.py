"""Compressible neo-Hookean energy, residual and tangent on P1 tetrahedra (float64).

Density (``tensormesh/assemble/builtin.py:285-315``)::

    F = I + grad u,  J = det F,  I1 = tr(F^T F)
    psi = mu/2 (I1 - d) - mu ln J + lam/2 (ln J)^2

with ``mu = E / (2 (1 + nu))``, ``lam = E nu / ((1 + nu)(1 - 2 nu))``
(``:281-283``) and ``E`` given per element (``element_data``,
``element_assembler.py:604-618``).  Energy is ``sum_e sum_q jxw psi``
(``element_assembler.py:676-678``).  Closed forms (Bonet & Wood,
*Nonlinear Continuum Mechanics for FEA*, 2nd ed., 6.4)::

    P = mu F + (lam ln J - mu) F^-T
    A_ijkl = mu d_ik d_jl + lam Finv_lk Finv_ji - (lam ln J - mu) Finv_jk Finv_li

Residual ``R_(a,i) = sum_e vol_e P_ij grad_a_j``; tangent block for nodes
(a, b) is ``vol_e A_ijkl grad_a_j grad_b_l`` at ``[i, k]``, stored per
scalar edge as a 3x3 block in the order of ``SparseMatrix.from_block_coo``
(``sparse/matrix.py:325-375``: row = node*3 + i, col = node*3 + k).
"""
from __future__ import annotations

import numpy as np

from .assemble import geometry


def lame(E, nu):
    E = np.asarray(E, np.float64)
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def element_kinematics(points, conn, u):
    phi, grad, jxw = geometry(points, conn, "tetra", 2)
    g = grad[:, 0]                                       # P1: constant per element [E, 4, 3]
    vol = jxw.sum(1)
    H = np.einsum("ebi,ebj->eij", np.asarray(u, np.float64)[np.asarray(conn)], g)
    F = np.eye(3)[None] + H
    return g, vol, F


def psi(F, mu, lam):
    J = np.linalg.det(F)
    lj = np.log(J)
    return mu / 2 * ((F * F).sum((1, 2)) - 3.0) - mu * lj + lam / 2 * lj ** 2


def energy(points, conn, u, E, nu=0.3):
    g, vol, F = element_kinematics(points, conn, u)
    mu, lam = lame(E, nu)
    return float((vol * psi(F, mu, lam)).sum())


def piola(F, mu, lam):
    Finv = np.linalg.inv(F)
    lj = np.log(np.linalg.det(F))
    return mu[:, None, None] * F + (lam * lj - mu)[:, None, None] * np.transpose(Finv, (0, 2, 1))


def dpdf(F, mu, lam):
    Finv = np.linalg.inv(F)
    lj = np.log(np.linalg.det(F))
    I = np.eye(3)
    A = mu[:, None, None, None, None] * np.einsum("ik,jl->ijkl", I, I)[None]
    A = A + lam[:, None, None, None, None] * np.einsum("elk,eji->eijkl", Finv, Finv)
    A = A - (lam * lj - mu)[:, None, None, None, None] * np.einsum("ejk,eli->eijkl", Finv, Finv)
    return A


def element_residual(points, conn, u, E, nu=0.3):
    g, vol, F = element_kinematics(points, conn, u)
    mu, lam = lame(E, nu)
    P = piola(F, mu, lam)
    return np.einsum("e,eij,eaj->eai", vol, P, g)           # [E, 4, 3]


def element_tangent(points, conn, u, E, nu=0.3):
    g, vol, F = element_kinematics(points, conn, u)
    mu, lam = lame(E, nu)
    A = dpdf(F, mu, lam)
    return np.einsum("e,eijkl,eaj,ebl->eabik", vol, A, g, g)  # [E, 4, 4, 3, 3]


def residual(points, conn, u, E, n_points, nu=0.3):
    re = element_residual(points, conn, u, E, nu)
    out = np.zeros((n_points, 3))
    np.add.at(out, np.asarray(conn).ravel(), re.reshape(-1, 3))
    return out


def tangent_blocks(points, conn, u, E, pe, nnz, nu=0.3):
    """Assembled tangent as ``[nnz, 3, 3]`` blocks on the scalar pattern."""
    ke = element_tangent(points, conn, u, E, nu).reshape(-1, 9)
    out = np.zeros((nnz, 9))
    for c in range(9):
        out[:, c] = np.bincount(pe, weights=ke[:, c], minlength=nnz)
    return out.reshape(nnz, 3, 3)


def dresidual_dE(points, conn, u, E, lam_adj, nu=0.3):
    """``d(lam_adj . R)/dE_e``; R_e is linear in E_e (mu, lam proportional to E)."""
    re = element_residual(points, conn, u, np.ones_like(np.asarray(E, np.float64)), nu)
    la = np.asarray(lam_adj, np.float64)[np.asarray(conn)]   # [E, 4, 3]
    return np.einsum("eai,eai->e", re, la)

"""Float64 reference solvers for the condensed systems.

* :func:`jacobi_pcg` - textbook preconditioned CG (Saad, *Iterative Methods
  for Sparse Linear Systems*, 2nd ed., Alg. 9.1) with ``M = diag(A)``;
  the diagonal follows ``torch_solve.py:24-59`` (missing diagonal -> 1).
  Stops when ``||r_k||_2 <= rtol * ||b||_2``.
* :func:`bicgstab` - Jacobi right-preconditioned BiCGSTAB in the order of
  the reference's ``bicgstab_py`` (``sparse/solve/torch_solve.py:130-248``:
  ``p_hat = M^-1 p``, ``v = A p_hat``, ``s = r - alpha v``, ``s_hat = M^-1 s``,
  ``t = A s_hat``, ``omega = (t,s)/(t,t)``, ``beta = (rho'/rho)(alpha/omega)``),
  with the stopping rule ``||r|| <= max(rtol ||b||, atol)`` tested after each
  full step.  Pinned against :func:`direct` on non-symmetric systems and by
  reducing to one step on a diagonal system (tests/test_oracle.py).
* :func:`direct` - ``scipy.sparse.linalg.spsolve`` (SuperLU), the CPU
  backend the reference's quickstart reports (``README.md:132``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse
import scipy.sparse.linalg


def csr(values, row, col, n):
    return scipy.sparse.csr_matrix((np.asarray(values, np.float64), (row, col)), shape=(n, n))


def diagonal(A: scipy.sparse.csr_matrix):
    """Jacobi diagonal: a missing entry is 1 (``torch_solve.py:24-59``) and so
    is a stored one below 1e-12 in magnitude (``torch_solve.py:177-182``)."""
    d = np.ones(A.shape[0])
    coo = A.tocoo()
    m = coo.row == coo.col
    d[coo.row[m]] = 0.0
    np.add.at(d, coo.row[m], coo.data[m])
    d[np.abs(d) < 1e-12] = 1.0
    return d


def jacobi_pcg(A, b, rtol=1e-8, maxiter=20000, x0=None):
    """Return ``(x, iterations, relres_history)``."""
    b = np.asarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
    dinv = 1.0 / diagonal(A)
    r = b - A @ x
    z = dinv * r
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    hist = [np.linalg.norm(r) / bn]
    it = 0
    while hist[-1] > rtol and it < maxiter:
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = dinv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
        hist.append(np.linalg.norm(r) / bn)
    return x, it, np.asarray(hist)


def bicgstab(A, b, rtol=1e-8, atol=0.0, maxiter=20000, x0=None):
    """Return ``(x, iterations, relres)``."""
    b = np.asarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
    minv = 1.0 / diagonal(A)
    r = b - A @ x
    r0h = r.copy()
    p = r.copy()
    rho = r0h @ r
    bn = np.linalg.norm(b)
    tol = max(rtol * bn, atol)
    it = 0
    while np.linalg.norm(r) > tol and it < maxiter:
        phat = minv * p
        v = A @ phat
        alpha = rho / (r0h @ v)
        s = r - alpha * v
        shat = minv * s
        t = A @ shat
        tt = t @ t
        omega = (t @ s) / tt if tt != 0.0 else 0.0
        x += alpha * phat + omega * shat
        r = s - omega * t
        it += 1
        rho_new = r0h @ r
        if omega == 0.0:
            break
        p = r + (rho_new / rho) * (alpha / omega) * (p - omega * v)
        rho = rho_new
    return x, it, np.linalg.norm(r) / bn if bn > 0 else 0.0


def direct(A, b):
    return scipy.sparse.linalg.spsolve(A.tocsc(), np.asarray(b, np.float64))


def relres(A, x, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(b - A @ np.asarray(x, np.float64)) / np.linalg.norm(b))

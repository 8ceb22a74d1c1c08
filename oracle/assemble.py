"""Float64 Map-Reduce assembly, restated from the reference's tensor sequence.

Geometry (``tensormesh/element/element.py:958-966``,
``tensormesh/element/transformation.py:1236-1293``)::

    X    = points[conn]                           # [E, nb, d]
    J    = einsum("ebj,qib->eqij", X, G)          # G[q, i, b] = d phi_b / d xi_i
    Jinv = inv(J)
    grad = einsum("qib,eqji->eqbj", G, Jinv)      # physical gradients
    jxw  = w_q * |det J|

Map (``element_assembler.py:402-460``): the integrand of each form is
evaluated at every (element, quadrature point, a, b) and summed with jxw
(``_integrate``, ``:196-221``).  Reduce (``projector.py:95-130``):
``values[P_E[k]] += local[k]`` in float64.  Load vectors follow
``node_assembler.py:242-409`` with projector indices ``conn.flatten()``
(``:665-670``).
"""
from __future__ import annotations

import numpy as np

# Quadrature rules of tensormesh/element/quadrature.py:93,102 (triangle) and
# :143,152 (tetrahedron), orders 1 and 2; reference node order phi_0 = 1 - sum(xi).
_A, _B = 0.5854101966249685, 0.1381966011250105
_RULES = {
    ("triangle", 1): ([0.5], [[1 / 3, 1 / 3]]),
    ("triangle", 2): ([1 / 6] * 3, [[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]),
    # order 3 (quadrature.py:94,103; Strang-Fix 4-point rule, one negative weight)
    ("triangle", 3): ([-27 / 96, 25 / 96, 25 / 96, 25 / 96],
                      [[1 / 3, 1 / 3], [0.2, 0.6], [0.6, 0.2], [0.2, 0.2]]),
    ("tetra", 1): ([1 / 6], [[0.25, 0.25, 0.25]]),
    ("tetra", 2): ([1 / 24] * 4, [[_A, _B, _B], [_B, _B, _B], [_B, _B, _A], [_B, _A, _B]]),
    # order 3 (quadrature.py:144,153; Keast 5-point rule)
    ("tetra", 3): ([-2 / 15, 3 / 40, 3 / 40, 3 / 40, 3 / 40],
                   [[0.25, 0.25, 0.25], [0.5, 1 / 6, 1 / 6], [1 / 6, 1 / 6, 1 / 6], [1 / 6, 1 / 6, 0.5],
                    [1 / 6, 0.5, 1 / 6]]),
}


def rule(element_type: str, order: int):
    w, q = _RULES[(element_type, order)]
    q = np.asarray(q, dtype=np.float64)
    phi = np.concatenate([1.0 - q.sum(1, keepdims=True), q], 1)        # [Q, nb]
    d = q.shape[1]
    dphi = np.concatenate([-np.ones((d, 1)), np.eye(d)], 1)              # [i, b]
    return np.asarray(w, dtype=np.float64), phi, np.broadcast_to(dphi, (len(w), d, d + 1))


def geometry(points: np.ndarray, conn: np.ndarray, element_type: str, order: int = 2):
    """Return ``(shape_val [Q, nb], grad [E, Q, nb, d], jxw [E, Q])`` in float64."""
    w, phi, G = rule(element_type, order)
    X = np.asarray(points, dtype=np.float64)[np.asarray(conn)]
    J = np.einsum("ebj,qib->eqij", X, G)
    Jinv = np.linalg.inv(J)
    grad = np.einsum("qib,eqji->eqbj", G, Jinv)
    jxw = w[None, :] * np.abs(np.linalg.det(J))
    return phi, grad, jxw


def element_matrices(points, conn, element_type: str, form: str, kappa=None, order: int = 2,
                     chunk: int = 1 << 18):
    """Per-element matrices ``[E, nb, nb]`` for ``form`` in {laplace, kappa_laplace, mass}.

    ``kappa`` is nodal (``point_data``), interpolated to quadrature points
    as ``einsum("eb,qb->eq", kappa[conn], shape_val)`` (``element_assembler.py:440``).
    Elements are processed in chunks to bound the ``[E, Q, nb, nb]`` intermediate.
    """
    conn = np.asarray(conn)
    E, nb = conn.shape
    out = np.empty((E, nb, nb), dtype=np.float64)
    for s in range(0, E, chunk):
        c = conn[s:s + chunk]
        phi, grad, jxw = geometry(points, c, element_type, order)
        if form == "laplace":
            f = np.einsum("eqad,eqbd->eqab", grad, grad)
        elif form == "kappa_laplace":
            kq = np.einsum("eb,qb->eq", np.asarray(kappa, dtype=np.float64)[c], phi)
            f = kq[:, :, None, None] * np.einsum("eqad,eqbd->eqab", grad, grad)
        elif form == "mass":
            f = np.broadcast_to(np.einsum("qa,qb->qab", phi, phi)[None], (c.shape[0],) + (phi.shape[0], nb, nb))
        else:
            raise ValueError(form)
        out[s:s + chunk] = np.einsum("eqab,eq->eab", f, jxw)
    return out


def element_matrices_general(points, conn, element_type: str, D, order: int = 2) -> np.ndarray:
    """Per-element ``[E, nb, nb]`` of the constant bilinear integrand ``U_a^T D V_b``.

    At every quadrature point ``U_a = (phi_a(q), grad phi_a)`` for the u basis
    function ``a`` (the row, ``element_assembler.py:371-398, 838-851``) and
    ``V_b`` likewise for the v basis function ``b``; the products are summed
    with jxw (``_integrate``, ``:196-221``).
    """
    conn = np.asarray(conn)
    phi, grad, jxw = geometry(points, conn, element_type, order)            # [Q, nb], [E, Q, nb, d], [E, Q]
    E, Q, nb, d = grad.shape
    U = np.concatenate([np.broadcast_to(phi[None, :, :, None], (E, Q, nb, 1)), grad], -1)   # [E, Q, nb, d+1]
    D = np.asarray(D, dtype=np.float64)
    return np.einsum("eqai,ij,eqbj,eq->eab", U, D, U, jxw)


def reduce(pe: np.ndarray, local: np.ndarray, nnz: int) -> np.ndarray:
    """``index_add_`` of ``local.ravel()`` into ``nnz`` slots (float64 accumulator)."""
    return np.bincount(np.asarray(pe), weights=np.asarray(local, dtype=np.float64).ravel(),
                       minlength=nnz)


def load_vector(points, conn, element_type: str, n_points: int, f=None, c=None, order: int = 2):
    """``b_a = sum_e sum_q w_q |J| g(q) phi_a(q)`` for ``g = f`` (nodal) or constant ``c``."""
    conn = np.asarray(conn)
    phi, grad, jxw = geometry(points, conn, element_type, order)
    if f is not None:
        fq = np.einsum("eb,qb->eq", np.asarray(f, dtype=np.float64)[conn], phi)
        be = np.einsum("eq,qa,eq->ea", fq, phi, jxw)
    else:
        be = float(c) * np.einsum("qa,eq->ea", phi, jxw)
    return np.bincount(conn.ravel(), weights=be.ravel(), minlength=n_points)


def kappa_gradient(points, conn, element_type, g_values: np.ndarray, pe: np.ndarray, n_points: int,
                   order: int = 2):
    """d(sum_s g_s K_s)/d kappa for the kappa-Laplace form (closed form, float64).

    ``K_s = sum_{e,a,b: P_E=s} sum_q w_q |J| kappa_q grad_a.grad_b`` is linear in
    nodal kappa with ``dK_s/dkappa_c = sum_q w_q |J| phi_c(q) grad_a.grad_b``.
    """
    conn = np.asarray(conn)
    E, nb = conn.shape
    phi, grad, jxw = geometry(points, conn, element_type, order)
    gg = np.asarray(g_values, dtype=np.float64)[np.asarray(pe)].reshape(E, nb, nb)
    dots = np.einsum("eqad,eqbd->eqab", grad, grad)
    s = np.einsum("eab,eqab,eq->eq", gg, dots, jxw)      # [E, Q]
    dk = np.einsum("eq,qc->ec", s, phi)
    return np.bincount(conn.ravel(), weights=dk.ravel(), minlength=n_points)


def _at_quadrature(points, conn, element_type, order, fields):
    """Geometry plus ``x`` and nodal fields (and their gradients, key ``grad<name>``)
    interpolated to the quadrature points (``element_assembler.py:416-457``)."""
    conn = np.asarray(conn)
    phi, grad, jxw = geometry(points, conn, element_type, order)
    X = np.asarray(points, dtype=np.float64)[conn]
    at = {"x": np.einsum("ebd,qb->eqd", X, phi)}
    for k, v in (fields or {}).items():
        v = np.asarray(v, dtype=np.float64)[conn]                          # [E, nb, ...]
        at[k] = np.einsum("eb...,qb->eq...", v, phi)
        at["grad" + k] = np.einsum("eb...,eqbd->eq...d", v, grad)
    return phi, grad, jxw, at


def element_matrices_quadrature(points, conn, element_type: str, integrand, order: int = 2, fields=None,
                                element_data=None) -> np.ndarray:
    """Per-element ``[E, nb, nb]`` for an integrand evaluated at every (element,
    quadrature point, a, b) - the reference's Map (``element_assembler.py:371-466``)
    without vmap: ``integrand(u, gradu, v, gradv, at)`` receives ``u``/``v``
    shaped ``[E, Q, nb, nb]``, ``gradu``/``gradv`` ``[E, Q, nb, nb, d]`` (a = row
    basis function u, b = column basis function v) and ``at`` = the
    coefficients at the quadrature points broadcast to ``[E, Q, 1, 1, ...]``
    (``x``, ``fields``, ``grad<field>``; ``element_data`` per element)."""
    phi, grad, jxw, at = _at_quadrature(points, conn, element_type, order, fields)
    E, Q, nb, d = grad.shape
    u = np.broadcast_to(phi[None, :, :, None], (E, Q, nb, nb))
    v = np.broadcast_to(phi[None, :, None, :], (E, Q, nb, nb))
    gu = np.broadcast_to(grad[:, :, :, None, :], (E, Q, nb, nb, d))
    gv = np.broadcast_to(grad[:, :, None, :, :], (E, Q, nb, nb, d))
    ab = {k: a.reshape(a.shape[:2] + (1, 1) + a.shape[2:]) for k, a in at.items()}
    for k, a in (element_data or {}).items():
        a = np.asarray(a, dtype=np.float64)
        ab[k] = a.reshape((E, 1, 1, 1) + a.shape[1:])
    f = integrand(u, gu, v, gv, ab)
    return np.einsum("eqab,eq->eab", np.broadcast_to(f, (E, Q, nb, nb)), jxw)


def load_vector_quadrature(points, conn, element_type: str, n_points: int, integrand, order: int = 2,
                           fields=None) -> np.ndarray:
    """``b_a = sum_e sum_q w_q |J| g(v = phi_a, gradv = grad phi_a, coefficients at q)``
    (``node_assembler.py:242-409``, projector indices ``conn.flatten()``):
    ``integrand(v [E, Q, nb], gradv [E, Q, nb, d], at)`` with ``at`` broadcast to
    ``[E, Q, 1, ...]``."""
    conn = np.asarray(conn)
    phi, grad, jxw, at = _at_quadrature(points, conn, element_type, order, fields)
    E, Q, nb, d = grad.shape
    v = np.broadcast_to(phi[None], (E, Q, nb))
    ab = {k: a.reshape(a.shape[:2] + (1,) + a.shape[2:]) for k, a in at.items()}
    g = np.broadcast_to(integrand(v, grad, ab), (E, Q, nb))
    be = np.einsum("eqa,eq->ea", g, jxw)
    return np.bincount(conn.ravel(), weights=be.ravel(), minlength=n_points)

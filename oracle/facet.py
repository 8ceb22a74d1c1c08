"""Float64 boundary-facet load vectors, restated from the reference's
FacetAssembler (``tensormesh/assemble/facet_assembler.py:123-277, 368-515``).

* Selection: facet ``k`` of element ``e`` (the face opposite local vertex
  ``k``) is integrated iff every one of its nodes is flagged (``:480-496``).
* Quadrature: a facet rule in facet-local coordinates pushed into the cell's
  reference coordinates; the cell basis is evaluated there (all ``nb``
  functions, zero off the facet), ``x`` and point data are interpolated with
  it, and the measure is ``w_q |F|`` with ``|F| = sqrt(det(F^T F))`` for the
  physical facet Jacobian ``F`` (``transformation.py:990-1112``).  Edges use
  Gauss-Legendre on [0, 1] with ``order`` points (``numpy.polynomial.legendre``,
  weights halved), faces the triangle rules of :mod:`oracle.assemble`.
* Reduce: ``b[cell nodes] += sum_q w_q |F| g(q)`` (``:245-277``).
"""
from __future__ import annotations

import numpy as np

from .assemble import rule as cell_rule


def select(conn: np.ndarray, mask: np.ndarray):
    conn = np.asarray(conn)
    nb = conn.shape[1]
    out = []
    for k in range(nb):
        others = [a for a in range(nb) if a != k]
        for e in np.nonzero(np.asarray(mask)[conn[:, others]].all(1))[0]:
            out.append((int(e), k))
    return out


def facet_rule(element_type: str, order: int):
    """``(weights [Q], mu [Q, d])``: facet barycentric coordinates."""
    if element_type == "triangle":
        p, w = np.polynomial.legendre.leggauss(order)
        t = (p + 1.0) / 2.0
        return w / 2.0, np.stack([1.0 - t, t], 1)
    w, phi, _ = cell_rule("triangle", order)
    return w, phi


def load_vector(points, conn, element_type: str, mask, integrand, order: int = 2, fields=None, H: int = 1):
    """``integrand(v [Q, nb], gradv [Q, nb, d], at)`` for one facet, ``at`` with
    ``x`` [Q, d], every field [Q, ...] and ``grad<field>``; returns ``[Q, nb]``
    or ``[Q, nb, H]``.  Output: ``b`` of length ``N * H`` (node-major)."""
    points = np.asarray(points, dtype=np.float64)
    conn = np.asarray(conn)
    nb = conn.shape[1]
    d = nb - 1
    ref_vertices = np.concatenate([np.zeros((1, d)), np.eye(d)], 0)      # cell reference coordinates
    w, mu = facet_rule(element_type, order)
    b = np.zeros((len(points), H))
    for e, k in select(conn, mask):
        fv = [a for a in range(nb) if a != k]
        xi = mu @ ref_vertices[fv]                                        # [Q, d] in the cell's reference space
        phi = np.concatenate([1.0 - xi.sum(1, keepdims=True), xi], 1)     # [Q, nb]
        X = points[conn[e]]                                               # [nb, d]
        J = (X[1:] - X[0]).T
        grad = np.concatenate([-np.ones((1, d)), np.eye(d)], 0) @ np.linalg.inv(J)   # [nb, d]
        Fj = np.stack([X[fv[i]] - X[fv[0]] for i in range(1, d)], 1)      # [d, d-1] physical facet Jacobian
        detF = np.sqrt(np.linalg.det(Fj.T @ Fj))
        at = {"x": phi @ X}
        for key, val in (fields or {}).items():
            v = np.asarray(val, dtype=np.float64)[conn[e]]
            at[key] = np.einsum("b...,qb->q...", v, phi)
            at["grad" + key] = np.broadcast_to(np.einsum("b...,bd->...d", v, grad),
                                               (len(w),) + v.shape[1:] + (d,))
        g = np.asarray(integrand(phi, np.broadcast_to(grad, (len(w), nb, d)), at), dtype=np.float64)
        g = g.reshape(len(w), nb, H)
        np.add.at(b, conn[e], np.einsum("qah,q->ah", g, w * detF))
    return b.reshape(-1)

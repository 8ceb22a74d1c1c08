"""P1 reference elements: quadrature rules, shape values and reference gradients.

Only the linear simplices used on the assembly path are provided
(``triangle`` and ``tetra``).  The quadrature tables are the standard
published symmetric rules (Strang-Fix for triangles, Keast for tetrahedra)
and are the same rules the reference ships for orders 1-3
(``tensormesh/element/quadrature.py:68-116`` triangle,
``tensormesh/element/quadrature.py:118-166`` tetrahedron): weights sum to
the reference measure (1/2, 1/6).

Node ordering follows the reference P1 elements
(``tensormesh/element/element.py:1673-1740`` Triangle,
``tensormesh/element/element.py:1936-2016`` Tetrahedron): the vertices in
the order (0,0[,0]), e_x, e_y[, e_z], so ``phi_0 = 1 - sum(xi)`` and
``phi_k = xi_{k-1}``.
"""
from __future__ import annotations

import numpy as np

ELEMENT_DIM = {"triangle": 2, "tetra": 3}
ELEMENT_NB = {"triangle": 3, "tetra": 4}

_TRI_RULES = {
    1: ([0.5], [[1.0 / 3.0, 1.0 / 3.0]]),
    2: ([1.0 / 6.0] * 3,
        [[1.0 / 6.0, 1.0 / 6.0], [2.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 2.0 / 3.0]]),
    3: ([-0.28125, 0.260416666666667, 0.260416666666667, 0.260416666666667],
        [[1.0 / 3.0, 1.0 / 3.0], [0.2, 0.6], [0.6, 0.2], [0.2, 0.2]]),
}

_A = 0.5854101966249685
_B = 0.1381966011250105
_TET_RULES = {
    1: ([1.0 / 6.0], [[0.25, 0.25, 0.25]]),
    2: ([1.0 / 24.0] * 4,
        [[_A, _B, _B], [_B, _B, _B], [_B, _B, _A], [_B, _A, _B]]),
    3: ([-2.0 / 15.0, 0.075, 0.075, 0.075, 0.075],
        [[0.25, 0.25, 0.25], [0.5, 1.0 / 6.0, 1.0 / 6.0], [1.0 / 6.0] * 3,
         [1.0 / 6.0, 1.0 / 6.0, 0.5], [1.0 / 6.0, 0.5, 1.0 / 6.0]]),
}


def quadrature(element_type: str, order: int = 2):
    """Return ``(weights [Q], points [Q, d])`` as float64 numpy arrays."""
    table = {"triangle": _TRI_RULES, "tetra": _TET_RULES}[element_type]
    if order not in table:
        raise ValueError(f"quadrature order {order} not available for {element_type} "
                         f"(have {sorted(table)})")
    w, p = table[order]
    return np.asarray(w, dtype=np.float64), np.asarray(p, dtype=np.float64)


def shape_val(element_type: str, points: np.ndarray) -> np.ndarray:
    """P1 shape values at reference points: ``[Q, nb]``."""
    p = np.asarray(points, dtype=np.float64)
    return np.concatenate([1.0 - p.sum(1, keepdims=True), p], axis=1)


def ref_grad(element_type: str) -> np.ndarray:
    """Constant P1 reference gradients ``d phi_b / d xi_i`` as ``[nb, d]``."""
    d = ELEMENT_DIM[element_type]
    return np.concatenate([-np.ones((1, d)), np.eye(d)], axis=0)


def quadrature_moments(element_type: str, order: int = 2):
    """Moments of the rule used by the P1 kernels.

    Returns ``(wsum, W, M)`` with ``wsum = sum_q w_q``,
    ``W[c] = sum_q w_q phi_c(q)`` and ``M[a, b] = sum_q w_q phi_a(q) phi_b(q)``.
    For P1 simplices ``grad phi`` and ``|det J|`` are constant per element, so
    every supported form reduces to these three reference-element moments
    times per-element geometry.
    """
    w, p = quadrature(element_type, order)
    phi = shape_val(element_type, p)
    wsum = float(w.sum())
    W = (w[:, None] * phi).sum(0)
    M = np.einsum("q,qa,qb->ab", w, phi, phi)
    return wsum, W, M


# Gauss-Legendre rules on [0, 1] for the edges of triangles, indexed like the
# reference's lin_quadrature (tensormesh/element/quadrature.py:13-66): "order"
# n is the number of points (exact to degree 2n - 1), weights sum to 1.
_GL = {
    1: ([2.0], [0.0]),
    2: ([1.0, 1.0], [-1.0 / np.sqrt(3.0), 1.0 / np.sqrt(3.0)]),
    3: ([5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0], [-np.sqrt(0.6), 0.0, np.sqrt(0.6)]),
}


def facet_quadrature(element_type: str, order: int = 2):
    """Quadrature on the facets of a P1 element as ``(weights [Q], bary [Q, d])``:
    barycentric coordinates of each point with respect to the facet's ``d``
    vertices.  Triangle facets (edges) use the Gauss-Legendre rule above
    (weights sum to 1, the edge Jacobian is its length); tetrahedron facets use
    the triangle rules of :func:`quadrature` (weights sum to 1/2, the face
    Jacobian is twice its area) - the facet rules of
    ``tensormesh/element/quadrature.py:321-410``."""
    if element_type == "triangle":
        if order not in _GL:
            raise ValueError(f"facet quadrature order {order} not available for triangle (have {sorted(_GL)})")
        w, p = _GL[order]
        t = (np.asarray(p, dtype=np.float64) + 1.0) / 2.0
        return np.asarray(w, dtype=np.float64) / 2.0, np.stack([1.0 - t, t], 1)
    w, p = quadrature("triangle", order)
    return w, np.concatenate([1.0 - p.sum(1, keepdims=True), p], 1)

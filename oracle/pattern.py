"""Sparsity pattern and element->edge map, restated with SciPy.

Follows ``tensormesh/assemble/element_assembler.py:838-861``: all
``(conn[e, i], conn[e, j])`` pairs in element-major, then i, then j order,
deduplicated through ``coo_matrix(...).tocsr().tocoo()`` (canonical CSR:
rows ascending, columns ascending and unique per row), and
``P_E[e*nb*nb + i*nb + j]`` = position of that pair in the edge list.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse


def element_pairs(conn: np.ndarray):
    conn = np.asarray(conn)
    E, nb = conn.shape
    u = np.repeat(conn, nb, axis=1).reshape(-1)        # conn[e, i] repeated over j
    v = np.tile(conn, (1, nb)).reshape(-1)             # conn[e, j]
    return u, v


def pattern_and_slots(conn: np.ndarray, n_points: int):
    """Return ``(edge_u, edge_v, P_E)``: int32 edges in canonical CSR order, int64 slots."""
    u, v = element_pairs(conn)
    tmp = scipy.sparse.coo_matrix((np.ones_like(u), (u, v)),
                                  shape=(n_points, n_points)).tocsr().tocoo()
    edge_u, edge_v = tmp.row, tmp.col
    nnz = len(edge_u)
    eids = scipy.sparse.coo_matrix((np.arange(nnz), (edge_u, edge_v)),
                                   shape=(n_points, n_points)).tocsr()
    pe = np.asarray(eids[u, v]).ravel().astype(np.int64)
    return edge_u, edge_v, pe


def csr_from_edges(edge_u: np.ndarray, n_points: int):
    """Row pointer of a canonical-order edge list."""
    counts = np.bincount(edge_u, minlength=n_points)
    rp = np.zeros(n_points + 1, dtype=np.int64)
    np.cumsum(counts, out=rp[1:])
    return rp


def brute_force_pairs(conn: np.ndarray):
    """Set of all node pairs that share an element (pure-Python, tiny meshes)."""
    s = set()
    for row in np.asarray(conn).tolist():
        for a in row:
            for b in row:
                s.add((a, b))
    return s


def kuhn_nnz(n: int) -> int:
    """Closed-form nnz of the P1 pattern on the ``n^3`` Kuhn grid.

    Edges of the Kuhn split: the 3 axis directions, the 3 face diagonals
    (one per face orientation) and the main diagonal, each counted on the
    lattice: 3 n (n+1)^2 axis edges, 3 n^2 (n+1) face diagonals and n^3 main
    diagonals.  nnz = N + 2 * edges.
    """
    N = (n + 1) ** 3
    edges = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    return N + 2 * edges

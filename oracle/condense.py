"""Static condensation, restated from ``tensormesh/operator/condense.py``.

Layout (``_compute_layout``, ``:182-220``): inner DOFs are ``~mask``; an
edge is inner-inner when both ends are inner and ``ou2in`` when the row is
inner and the column outer; local ids are the rank of a DOF among the inner
(resp. outer) DOFs in increasing global order.  ``__call__`` (``:238-309``)
keeps the edge order (so canonical CSR order is preserved) and returns
``(K_ii, f_i - K_io u_o)``; ``recover`` (``:429-458``) scatters back and
writes the Dirichlet values.
"""
from __future__ import annotations

import numpy as np


def layout(row: np.ndarray, col: np.ndarray, mask: np.ndarray):
    mask = np.asarray(mask, dtype=bool)
    inner = ~mask
    is_inner_edge = inner[row] & inner[col]
    is_ou2in_edge = inner[row] & mask[col]
    local = np.full(mask.shape[0], -1, dtype=np.int64)
    local[inner] = np.arange(int(inner.sum()))
    local[mask] = np.arange(int(mask.sum()))
    return dict(
        is_inner_edge=is_inner_edge, is_ou2in_edge=is_ou2in_edge,
        inner_row=local[row[is_inner_edge]], inner_col=local[col[is_inner_edge]],
        ou2in_row=local[row[is_ou2in_edge]], ou2in_col=local[col[is_ou2in_edge]],
        n_inner=int(inner.sum()), n_outer=int(mask.sum()), inner=inner,
    )


def condense(values, row, col, mask, rhs=None, dirichlet_value=None):
    """Return ``(Kii_vals, Kii_row, Kii_col, n_inner, f_i, lay)``."""
    lay = layout(row, col, mask)
    values = np.asarray(values, dtype=np.float64)
    n = len(mask)
    rhs = np.zeros(n) if rhs is None else np.asarray(rhs, dtype=np.float64)
    uo = np.zeros(lay["n_outer"]) if dirichlet_value is None else np.asarray(dirichlet_value, np.float64)
    if uo.shape[0] == n:
        uo = uo[np.asarray(mask, bool)]
    kio = values[lay["is_ou2in_edge"]]
    minus = np.bincount(lay["ou2in_row"], weights=kio * uo[lay["ou2in_col"]], minlength=lay["n_inner"])
    fi = rhs[lay["inner"]] - minus
    return values[lay["is_inner_edge"]], lay["inner_row"], lay["inner_col"], lay["n_inner"], fi, lay


def recover(u_inner, mask, dirichlet_value=None):
    mask = np.asarray(mask, dtype=bool)
    out = np.zeros(mask.shape[0])
    out[~mask] = u_inner
    if dirichlet_value is not None:
        dv = np.asarray(dirichlet_value, np.float64)
        out[mask] = dv[mask] if dv.shape[0] == mask.shape[0] else dv
    return out

"""Test matrices shared by the CPU and GPU solver tests (SciPy, float64)."""
import numpy as np
import scipy.sparse


def nonsym_banded(n, seed, density=0.1, band=40):
    """Non-symmetric, diagonally dominant, ragged rows inside a band."""
    rng = np.random.default_rng(seed)
    offs = np.arange(-band, band + 1)
    rows = np.repeat(np.arange(n), len(offs))
    cols = rows + np.tile(offs, n)
    keep = (cols >= 0) & (cols < n) & (cols != rows) & (rng.random(len(rows)) < density)
    B = scipy.sparse.csr_matrix((rng.uniform(-1.0, 1.0, keep.sum()), (rows[keep], cols[keep])), shape=(n, n))
    d = np.asarray(abs(B).sum(1)).ravel() + 0.5
    return (B + scipy.sparse.diags(d)).tocsr()


def zero_diagonal_rows(n, seed, npairs):
    """Non-singular, non-symmetric matrix with stored zero diagonal entries:
    a diagonally dominant band matrix whose rows i, i+1 are swapped for
    ``npairs`` even i, with (i, i+1) and (i+1, i) set to explicit zeros first,
    so both swapped rows carry a stored 0 on the diagonal (a saddle-point-like
    row structure).  Returns ``(A csr, first rows of the pairs)``."""
    M = nonsym_banded(n, seed, density=0.05, band=20).tocoo()
    rng = np.random.default_rng(seed + 1)
    pr = np.sort(rng.choice(np.arange(0, n - 1, 2), npairs, replace=False))
    isp = np.zeros(n, bool)
    isp[pr] = True
    r, c, v = M.row, M.col, M.data
    lo = np.minimum(r, c)
    drop = isp[lo] & (np.abs(r - c) == 1) & (lo % 2 == 0)
    r = np.concatenate([r[~drop], pr + 1, pr])
    c = np.concatenate([c[~drop], pr, pr + 1])
    v = np.concatenate([v[~drop], np.zeros(2 * npairs)])
    perm = np.arange(n)
    perm[pr], perm[pr + 1] = pr + 1, pr
    A = scipy.sparse.csr_matrix((v, (perm[r], c)), shape=(n, n))
    A.sort_indices()
    return A, pr

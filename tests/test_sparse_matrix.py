"""SparseMatrix host logic (CPU): the symmetry flag of sums and the
tolerance of the symmetry check."""
import numpy as np
import torch

import tensormesh_amd as tm


def _pair():
    row = torch.tensor([0, 0, 1, 1, 2, 2], dtype=torch.long)
    col = torch.tensor([0, 1, 0, 1, 1, 2], dtype=torch.long)
    return row, col


def test_sum_symmetry_flag():
    row, col = _pair()
    a = tm.SparseMatrix(torch.ones(6, dtype=torch.float64), row, col, (3, 3))
    b = tm.SparseMatrix(torch.arange(6, dtype=torch.float64), row, col, (3, 3))
    for fa, fb, want in [(True, True, True), (True, False, False), (False, True, False),
                         (False, False, None), (None, True, None), (True, None, None), (None, None, None)]:
        a.symmetric, b.symmetric = fa, fb
        assert (a + b).symmetric is want, (fa, fb)
        assert (a - b).symmetric is want, (fa, fb)
    a.symmetric = True
    assert (2.0 * a).symmetric is True and (-a).symmetric is True


def test_is_symmetric_tolerates_last_bit_differences():
    """Duplicate COO entries summed in different orders may differ in the last
    bits: such a matrix still counts as symmetric; a real asymmetry does not."""
    rng = np.random.default_rng(0)
    n = 40
    S = rng.standard_normal((n, n))
    S = S + S.T
    r, c = np.nonzero(np.abs(S) > 1.0)
    for dt, eps in [(torch.float64, 2.2e-16), (torch.float32, 1.2e-7)]:
        v = torch.tensor(S[r, c], dtype=dt)
        upper = torch.tensor(r < c)
        v_pert = torch.where(upper, v * (1 + 2 * eps), v)
        m = tm.SparseMatrix(v_pert, torch.tensor(r), torch.tensor(c), (n, n))
        assert m.is_symmetric()
        v_bad = torch.where(upper, v * (1 + 1e-3), v)
        m = tm.SparseMatrix(v_bad, torch.tensor(r), torch.tensor(c), (n, n))
        assert not m.is_symmetric()

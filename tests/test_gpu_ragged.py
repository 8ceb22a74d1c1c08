"""Ragged CSR on the GPU: row lengths 0..k (empty rows, lengths not a multiple of
four, rows ending at a workgroup's last entry), for the scalar and 3x3-block SpMV
row loops and Jacobi-PCG, against SciPy in float64; plus the trivial solves."""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg
import torch

from tensormesh_amd.sparse.ops import CSR, _spmv_raw, pcg_raw

pytestmark = pytest.mark.gpu


def _ragged(n, max_len, seed, n_cols=None, empty_frac=0.1):
    rng = np.random.default_rng(seed)
    n_cols = n_cols or n
    lens = rng.integers(1, max_len + 1, n)
    lens[rng.random(n) < empty_frac] = 0
    rows, cols = [], []
    for i, ln in enumerate(lens):
        c = np.sort(rng.choice(n_cols, size=min(int(ln), n_cols), replace=False))
        rows.append(np.full(len(c), i))
        cols.append(c)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp).astype(np.int32), cols.astype(np.int32)


def _csr(rp, ci, n, cuda):
    return CSR(torch.tensor(rp, device=cuda), torch.tensor(ci, device=cuda), n, n)


@pytest.mark.parametrize("max_len", [1, 3, 7, 13, 30, 60, 130])
@pytest.mark.parametrize("vdt", [torch.float32, torch.float64])
def test_scalar_spmv_ragged(cuda, max_len, vdt):
    print(f"case ragged spmv max_len={max_len} {vdt}", flush=True)
    n = 20011
    rp, ci = _ragged(n, max_len, seed=max_len)
    rng = np.random.default_rng(1)
    v = rng.standard_normal(len(ci)).astype(np.float32 if vdt == torch.float32 else np.float64)
    x = rng.standard_normal(n)
    y = _spmv_raw(_csr(rp, ci, n, cuda), torch.tensor(v, device=cuda), torch.tensor(x, device=cuda)).cpu().numpy()
    ref = scipy.sparse.csr_matrix((v.astype(np.float64), ci, rp), shape=(n, n)) @ x
    scale = np.abs(scipy.sparse.csr_matrix((np.abs(v).astype(np.float64), ci, rp), shape=(n, n)) @ np.abs(x))
    assert np.all(np.abs(y - ref) <= 1e-13 * scale + 1e-300)
    assert np.all(y[np.diff(rp) == 0] == 0.0)


@pytest.mark.parametrize("max_len", [2, 9, 30])
def test_block_spmv_ragged(cuda, max_len):
    print(f"case ragged bsr3 max_len={max_len}", flush=True)
    n = 5003
    rp, ci = _ragged(n, max_len, seed=100 + max_len)
    rng = np.random.default_rng(2)
    v = rng.standard_normal((len(ci), 3, 3)).astype(np.float32)
    x = rng.standard_normal(3 * n)
    y = _spmv_raw(_csr(rp, ci, n, cuda), torch.tensor(v.reshape(-1), device=cuda), torch.tensor(x, device=cuda),
                  block_size=3).cpu().numpy()
    A = scipy.sparse.bsr_matrix((v.astype(np.float64), ci, rp), shape=(3 * n, 3 * n))
    ref = A @ x
    scale = abs(A) @ np.abs(x)
    assert np.all(np.abs(y - ref) <= 1e-13 * scale + 1e-300)


def _spd_from(rp, ci, n, seed):
    """Symmetric, strictly diagonally dominant matrix on the symmetrised pattern (plus the diagonal)."""
    B = scipy.sparse.csr_matrix((np.random.default_rng(seed).uniform(-1, 1, len(ci)), ci, rp), shape=(n, n))
    S = (B + B.T).tocsr()
    S.setdiag(0.0)
    S.eliminate_zeros()
    d = np.asarray(abs(S).sum(1)).ravel() + 1.0
    A = (S + scipy.sparse.diags(d)).tocsr().astype(np.float32)
    A.sort_indices()
    return A


@pytest.mark.parametrize("max_len", [3, 12, 40])
def test_pcg_ragged_vs_direct(cuda, max_len):
    print(f"case ragged pcg max_len={max_len}", flush=True)
    n = 8009
    rp, ci = _ragged(n, max_len, seed=200 + max_len, empty_frac=0.0)
    A = _spd_from(rp, ci, n, seed=max_len)
    b = np.random.default_rng(3).standard_normal(n)
    csr = _csr(A.indptr.astype(np.int32), A.indices.astype(np.int32), n, cuda)
    x, it, rel, ok = pcg_raw(csr, torch.tensor(A.data, device=cuda), torch.tensor(b, device=cuda), rtol=1e-12)
    A64 = A.astype(np.float64)
    xd = scipy.sparse.linalg.spsolve(A64.tocsc(), b)
    assert ok and rel <= 1e-12
    assert np.linalg.norm(x.cpu().numpy() - xd) <= 1e-10 * np.linalg.norm(xd)


def test_pcg_trivial_systems(cuda):
    print("case trivial pcg", flush=True)
    # zero right-hand side: x = 0 without an iteration
    n = 1000
    rp, ci = _ragged(n, 5, seed=7, empty_frac=0.0)
    A = _spd_from(rp, ci, n, seed=7)
    csr = _csr(A.indptr.astype(np.int32), A.indices.astype(np.int32), n, cuda)
    vals = torch.tensor(A.data, device=cuda)
    x, it, rel, ok = pcg_raw(csr, vals, torch.zeros(n, dtype=torch.float64, device=cuda), rtol=1e-8)
    assert ok and it == 0 and torch.count_nonzero(x).item() == 0
    # diagonal system: one iteration
    d = np.linspace(1.0, 5.0, 777).astype(np.float32)
    D = scipy.sparse.diags(d).tocsr()
    csr = _csr(D.indptr.astype(np.int32), D.indices.astype(np.int32), 777, cuda)
    b = np.random.default_rng(8).standard_normal(777)
    x, it, rel, ok = pcg_raw(csr, torch.tensor(D.data.astype(np.float32), device=cuda), torch.tensor(b, device=cuda),
                             rtol=1e-12)
    assert ok and it == 1 and np.abs(x.cpu().numpy() - b / d.astype(np.float64)).max() <= 1e-15 * np.abs(b / d).max()
    # a single unknown
    one = _csr(np.array([0, 1], np.int32), np.array([0], np.int32), 1, cuda)
    x, it, rel, ok = pcg_raw(one, torch.tensor([4.0], dtype=torch.float32, device=cuda),
                             torch.tensor([2.0], dtype=torch.float64, device=cuda), rtol=1e-12)
    assert ok and x.item() == 0.5

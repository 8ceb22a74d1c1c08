"""Solver selection and BiCGSTAB edge cases on the GPU: the symmetry flag of a
sum of a symmetric and a non-symmetric matrix (BiCGSTAB, adjoint against
dense autograd), symmetric COO input with duplicate entries (PCG), stored zero
diagonals (the 1e-12 guard of bicgstab_py, torch_solve.py:177-182, against
SuperLU), CG breakdown on a symmetric indefinite matrix (BiCGSTAB takes over)
and the true residual reported when the recursive one meets a tolerance the
true one cannot."""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg
import torch

import tensormesh_amd as tm
from oracle import solve as osv
from tensormesh_amd import _lib
from tensormesh_amd.mesh import structured_rectangle_arrays
from tensormesh_amd.sparse.ops import CSR, _spmv_raw, bicgstab_raw
from tests.matrices import nonsym_banded, zero_diagonal_rows

pytestmark = pytest.mark.gpu


def _csr(A, cuda):
    A = A.tocsr()
    A.sort_indices()
    return CSR(torch.tensor(A.indptr.astype(np.int32), device=cuda),
               torch.tensor(A.indices.astype(np.int32), device=cuda), A.shape[0], A.shape[1]), A


def test_symmetric_plus_nonsymmetric_takes_bicgstab(cuda):
    print("case sum symmetric + convection", flush=True)
    pts, conn, bnd = structured_rectangle_arrays(8, jitter=0.2, seed=2)
    mesh = tm.Mesh(torch.from_numpy(pts), {"triangle": torch.from_numpy(conn)},
                   {"is_boundary": torch.from_numpy(bnd)})
    beta = torch.tensor([3.0, -2.0], dtype=torch.float64)

    class Conv(tm.ElementAssembler):
        def forward(self, gradu, v):
            return (beta.to(gradu.dtype) @ gradu) * v

    Kd = tm.LaplaceElementAssembler.from_mesh(mesh).to(cuda)()
    Kc = Conv.from_mesh(mesh).to(cuda)(mesh.points.to(cuda))
    assert Kd.has_same_layout(Kc) and Kd.symmetric is True and Kc.symmetric is False
    for K in (Kd + Kc, Kc + Kd, Kd - Kc):
        assert K.symmetric is False
    n = Kd.shape[0]
    shift = tm.MassElementAssembler.from_mesh(mesh).to(cuda)()   # makes the sum non-singular
    row, col = Kd.row, Kd.col
    vd = Kd.values.detach().double().clone().requires_grad_(True)
    vc = Kc.values.detach().double().clone().requires_grad_(True)
    A1 = tm.SparseMatrix(vd + shift.values.double(), row, col, (n, n))
    A1.symmetric = True
    A2 = tm.SparseMatrix(vc, row, col, (n, n))
    A2.symmetric = False
    K = A1 + A2
    b = torch.linspace(-1.0, 2.0, n, dtype=torch.float64, device=cuda).requires_grad_(True)
    x = K.solve(b, tol=1e-13, atol=0.0)
    assert K.last_solve_info["method"] == "bicgstab" and K.last_solve_info["converged"]
    w = torch.cos(torch.arange(n, dtype=torch.float64, device=cuda))
    (x * w).sum().backward()
    vd_c = vd.detach().cpu().requires_grad_(True)
    vc_c = vc.detach().cpu().requires_grad_(True)
    b_c = b.detach().cpu().requires_grad_(True)
    dense = torch.zeros(n, n, dtype=torch.float64).index_put(
        (row.cpu(), col.cpu()), vd_c + shift.values.double().cpu() + vc_c, accumulate=True)
    xd = torch.linalg.solve(dense, b_c)
    (xd * w.cpu()).sum().backward()
    assert torch.allclose(x.detach().cpu(), xd.detach(), rtol=1e-9, atol=1e-12)
    assert torch.allclose(vd.grad.cpu(), vd_c.grad, rtol=1e-7, atol=1e-10)
    assert torch.allclose(vc.grad.cpu(), vc_c.grad, rtol=1e-7, atol=1e-10)
    assert torch.allclose(b.grad.cpu(), b_c.grad, rtol=1e-7, atol=1e-10)


def test_symmetric_coo_with_duplicates_takes_pcg(cuda):
    """Element-COO input (every element's 3x3 block, duplicates summed by the
    canonicalisation): classified symmetric on every call, solved by PCG."""
    print("case duplicated symmetric COO", flush=True)
    pts, conn, _ = structured_rectangle_arrays(12, jitter=0.2, seed=4)
    n = len(pts)
    rng = np.random.default_rng(5)
    r, c, v = [], [], []
    for e in conn:
        w = rng.uniform(0.5, 2.0)
        blk = w * (2.0 * np.eye(3) + 0.5)          # SPD, well conditioned
        for a in range(3):
            for bb in range(3):
                r.append(e[a]); c.append(e[bb]); v.append(blk[a, bb])
    r, c = np.array(r), np.array(c)
    v = np.array(v)
    for trial in range(3):
        K = tm.SparseMatrix(torch.tensor(v, dtype=torch.float32, device=cuda), torch.tensor(r, device=cuda),
                            torch.tensor(c, device=cuda), (n, n))
        assert K.is_symmetric(), f"trial {trial}"
    # solve in FP64 (FP32 duplicate sums are not reproducible to the last bit
    # from call to call, which a comparison at 1e-8 would see)
    K = tm.SparseMatrix(torch.tensor(v, dtype=torch.float64, device=cuda), torch.tensor(r, device=cuda),
                        torch.tensor(c, device=cuda), (n, n))
    x = K.solve(torch.ones(n, dtype=torch.float64, device=cuda), tol=1e-12, atol=0.0)
    assert K.last_solve_info["method"] == "jacobi-pcg" and K.last_solve_info["converged"]
    A = scipy.sparse.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    xd = scipy.sparse.linalg.spsolve(A.tocsc(), np.ones(n))
    assert np.linalg.norm(x.cpu().numpy() - xd) <= 1e-10 * np.linalg.norm(xd)


@pytest.mark.parametrize("n,npairs", [(4000, 40), (20000, 200)])
def test_bicgstab_stored_zero_diagonal(cuda, n, npairs):
    print(f"case bicgstab zero diagonal n={n}", flush=True)
    A, pr = zero_diagonal_rows(n, 9, npairs)
    csr, A = _csr(A, cuda)
    b = np.random.default_rng(3).standard_normal(n)
    x, it, rel, ok = bicgstab_raw(csr, torch.tensor(A.data, device=cuda), torch.tensor(b, device=cuda),
                                  rtol=1e-12, maxiter=5000)
    xd = scipy.sparse.linalg.spsolve(A.tocsc(), b)
    assert ok and rel <= 1e-12
    assert np.linalg.norm(x.cpu().numpy() - xd) <= 1e-9 * np.linalg.norm(xd)
    _, it_ref, _ = osv.bicgstab(A, b, rtol=1e-12, maxiter=5000)
    assert abs(it - it_ref) <= max(3, 0.1 * it_ref)
    # through K.solve (non-symmetric: BiCGSTAB is picked)
    Ac = A.tocoo()
    K = tm.SparseMatrix(torch.tensor(Ac.data, device=cuda), torch.tensor(Ac.row, dtype=torch.long, device=cuda),
                        torch.tensor(Ac.col, dtype=torch.long, device=cuda), (n, n))
    xs = K.solve(torch.tensor(b, device=cuda), tol=1e-12, atol=0.0)
    assert K.last_solve_info["method"] == "bicgstab"
    assert np.linalg.norm(xs.cpu().numpy() - xd) <= 1e-9 * np.linalg.norm(xd)


def test_pcg_breakdown_hands_over_to_bicgstab(cuda):
    """Symmetric indefinite tridiagonal matrix (diagonal +4/-4, off-diagonal 1):
    with b = e_1 (a row with diagonal -4) the first p.Ap is negative, CG breaks
    down and K.solve finishes with BiCGSTAB."""
    print("case pcg breakdown", flush=True)
    n = 301
    d = np.where(np.arange(n) % 2 == 0, 4.0, -4.0)
    A = scipy.sparse.diags([np.ones(n - 1), d, np.ones(n - 1)], [-1, 0, 1]).tocoo()
    K = tm.SparseMatrix(torch.tensor(A.data, device=cuda), torch.tensor(A.row, dtype=torch.long, device=cuda),
                        torch.tensor(A.col, dtype=torch.long, device=cuda), (n, n))
    assert K.is_symmetric()
    b = np.zeros(n)
    b[1] = 1.0
    with pytest.warns(UserWarning, match="breakdown"):
        x = K.solve(torch.tensor(b, device=cuda), tol=1e-12, atol=0.0)
    info = K.last_solve_info
    assert info["method"] == "bicgstab" and info["converged"]
    assert info["after_pcg_breakdown"]["breakdown"] is True
    xd = scipy.sparse.linalg.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x.cpu().numpy() - xd) <= 1e-10 * np.linalg.norm(xd)


def test_bicgstab_reports_true_residual_on_a_miss(cuda):
    """rtol 1e-17 lies below what FP64 can attain: the recursive residual meets
    it, the true residual does not; after three restarts the solve ends with
    TM_ERR_NOTCONV and reports the true residual of the returned x."""
    print("case bicgstab true-residual miss", flush=True)
    A = nonsym_banded(3000, 21)
    csr, A = _csr(A, cuda)
    vals = torch.tensor(A.data, device=cuda)
    b = torch.tensor(np.random.default_rng(22).standard_normal(3000), device=cuda)
    with pytest.warns(UserWarning, match="true residual misses"):
        x, it, rel, ok = bicgstab_raw(csr, vals, b, rtol=1e-17, maxiter=3000)
    assert not ok
    true_rel = float(torch.linalg.vector_norm(b - _spmv_raw(csr, vals, x)) / torch.linalg.vector_norm(b))
    assert true_rel > 1e-17
    assert abs(rel - true_rel) <= 1e-6 * true_rel

"""GPU BiCGSTAB (``tm_bicgstab_jacobi``, the reference's ``lse_solver``
``bicgstab_py``, torch_solve.py:130-248) on non-symmetric systems: a
convection-diffusion matrix from the general-form kernel through the Condenser
and ``K.solve`` (auto-selected), random ragged systems against SuperLU and the
float64 oracle, trivial systems, and the adjoint against dense autograd."""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg
import torch

import tensormesh_amd as tm
from oracle import solve as osv
from tensormesh_amd.mesh import structured_rectangle_arrays
from tensormesh_amd.sparse.ops import CSR, bicgstab_raw

pytestmark = pytest.mark.gpu


def _csr(A, cuda):
    A = A.tocsr()
    A.sort_indices()
    return (CSR(torch.tensor(A.indptr.astype(np.int32), device=cuda), torch.tensor(A.indices.astype(np.int32),
                device=cuda), A.shape[0], A.shape[1]), A)


def _nonsym(n, seed, density=0.1, band=40):
    """Non-symmetric, diagonally dominant, ragged rows inside a band (SuperLU stays cheap)."""
    rng = np.random.default_rng(seed)
    offs = np.arange(-band, band + 1)
    rows = np.repeat(np.arange(n), len(offs))
    cols = rows + np.tile(offs, n)
    keep = (cols >= 0) & (cols < n) & (cols != rows) & (rng.random(len(rows)) < density)
    B = scipy.sparse.csr_matrix((rng.uniform(-1.0, 1.0, keep.sum()), (rows[keep], cols[keep])), shape=(n, n))
    d = np.asarray(abs(B).sum(1)).ravel() + 0.5
    return (B + scipy.sparse.diags(d)).tocsr()


@pytest.mark.parametrize("vdt", [np.float32, np.float64])
def test_bicgstab_vs_superlu_and_oracle(cuda, vdt):
    print(f"case bicgstab random {vdt.__name__}", flush=True)
    A = _nonsym(20000, 3).astype(vdt)
    csr, A = _csr(A, cuda)
    b = np.random.default_rng(4).standard_normal(A.shape[0])
    x, it, rel, ok = bicgstab_raw(csr, torch.tensor(A.data, device=cuda), torch.tensor(b, device=cuda), rtol=1e-12)
    A64 = A.astype(np.float64)
    xd = scipy.sparse.linalg.spsolve(A64.tocsc(), b)
    assert ok and rel <= 1e-12
    assert np.linalg.norm(x.cpu().numpy() - xd) <= 1e-9 * np.linalg.norm(xd)
    _, it_ref, _ = osv.bicgstab(A64, b, rtol=1e-12)
    assert abs(it - it_ref) <= max(3, 0.1 * it_ref)


def test_convection_diffusion_solve(cuda):
    """-div(grad u) + beta . grad u = 1 on the unit square, zero Dirichlet: the general
    kernel assembles a non-symmetric K; K.solve picks BiCGSTAB."""
    print("case bicgstab convection-diffusion", flush=True)
    pts, conn, bnd = structured_rectangle_arrays(96, jitter=0.2, seed=1)
    mesh = tm.Mesh(torch.from_numpy(pts), {"triangle": torch.from_numpy(conn)},
                   {"is_boundary": torch.from_numpy(bnd)})
    beta = torch.tensor([20.0, -10.0], dtype=torch.float64)

    class ConvDiff(tm.ElementAssembler):
        def forward(self, gradu, gradv, v):
            return gradu @ gradv + (beta.to(gradu.dtype) @ gradu) * v

    asm = ConvDiff.from_mesh(mesh).to(cuda)
    K = asm(mesh.points.to(cuda))
    assert K.symmetric is False
    f = tm.const_node_assembler(1.0).from_mesh(mesh).to(cuda)()
    Ki, fi = tm.Condenser(mesh.boundary_mask.to(cuda))(K, f)
    x = Ki.solve(fi, tol=1e-10, atol=0.0)
    assert Ki.last_solve_info["method"] == "bicgstab" and Ki.last_solve_info["converged"]
    A = Ki.to_scipy_coo().tocsr()
    xd = scipy.sparse.linalg.spsolve(A.tocsc(), fi.double().cpu().numpy())
    assert np.linalg.norm(x.cpu().numpy() - xd) <= 1e-8 * np.linalg.norm(xd)
    # the reported residual is the true one, b - A x recomputed after convergence
    bh = fi.double().cpu().numpy()
    true_rel = np.linalg.norm(bh - A @ x.cpu().numpy()) / np.linalg.norm(bh)
    assert true_rel <= 1e-10 * 1.01
    assert abs(Ki.last_solve_info["relres"] - true_rel) <= 1e-3 * true_rel + 1e-16


def test_bicgstab_trivial_systems(cuda):
    print("case bicgstab trivial", flush=True)
    A = _nonsym(3000, 5)
    csr, A = _csr(A, cuda)
    vals = torch.tensor(A.data.astype(np.float32), device=cuda)
    x, it, rel, ok = bicgstab_raw(csr, vals, torch.zeros(A.shape[0], dtype=torch.float64, device=cuda))
    assert ok and it == 0 and torch.count_nonzero(x).item() == 0
    d = np.linspace(1.0, 4.0, 500)
    csr, D = _csr(scipy.sparse.diags(d), cuda)
    b = np.random.default_rng(6).standard_normal(500)
    x, it, rel, ok = bicgstab_raw(csr, torch.tensor(D.data, device=cuda), torch.tensor(b, device=cuda), rtol=1e-14)
    assert ok and it == 1 and np.abs(x.cpu().numpy() - b / d).max() <= 1e-14 * np.abs(b / d).max()


def test_bicgstab_gradients_vs_dense(cuda):
    """Adjoint (tests/sparse/test_spsolve.py:46-80): values.grad and b.grad match dense
    autograd masked to the pattern, for a non-symmetric matrix."""
    print("case bicgstab gradients", flush=True)
    n = 24
    A = _nonsym(n, 7, density=0.3, band=6).tocoo()
    row, col = torch.tensor(A.row, dtype=torch.long), torch.tensor(A.col, dtype=torch.long)
    vals = torch.tensor(A.data, dtype=torch.float64, device=cuda, requires_grad=True)
    b = torch.tensor(np.random.default_rng(8).standard_normal(n), device=cuda, requires_grad=True)
    K = tm.SparseMatrix(vals, row.to(cuda), col.to(cuda), (n, n))
    x = K.solve(b, method="bicgstab", tol=1e-14, atol=0.0)
    w = torch.linspace(-1.0, 1.0, n, dtype=torch.float64, device=cuda)
    (x * w).sum().backward()
    vd = vals.detach().clone().cpu().requires_grad_(True)
    bd = b.detach().clone().cpu().requires_grad_(True)
    Ad = torch.zeros(n, n, dtype=torch.float64).index_put((row, col), vd, accumulate=True)
    xd = torch.linalg.solve(Ad, bd)
    (xd * w.cpu()).sum().backward()
    assert torch.allclose(x.detach().cpu(), xd.detach(), rtol=1e-10, atol=1e-12)
    assert torch.allclose(vals.grad.cpu(), vd.grad, rtol=1e-8, atol=1e-10)
    assert torch.allclose(b.grad.cpu(), bd.grad, rtol=1e-8, atol=1e-10)

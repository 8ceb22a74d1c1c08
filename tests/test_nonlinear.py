"""``nonlinear_solve`` (``sparse/nonlinear_solve.py:21-175``): Newton with the
adjoint backward, on a small sparse nonlinear system ``F(u) = A u + u^3 - b``
(A non-symmetric on the GPU case, so the linear solves use BiCGSTAB there)."""
import numpy as np
import pytest
import scipy.sparse
import torch

import tensormesh_amd as tm
from tensormesh_amd.sparse import nonlinear_solve


def _system(n, device, symmetric):
    rng = np.random.default_rng(0)
    A = scipy.sparse.diags([-1.0, 4.0, -1.0 if symmetric else -1.5], [-1, 0, 1], shape=(n, n)).tocoo()
    row = torch.tensor(A.row, dtype=torch.long, device=device)
    col = torch.tensor(A.col, dtype=torch.long, device=device)
    vals = torch.tensor(A.data, dtype=torch.float64, device=device)

    def f(u, b):
        Au = torch.zeros_like(u).index_add(0, row, vals * u[col])
        return Au + u ** 3 - b

    def j(u, b):
        diag = torch.arange(n, device=device)
        r = torch.cat([row, diag])
        c = torch.cat([col, diag])
        v = torch.cat([vals, 3.0 * u.detach() ** 2])
        return tm.SparseMatrix(v, r, c, (n, n))

    b = torch.tensor(rng.uniform(0.5, 1.5, n), dtype=torch.float64, device=device, requires_grad=True)
    return f, j, b


def _check(device, symmetric):
    n = 60
    f, j, b = _system(n, device, symmetric)
    u0 = torch.zeros(n, dtype=torch.float64, device=device)
    u = nonlinear_solve(f, j, u0, (b,), tol=1e-12)
    assert torch.linalg.vector_norm(f(u.detach(), b.detach())).item() < 1e-10
    w = torch.linspace(-1.0, 1.0, n, dtype=torch.float64, device=device)
    (u * w).sum().backward()
    # central differences of L(b) = w . u(b) in two directions
    for k in (3, 41):
        h = 1e-6
        bp, bm = b.detach().clone(), b.detach().clone()
        bp[k] += h
        bm[k] -= h
        up = nonlinear_solve(f, j, u0, (bp,), tol=1e-13)
        um = nonlinear_solve(f, j, u0, (bm,), tol=1e-13)
        fd = ((up - um) * w).sum().item() / (2 * h)
        assert abs(fd - b.grad[k].item()) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("symmetric", [True, False])
def test_nonlinear_solve_cpu(symmetric):
    _check("cpu", symmetric)


@pytest.mark.gpu
@pytest.mark.parametrize("symmetric", [True, False])
def test_nonlinear_solve_gpu(cuda, symmetric):
    print(f"case nonlinear solve symmetric={symmetric}", flush=True)
    _check(cuda, symmetric)

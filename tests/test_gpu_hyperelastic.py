"""GPU parity of the neo-Hookean kernels (energy, residual, tangent, adjoint, Newton)."""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg
import torch

import tensormesh_amd as tm
from oracle import hyperelastic as oh
from oracle import pattern as op
from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays

pytestmark = pytest.mark.gpu


def _case(n=4, seed=0, amp=0.01):
    pts, conn, bnd = kuhn_cube_arrays(n)
    rng = np.random.default_rng(seed)
    pts = pts + (~bnd)[:, None] * rng.uniform(-0.1, 0.1, pts.shape) / n
    u = amp * rng.standard_normal(pts.shape)
    E = 1.0 + rng.uniform(-0.1, 0.1, len(conn))
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    return m, pts, conn, bnd, u, E


@pytest.mark.parametrize("edt", [torch.float32, torch.float64, None])
def test_energy_and_residual(cuda, edt):
    print(f"case nh residual E={edt}", flush=True)
    m, pts, conn, bnd, u, E = _case()
    if edt is None:
        E = np.full(len(conn), 1.3)
        model = tm.NeoHookeanModel.from_mesh(m, E=1.3, nu=0.3).to(cuda)
        Et = None
    else:
        Et = torch.tensor(E, dtype=edt, device=cuda)
        E = Et.double().cpu().numpy()
        model = tm.NeoHookeanModel.from_mesh(m, E=Et, nu=0.3).to(cuda)
    ut = torch.tensor(u, device=cuda)
    W = model.energy(ut).item()
    Wr = oh.energy(pts, conn, u, E)
    assert abs(W - Wr) <= 1e-12 * abs(Wr)
    R = model.residual(ut).cpu().numpy()
    Rr = oh.residual(pts, conn, u, E, len(pts))
    assert np.abs(R - Rr).max() <= 1e-12 * np.abs(Rr).max()


@pytest.mark.parametrize("vdt", [torch.float32, torch.float64])
def test_tangent_blocks_and_bsr_products(cuda, vdt):
    print(f"case nh tangent {vdt}", flush=True)
    m, pts, conn, bnd, u, E = _case(seed=1)
    model = tm.NeoHookeanModel.from_mesh(m, E=torch.tensor(E, device=cuda), nu=0.3).to(cuda)
    K = model.tangent(torch.tensor(u, device=cuda), value_dtype=vdt)
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    T = oh.tangent_blocks(pts, conn, u, E, pe, len(eu))
    got = K.values.double().cpu().numpy().reshape(-1, 3, 3)
    # row-scaled: relative to the largest diagonal entry of the block row
    diag = np.abs(T[eu == ev][:, [0, 1, 2], [0, 1, 2]]).max(1)
    err = (np.abs(got - T).max(axis=2) / diag[eu][:, None]).max()
    assert err <= (1e-6 if vdt == torch.float32 else 1e-12)
    # expanded COO equals from_block_coo of the same blocks
    ref = tm.SparseMatrix.from_block_coo(K.values.view(-1, 3, 3), model.edges[0], model.edges[1],
                                         (len(pts), len(pts)))
    assert torch.equal(ref.row, K.row) and torch.equal(ref.col, K.col)
    # the reference's dense-placement check of block COO (tests/sparse/test_block_coo.py:27-44)
    # on the assembled tangent: block k lands at [3 eu_k : 3 eu_k + 3, 3 ev_k : 3 ev_k + 3]
    N = len(pts)
    dense = np.zeros((3 * N, 3 * N))
    for i, j, blk in zip(eu, ev, got):
        dense[3 * i:3 * i + 3, 3 * j:3 * j + 3] += blk
    assert np.array_equal(K.to_dense().double().cpu().numpy(), dense)
    x = torch.randn(3 * len(pts), dtype=torch.float64, device=cuda)
    y = (K @ x).cpu().numpy()
    yr = K.to_scipy_coo().tocsr() @ x.cpu().numpy()
    assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()
    # matrix-free tangent action agrees with the assembled tangent (float64 values)
    if vdt == torch.float64:
        ya = model.apply_tangent(torch.tensor(u, device=cuda), x.view(-1, 3)).reshape(-1).cpu().numpy()
        assert np.abs(ya - yr).max() <= 1e-12 * np.abs(yr).max()


def test_residual_adjoint(cuda):
    """dL/du and dL/dE of L = g . R(u, E) against the oracle and central differences."""
    print("case nh adjoint", flush=True)
    m, pts, conn, bnd, u, E = _case(seed=2)
    model = tm.NeoHookeanModel.from_mesh(m, nu=0.3).to(cuda)
    ut = torch.tensor(u, device=cuda, requires_grad=True)
    Et = torch.tensor(E, device=cuda, requires_grad=True)
    g = torch.tensor(np.random.default_rng(5).standard_normal(u.shape), device=cuda)
    (model.residual(ut, Et) * g).sum().backward()
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    T = oh.tangent_blocks(pts, conn, u, E, pe, len(eu))
    A = scipy.sparse.bsr_matrix((T, ev, op.csr_from_edges(eu, len(pts))), shape=(3 * len(pts),) * 2)
    du_ref = A.T @ g.cpu().numpy().reshape(-1)
    assert np.abs(ut.grad.cpu().numpy().reshape(-1) - du_ref).max() <= 1e-11 * np.abs(du_ref).max()
    dE_ref = oh.dresidual_dE(pts, conn, u, E, g.cpu().numpy())
    assert np.abs(Et.grad.cpu().numpy() - dE_ref).max() <= 1e-11 * np.abs(dE_ref).max()
    h = 1e-6
    with torch.no_grad():
        for (i, c) in [(5, 0), (40, 2)]:
            up, um = ut.detach().clone(), ut.detach().clone()
            up[i, c] += h
            um[i, c] -= h
            fd = ((model.residual(up, Et) - model.residual(um, Et)) * g).sum().item() / (2 * h)
            assert abs(fd - ut.grad[i, c].item()) <= 1e-6 * max(1.0, abs(fd))
        for e in [0, 100]:
            Ep, Em = Et.detach().clone(), Et.detach().clone()
            Ep[e] += h
            Em[e] -= h
            fd = ((model.residual(ut, Ep) - model.residual(ut, Em)) * g).sum().item() / (2 * h)
            assert abs(fd - Et.grad[e].item()) <= 1e-6 * max(1.0, abs(fd))


def test_newton_cube_compression(cuda):
    """Clamped bottom, top displaced by 0.1 in z (the 64^3 configuration at n=8)."""
    print("case nh newton", flush=True)
    n = 8
    pts, conn, bnd = kuhn_cube_arrays(n)
    E = 1.0 + np.random.default_rng(1).uniform(-0.1, 0.1, len(conn))
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    model = tm.NeoHookeanModel.from_mesh(m, E=torch.tensor(E, device=cuda), nu=0.3).to(cuda)
    z = pts[:, 2]
    fixed = torch.tensor((z == 0) | (z == 1), device=cuda)
    u0 = torch.zeros((len(pts), 3), dtype=torch.float64, device=cuda)
    u0[:, 2] = 0.1 * torch.tensor(z, device=cuda)
    u, info = model.newton(u0, fixed, rtol=1e-8)
    assert info["converged"] and info["iterations"] <= 12, info
    # float64 oracle Newton with a direct inner solve from the same start
    free = ~fixed.cpu().numpy()
    ur = u0.cpu().numpy().copy()
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    fdof = np.repeat(free, 3)
    for _ in range(20):
        R = oh.residual(pts, conn, ur, E, len(pts)).reshape(-1)
        if np.linalg.norm(R[fdof]) < 1e-13:
            break
        T = oh.tangent_blocks(pts, conn, ur, E, pe, len(eu))
        A = scipy.sparse.bsr_matrix((T, ev, op.csr_from_edges(eu, len(pts))), shape=(3 * len(pts),) * 2).tocsr()
        du = scipy.sparse.linalg.spsolve(A[fdof][:, fdof].tocsc(), -R[fdof])
        ur.reshape(-1)[fdof] += du
    assert np.abs(u.cpu().numpy() - ur).max() <= 1e-7 * np.abs(ur).max()

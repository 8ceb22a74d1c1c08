"""Pins the CPU oracle against closed forms, known answers, brute force and the
reference's own test fixtures (CPU only)."""
import math

import numpy as np
import pytest
import scipy.sparse

from oracle import assemble as oa
from oracle import condense as oc
from oracle import hyperelastic as oh
from oracle import pattern as op
from oracle import solve as osv
from tensormesh_amd import element as tel
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays


# ---- quadrature / reference element (tests/quadrature/test_tri.py:17-23, test_tetra.py:17-21)
@pytest.mark.parametrize("et,measure", [("triangle", 0.5), ("tetra", 1.0 / 6.0)])
@pytest.mark.parametrize("order", [1, 2])
def test_quadrature_weights_sum(et, measure, order):
    w, phi, _ = oa.rule(et, order)
    assert abs(w.sum() - measure) < 1e-15
    assert np.allclose(phi.sum(1), 1.0)                       # partition of unity


@pytest.mark.parametrize("et,fact", [("triangle", 24.0), ("tetra", 120.0)])
def test_order2_rule_integrates_quadratics(et, fact):
    # int phi_a phi_b over the reference simplex = (1 + delta_ab) / fact (closed form)
    w, phi, _ = oa.rule(et, 2)
    M = np.einsum("q,qa,qb->ab", w, phi, phi)
    nb = phi.shape[1]
    assert np.allclose(M, (np.ones((nb, nb)) + np.eye(nb)) / fact, atol=1e-15)


@pytest.mark.parametrize("et", ["triangle", "tetra"])
@pytest.mark.parametrize("order", [1, 2])
def test_package_tables_match_oracle(et, order):
    w, q = tel.quadrature(et, order)
    w2, phi2, _ = oa.rule(et, order)
    assert np.allclose(w, w2, atol=1e-15) and np.allclose(tel.shape_val(et, q), phi2, atol=1e-15)


# ---- single-element known answers
def test_unit_triangle_laplace_and_mass():
    # tests/assemble/test_mixed_element_type.py:126-138 ('2d_pure_tri')
    pts = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    conn = np.array([[0, 1, 2]])
    K = oa.element_matrices(pts, conn, "triangle", "laplace")[0]
    assert np.allclose(K, [[1, -0.5, -0.5], [-0.5, 0.5, 0], [-0.5, 0, 0.5]], atol=1e-15)
    M = oa.element_matrices(pts, conn, "triangle", "mass")[0]
    assert np.allclose(M, np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]]) / 24.0, atol=1e-15)


def test_kuhn_tet_laplace():
    pts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]], dtype=float)
    K = oa.element_matrices(pts, np.array([[0, 1, 2, 3]]), "tetra", "laplace")[0]
    ref = np.array([[1, -1, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -1, 1]]) / 6.0
    assert np.allclose(K, ref, atol=1e-15)


def test_kappa_laplace_constant_kappa_is_scaled_laplace():
    pts, conn, _ = kuhn_cube_arrays(2)
    K1 = oa.element_matrices(pts, conn, "tetra", "laplace")
    K3 = oa.element_matrices(pts, conn, "tetra", "kappa_laplace", kappa=np.full(len(pts), 3.0))
    assert np.allclose(K3, 3.0 * K1, rtol=1e-14, atol=1e-16)


# ---- invariants (tests/assemble/test_builtin.py:24-79, test_batch_size.py:191-212)
@pytest.fixture(scope="module")
def jittered_tri():
    pts, conn, bnd = structured_rectangle_arrays(7, jitter=0.25, seed=11)
    return pts, conn, bnd


def test_matrix_invariants(jittered_tri):
    pts, conn, _ = jittered_tri
    n = len(pts)
    u, v, pe = op.pattern_and_slots(conn, n)
    kappa = np.exp(np.random.default_rng(2).standard_normal(n))
    K = scipy.sparse.csr_matrix((oa.reduce(pe, oa.element_matrices(pts, conn, "triangle", "kappa_laplace",
                                                                   kappa=kappa), len(u)), (u, v)), shape=(n, n))
    assert np.abs(K @ np.ones(n)).max() < 1e-12                      # K 1 = 0
    assert abs(K - K.T).max() < 1e-14                                  # symmetric
    assert np.linalg.eigvalsh(K.toarray()).min() > -1e-10              # PSD
    M = scipy.sparse.csr_matrix((oa.reduce(pe, oa.element_matrices(pts, conn, "triangle", "mass"), len(u)),
                                 (u, v)), shape=(n, n))
    assert abs(np.ones(n) @ (M @ np.ones(n)) - 1.0) < 1e-12            # 1^T M 1 = |Omega|
    b1 = oa.load_vector(pts, conn, "triangle", n, c=1.0)
    assert abs(b1.sum() - 1.0) < 1e-12
    f = np.sin(pts[:, 0] * 3) + pts[:, 1]
    assert np.allclose(oa.load_vector(pts, conn, "triangle", n, f=f), M @ f, atol=1e-14)   # b = M f


def test_tet_mass_and_volume():
    pts, conn, _ = kuhn_cube_arrays(3)
    n = len(pts)
    u, v, pe = op.pattern_and_slots(conn, n)
    M = scipy.sparse.csr_matrix((oa.reduce(pe, oa.element_matrices(pts, conn, "tetra", "mass"), len(u)), (u, v)),
                                shape=(n, n))
    assert abs(M.sum() - 1.0) < 1e-12
    assert abs(oa.load_vector(pts, conn, "tetra", n, c=2.0).sum() - 2.0) < 1e-12


# ---- pattern (tests/mesh/test_node_adjacency.py:7-21)
@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_pattern_brute_force(mesh):
    pts, conn, _ = structured_rectangle_arrays(4, 0.1, 1) if mesh == "tri" else kuhn_cube_arrays(2)
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    pairs = sorted(op.brute_force_pairs(conn))
    assert list(zip(u.tolist(), v.tolist())) == pairs          # canonical (row, col) order
    uu, vv = op.element_pairs(conn)
    assert np.array_equal(u[pe], uu) and np.array_equal(v[pe], vv)


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_kuhn_nnz_closed_form(n):
    pts, conn, bnd = kuhn_cube_arrays(n)
    u, v, _ = op.pattern_and_slots(conn, len(pts))
    assert len(u) == op.kuhn_nnz(n)
    rp = op.csr_from_edges(u, len(pts))
    assert (np.diff(rp)[~bnd] == 15).all()                     # 14 Kuhn neighbours + diagonal


def test_kuhn_nnz_baseline_sizes():
    assert op.kuhn_nnz(128) == 31_802_497
    assert op.kuhn_nnz(96) == 13_465_441
    assert op.kuhn_nnz(64) == 4_018_753


# ---- convergence to the exact quickstart solution (README.md:94-133)
def _quickstart(m):
    pts, conn, bnd = structured_rectangle_arrays(m)
    n = len(pts)
    u, v, pe = op.pattern_and_slots(conn, n)
    vals = oa.reduce(pe, oa.element_matrices(pts, conn, "triangle", "laplace"), len(u))
    f = 2 * math.pi ** 2 * np.sin(math.pi * pts[:, 0]) * np.sin(math.pi * pts[:, 1])
    b = oa.load_vector(pts, conn, "triangle", n, f=f)
    kv, kr, kc, ni, fi, lay = oc.condense(vals, u, v, bnd, b)
    x = osv.direct(osv.csr(kv, kr, kc, ni), fi)
    uf = oc.recover(x, bnd)
    ue = np.sin(math.pi * pts[:, 0]) * np.sin(math.pi * pts[:, 1])
    return np.linalg.norm(uf - ue) / np.linalg.norm(ue)


def test_quickstart_second_order_convergence():
    e = [_quickstart(m) for m in (8, 16, 32)]
    assert 3.5 < e[0] / e[1] < 4.5 and 3.5 < e[1] / e[2] < 4.5
    assert _quickstart(20) < 1e-2                               # h = 0.05 (reference: 3.1e-3 on gmsh)


# ---- condensation (tests/operator/test_condenser.py:48-99)
def test_condense_tridiagonal_fixture():
    row = np.array([0, 0, 1, 1, 1, 2, 2, 2, 3, 3])
    col = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3])
    val = np.array([2., -1., -1., 2., -1., -1., 2., -1., -1., 2.])
    mask = np.array([True, False, False, True])
    kv, kr, kc, ni, fi, lay = oc.condense(val, row, col, mask, np.array([1., 0., 0., 1.]),
                                          np.array([1.0, 2.0]))
    assert ni == 2 and list(kr) == [0, 0, 1, 1] and list(kc) == [0, 1, 0, 1]
    assert np.allclose(kv, [2, -1, -1, 2])
    A = scipy.sparse.coo_matrix((val, (row, col))).toarray()
    uo = np.array([1.0, 2.0])
    assert np.allclose(fi, np.array([0., 0.]) - A[1:3][:, [0, 3]] @ uo)   # dense block elimination
    full = oc.recover(osv.direct(osv.csr(kv, kr, kc, ni), fi), mask, np.array([1.0, 0, 0, 2.0]))
    assert np.allclose(A @ full - np.array([0, 0, 0, 0]), [A[0] @ full, 0, 0, A[3] @ full])
    assert full[0] == 1.0 and full[3] == 2.0


# ---- solver
def test_jacobi_pcg_matches_direct(jittered_tri):
    pts, conn, bnd = jittered_tri
    n = len(pts)
    u, v, pe = op.pattern_and_slots(conn, n)
    kappa = np.exp(np.random.default_rng(5).standard_normal(n))
    vals = oa.reduce(pe, oa.element_matrices(pts, conn, "triangle", "kappa_laplace", kappa=kappa), len(u))
    kv, kr, kc, ni, fi, _ = oc.condense(vals, u, v, bnd, oa.load_vector(pts, conn, "triangle", n, c=1.0))
    A = osv.csr(kv, kr, kc, ni)
    x, it, hist = osv.jacobi_pcg(A, fi, rtol=1e-10)
    assert hist[-1] <= 1e-10 and osv.relres(A, x, fi) < 1e-9
    assert np.allclose(x, osv.direct(A, fi), rtol=1e-8, atol=1e-12)
    assert it <= ni + 5                                       # CG terminates in <= n steps (exact arithmetic)


# ---- hyperelasticity (builtin.py:285-315)
@pytest.fixture(scope="module")
def nh_case():
    pts, conn, bnd = kuhn_cube_arrays(2)
    rng = np.random.default_rng(3)
    u = 0.05 * rng.standard_normal(pts.shape)
    E = 1.0 + rng.uniform(-0.1, 0.1, len(conn))
    return pts, conn, u, E


def test_neohookean_zero_displacement(nh_case):
    pts, conn, u, E = nh_case
    assert oh.energy(pts, conn, np.zeros_like(pts), E) == 0.0
    assert np.abs(oh.residual(pts, conn, np.zeros_like(pts), E, len(pts))).max() < 1e-14


def test_neohookean_residual_is_energy_gradient(nh_case):
    pts, conn, u, E = nh_case
    R = oh.residual(pts, conn, u, E, len(pts))
    h = 1e-6
    rng = np.random.default_rng(0)
    for _ in range(6):
        i, c = rng.integers(len(pts)), rng.integers(3)
        up, um = u.copy(), u.copy()
        up[i, c] += h
        um[i, c] -= h
        fd = (oh.energy(pts, conn, up, E) - oh.energy(pts, conn, um, E)) / (2 * h)
        assert abs(fd - R[i, c]) < 1e-7 * max(1.0, abs(R[i, c]))


def test_neohookean_tangent_is_residual_jacobian(nh_case):
    pts, conn, u, E = nh_case
    n = len(pts)
    eu, ev, pe = op.pattern_and_slots(conn, n)
    T = oh.tangent_blocks(pts, conn, u, E, pe, len(eu))
    dense = np.zeros((3 * n, 3 * n))
    for s in range(len(eu)):
        dense[3 * eu[s]:3 * eu[s] + 3, 3 * ev[s]:3 * ev[s] + 3] = T[s]
    h = 1e-6
    for j in [0, 7, 20, 41, 3 * n - 1]:
        up, um = u.copy().ravel(), u.copy().ravel()
        up[j] += h
        um[j] -= h
        col = (oh.residual(pts, conn, up.reshape(-1, 3), E, n) - oh.residual(pts, conn, um.reshape(-1, 3), E, n)) / (2 * h)
        assert np.abs(col.ravel() - dense[:, j]).max() < 1e-6
    assert np.abs(dense - dense.T).max() < 1e-12


def test_neohookean_matches_torch_autograd(nh_case):
    torch = pytest.importorskip("torch")
    pts, conn, u, E = nh_case
    nu = 0.3
    mu = torch.tensor(E / (2 * (1 + nu)))
    lam = torch.tensor(E * nu / ((1 + nu) * (1 - 2 * nu)))
    phi, grad, jxw = oa.geometry(pts, conn, "tetra", 2)
    ut = torch.tensor(u, requires_grad=True)
    g = torch.tensor(grad[:, 0])
    H = torch.einsum("ebi,ebj->eij", ut[torch.tensor(conn)], g)
    F = torch.eye(3, dtype=torch.float64) + H
    J = torch.linalg.det(F)
    psi = mu / 2 * ((F * F).sum((1, 2)) - 3) - mu * torch.log(J) + lam / 2 * torch.log(J) ** 2
    (torch.tensor(jxw.sum(1)) * psi).sum().backward()
    assert np.allclose(ut.grad.numpy(), oh.residual(pts, conn, u, E, len(pts)), atol=1e-13)


def test_dresidual_dE_is_finite_difference(nh_case):
    pts, conn, u, E = nh_case
    lam_adj = np.random.default_rng(9).standard_normal(pts.shape)
    g = oh.dresidual_dE(pts, conn, u, E, lam_adj)
    h = 1e-6
    for e in [0, 5, 17]:
        Ep, Em = E.copy(), E.copy()
        Ep[e] += h
        Em[e] -= h
        fd = (np.sum(lam_adj * oh.residual(pts, conn, u, Ep, len(pts)))
              - np.sum(lam_adj * oh.residual(pts, conn, u, Em, len(pts)))) / (2 * h)
        assert abs(fd - g[e]) < 1e-8


def test_kappa_gradient_is_finite_difference(jittered_tri):
    pts, conn, _ = jittered_tri
    n = len(pts)
    u, v, pe = op.pattern_and_slots(conn, n)
    rng = np.random.default_rng(4)
    kappa = np.exp(rng.standard_normal(n))
    g = rng.standard_normal(len(u))
    dk = oa.kappa_gradient(pts, conn, "triangle", g, pe, n)

    def L(k):
        return g @ oa.reduce(pe, oa.element_matrices(pts, conn, "triangle", "kappa_laplace", kappa=k), len(u))
    for c in [0, 9, 30]:
        kp, km = kappa.copy(), kappa.copy()
        kp[c] += 1e-6
        km[c] -= 1e-6
        assert abs((L(kp) - L(km)) / 2e-6 - dk[c]) < 1e-7


def test_bicgstab_oracle_pins():
    """BiCGSTAB (torch_solve.py:130-248 order): non-symmetric systems against SuperLU,
    one step on a diagonal system, and no step for b = 0."""
    rng = np.random.default_rng(11)
    n = 400
    B = scipy.sparse.random(n, n, density=0.02, random_state=11, format="csr")
    d = np.asarray(abs(B).sum(1)).ravel() + 1.0
    A = (B + scipy.sparse.diags(d)).tocsr()
    assert abs(A - A.T).max() > 0.1                                  # genuinely non-symmetric
    b = rng.standard_normal(n)
    x, it, rel = osv.bicgstab(A, b, rtol=1e-12)
    assert rel <= 1e-12 and np.linalg.norm(x - osv.direct(A, b)) <= 1e-10 * np.linalg.norm(x)
    D = scipy.sparse.diags(np.linspace(1.0, 3.0, 50)).tocsr()
    bd = rng.standard_normal(50)
    x, it, rel = osv.bicgstab(D, bd, rtol=1e-14)
    assert it == 1 and np.allclose(x, bd / np.linspace(1.0, 3.0, 50), rtol=1e-14, atol=0)
    x, it, rel = osv.bicgstab(A, np.zeros(n))
    assert it == 0 and not x.any()


def test_bicgstab_oracle_zero_diagonal_guard():
    """A stored diagonal below 1e-12 counts as 1 in the Jacobi set-up
    (torch_solve.py:177-182): BiCGSTAB then solves a system with explicit zero
    diagonals (without the guard its first M^-1 p is inf)."""
    from tests.matrices import zero_diagonal_rows
    A, pr = zero_diagonal_rows(4000, 9, 40)
    assert np.all(A.diagonal()[pr] == 0.0) and all(i in set(A[i].indices) for i in pr)
    d = osv.diagonal(A)
    assert np.all(d[pr] == 1.0) and np.all(d[pr + 1] == 1.0)
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    x, it, rel = osv.bicgstab(A, b, rtol=1e-12, maxiter=5000)
    xd = osv.direct(A, b)
    assert rel <= 1e-12 and np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)


def test_torch_cpu_baseline_matches_oracle():
    """The torch-CPU baseline's Map + Reduce and condensation equal the numpy oracle."""
    import torch
    from oracle import torch_cpu as tc
    pts, conn, bnd = kuhn_cube_arrays(4)
    kappa = np.exp(np.random.default_rng(2).standard_normal(len(pts)))
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    got = tc.assemble_kappa_laplace(torch.tensor(pts), torch.tensor(conn), torch.tensor(kappa), torch.tensor(pe),
                                    len(u), chunk=100).numpy()
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, "tetra", "kappa_laplace", kappa=kappa), len(u))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    kv, kr, kc, ni, fi = tc.condense(torch.tensor(got), torch.tensor(u.astype(np.int64)),
                                     torch.tensor(v.astype(np.int64)), torch.tensor(bnd),
                                     torch.ones(len(pts), dtype=torch.float64))
    rv, rr, rc, rni, rfi, _ = oc.condense(ref, u, v, bnd, np.ones(len(pts)))
    assert ni == rni and np.array_equal(kr.numpy(), rr) and np.array_equal(kc.numpy(), rc)
    assert np.abs(kv.numpy() - rv).max() <= 1e-13 * np.abs(rv).max() and np.allclose(fi.numpy(), rfi)
    import scipy.sparse as sp
    A = sp.csr_matrix((rv, (rr, rc)), shape=(ni, ni))
    rp = torch.zeros(ni + 1, dtype=torch.int64)
    rp[1:] = torch.cumsum(torch.bincount(kr, minlength=ni), 0)
    At = torch.sparse_csr_tensor(rp, kc, kv, (ni, ni))
    d = torch.tensor(A.diagonal())
    x, it = tc.jacobi_pcg(At, fi, 1.0 / d, rtol=1e-10)
    xr, itr, _ = osv.jacobi_pcg(A, rfi, rtol=1e-10)
    assert abs(it - itr) <= 1 and np.abs(x.numpy() - xr).max() <= 1e-8 * np.abs(xr).max()

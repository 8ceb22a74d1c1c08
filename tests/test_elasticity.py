"""Linear elasticity (``LinearElasticityElementAssembler``, ``assemble/builtin.py:103-179``):
the oracle pinned by rigid-body modes, the patch-test energy and the
neo-Hookean tangent at rest, and the package's generic (torch) path against it."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import elasticity as oe
from oracle import hyperelastic as oh
from oracle import pattern as op
from tensormesh_amd.assemble.elasticity import voigt_shape_grad, voigt_stiffness
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays

NU = 0.3     # lam != mu, so a transposed block cannot pass the rigid-mode test


def _tet_case(n=2, seed=0):
    pts, conn, bnd = kuhn_cube_arrays(n)
    rng = np.random.default_rng(seed)
    pts = pts + (~bnd)[:, None] * rng.uniform(-0.1, 0.1, pts.shape) / n
    return pts, conn


def _tri_case(m=4, seed=0):
    pts, conn, _ = structured_rectangle_arrays(m, jitter=0.2, seed=seed)
    return pts, conn


def _rigid_modes(pts):
    d = pts.shape[1]
    modes = [np.tile(t, len(pts)) for t in np.eye(d)]
    if d == 3:
        modes += [np.cross(w, pts).reshape(-1) for w in np.eye(3)]
    else:
        modes.append(np.stack([-pts[:, 1], pts[:, 0]], 1).reshape(-1))
    return modes


def _assembled(pts, conn, et, E):
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    return oe.stiffness_blocks(pts, conn, et, E, pe, len(eu), NU), eu, ev


@pytest.mark.parametrize("et", ["tetra", "triangle"])
def test_oracle_rigid_modes_and_patch_energy(et):
    pts, conn = _tet_case() if et == "tetra" else _tri_case()
    d = pts.shape[1]
    E = 2.5
    blocks, eu, ev = _assembled(pts, conn, et, E)
    A = oe.dense(blocks, eu, ev, len(pts))
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()
    for r in _rigid_modes(pts):
        assert np.abs(A @ r).max() <= 1e-13 * np.abs(A).max()
    # patch test: u = G x has constant strain; u^T K u = vol (lam tr(eps)^2 + 2 mu eps:eps)
    rng = np.random.default_rng(1)
    G = rng.standard_normal((d, d))
    u = (pts @ G.T).reshape(-1)
    eps = 0.5 * (G + G.T)
    mu, lam = oe.lame(E, NU)
    vol = 1.0                                               # unit square / cube
    expect = vol * (lam * np.trace(eps) ** 2 + 2 * mu * (eps * eps).sum())
    assert abs(u @ A @ u - expect) <= 1e-12 * expect
    # a transposed block (lam and mu cross terms swapped) breaks rotation invariance
    At = oe.dense(np.transpose(blocks, (0, 2, 1)), eu, ev, len(pts))
    assert max(np.abs(At @ r).max() for r in _rigid_modes(pts)) > 1e-3 * np.abs(At).max()


def test_oracle_equals_neohookean_tangent_at_rest():
    pts, conn = _tet_case(seed=2)
    E = 1.0 + np.random.default_rng(2).uniform(-0.1, 0.1, len(conn))
    blocks, eu, ev = _assembled(pts, conn, "tetra", E)
    _, _, pe = op.pattern_and_slots(conn, len(pts))
    T = oh.tangent_blocks(pts, conn, np.zeros_like(pts), E, pe, len(eu), NU)
    assert np.abs(blocks - T).max() <= 1e-14 * np.abs(T).max()


def test_voigt_helpers():
    g = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64)
    B = voigt_shape_grad(g)
    assert torch.equal(B, torch.tensor([[1., 0, 0], [0, 2, 0], [0, 0, 3], [0, 3, 2], [3, 0, 1], [2, 1, 0]],
                                       dtype=torch.float64))
    B2 = voigt_shape_grad(torch.tensor([1.0, 2.0], dtype=torch.float64))
    assert torch.equal(B2, torch.tensor([[1., 0], [0, 2], [2, 1]], dtype=torch.float64))
    for d in (2, 3):
        C = voigt_stiffness(2.0, NU, d).numpy()
        assert np.allclose(C, oe.voigt_C(2.0, NU, d), rtol=0, atol=1e-15)


@pytest.mark.parametrize("et", ["tetra", "triangle"])
def test_generic_path_matches_oracle(et):
    """CPU tensors run the reference's own evaluation (``forward`` under vmap);
    block COO in the ``from_block_coo`` order."""
    pts, conn = _tet_case(seed=3) if et == "tetra" else _tri_case(seed=3)
    d = pts.shape[1]
    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    asm = tm.LinearElasticityElementAssembler.from_mesh(mesh, E=3.0, nu=NU)
    K = asm(mesh.points)
    blocks, eu, ev = _assembled(pts, conn, et, 3.0)
    assert K.shape == (len(pts) * d, len(pts) * d)
    ref = tm.SparseMatrix.from_block_coo(torch.from_numpy(blocks), torch.from_numpy(eu), torch.from_numpy(ev),
                                         (len(pts), len(pts)))
    assert torch.equal(K.row.long(), ref.row.long()) and torch.equal(K.col.long(), ref.col.long())
    got = K.values.double().numpy()
    assert np.abs(got - ref.values.numpy()).max() <= 1e-12 * np.abs(blocks).max()


def test_element_energy_is_half_strain_work():
    asm_cls = tm.LinearElasticityElementAssembler
    obj = asm_cls.__new__(asm_cls)
    obj.E, obj.nu = 2.0, NU
    G = torch.tensor([[0.1, 0.02, -0.03], [0.04, -0.05, 0.01], [0.0, 0.03, 0.02]], dtype=torch.float64)
    eps = 0.5 * (G + G.T)
    C = voigt_stiffness(2.0, NU, 3)
    ev = torch.stack([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[1, 2], 2 * eps[0, 2], 2 * eps[0, 1]])
    assert abs(float(obj.element_energy(G)) - 0.5 * float(ev @ C @ ev)) <= 1e-15

"""General constant-coefficient P1 forms (``forward(u, gradu, v, gradv)`` = U^T D V):
oracle pins, the form probe, and the generic (torch) path against the oracle."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import assemble as oa
from oracle import pattern as op
from tensormesh_amd import _lib
from tensormesh_amd.assemble import forms
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays


def _case(et, seed=0):
    if et == "tetra":
        pts, conn, bnd = kuhn_cube_arrays(2)
        rng = np.random.default_rng(seed)
        pts = pts + (~bnd)[:, None] * rng.uniform(-0.1, 0.1, pts.shape) / 2
    else:
        pts, conn, _ = structured_rectangle_arrays(4, jitter=0.2, seed=seed)
    return pts, conn


def _D(d, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((d + 1, d + 1))


@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_oracle_reductions_and_transpose(et):
    pts, conn = _case(et)
    d = pts.shape[1]
    Dl = np.diag([0.0] + [1.0] * d)
    assert np.allclose(oa.element_matrices_general(pts, conn, et, Dl), oa.element_matrices(pts, conn, et, "laplace"),
                       rtol=0, atol=1e-14)
    Dm = np.zeros((d + 1, d + 1))
    Dm[0, 0] = 1.0
    assert np.allclose(oa.element_matrices_general(pts, conn, et, Dm), oa.element_matrices(pts, conn, et, "mass"),
                       rtol=0, atol=1e-15)
    D = _D(d, 1)
    K, Kt = oa.element_matrices_general(pts, conn, et, D), oa.element_matrices_general(pts, conn, et, D.T)
    assert np.abs(Kt - np.transpose(K, (0, 2, 1))).max() <= 1e-14 * np.abs(K).max()


@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_oracle_convection_closed_form(et):
    """D[0, 1:] = beta: sum_ab X_a K_ab Y_b = int f1 (beta . grad f2) for linear f1, f2;
    with f1 = x_0 and grad f2 = G over the unit square / cube that is 0.5 beta . G."""
    pts, conn = _case(et, seed=2)
    d = pts.shape[1]
    beta = np.arange(1.0, d + 1.0)
    G = np.linspace(-1.0, 0.7, d)
    D = np.zeros((d + 1, d + 1))
    D[0, 1:] = beta
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    vals = oa.reduce(pe, oa.element_matrices_general(pts, conn, et, D), len(eu))
    X, Y = pts[:, 0], pts @ G
    total = float(np.sum(X[eu] * vals * Y[ev]))
    assert abs(total - 0.5 * beta @ G) <= 1e-13


def test_probe_recognises_and_rejects():
    A = np.array([[2.0, 0.5], [-0.3, 1.0]])
    b = np.array([0.7, -1.1])

    def f(u, gradu, v, gradv):
        return gradu @ torch.tensor(A) @ gradv + (torch.tensor(b) @ gradu) * v + 3.0 * u * v

    r = forms.classify_bilinear(f, 2, {}, {}, {})
    assert r is not None and r[0] == _lib.FORM_GENERAL
    D = r[1]
    assert np.allclose(D[1:, 1:], A, rtol=0, atol=1e-15) and np.allclose(D[1:, 0], b) and D[0, 0] == 3.0
    assert np.all(D[0, 1:] == 0.0)
    # isotropic Laplace keeps its dedicated kernel
    assert forms.classify_bilinear(lambda gradu, gradv: gradu @ gradv, 2, {}, {}, {})[0] == _lib.FORM_LAPLACE
    # not bilinear, or not constant: generic path
    assert forms.classify_bilinear(lambda gradu, gradv: (gradu @ gradv) ** 2, 2, {}, {}, {}) is None
    assert forms.classify_bilinear(lambda u, v: torch.sin(u) * v, 2, {}, {}, {}) is None


@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_generic_path_matches_oracle(et):
    pts, conn = _case(et, seed=3)
    d = pts.shape[1]
    D = _D(d, 4)
    Dt = torch.tensor(D)

    class Custom(tm.ElementAssembler):
        def forward(self, u, gradu, v, gradv):
            U = torch.cat([u.reshape(1), gradu])
            V = torch.cat([v.reshape(1), gradv])
            return U @ Dt.to(U.dtype) @ V

    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    K = Custom.from_mesh(mesh)(mesh.points)
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices_general(pts, conn, et, D), len(eu))
    assert np.array_equal(K.row.numpy(), eu) and np.array_equal(K.col.numpy(), ev)
    assert np.abs(K.values.double().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_form_cache_follows_coefficients():
    """A classification is reused only while what the forward reads is unchanged:
    closures, instance attributes, small module-level tensors."""
    pts, conn = _case("triangle")
    mesh = tm.Mesh(torch.from_numpy(pts), {"triangle": torch.from_numpy(conn)})
    asm = tm.LaplaceElementAssembler.from_mesh(mesh)

    def scaled(c):
        return lambda gradu, gradv: c * (gradu @ gradv)

    assert asm._classify(scaled(2.0), {}, {}, {})[1] == 2.0
    assert asm._classify(scaled(3.0), {}, {}, {})[1] == 3.0
    A = torch.eye(2, dtype=torch.float64)

    def aniso(gradu, gradv):
        return gradu @ A @ gradv

    r1 = asm._classify(aniso, {}, {}, {})
    A[0, 1] = 0.5
    r2 = asm._classify(aniso, {}, {}, {})
    assert r1 == (tm._lib.FORM_LAPLACE, 1.0, None)                 # isotropic: the Laplace kernel
    assert r2[0] == tm._lib.FORM_GENERAL and r2[1][1, 2] == 0.5 and r2[1][2, 1] == 0.0

    class Scaled(tm.ElementAssembler):
        def __post_init__(self, c=1.0):
            self.c = c

        def forward(self, gradu, gradv):
            return self.c * (gradu @ gradv)

    s = Scaled.from_mesh(mesh, c=2.0)
    assert s._classify(s.forward, {}, {}, {})[1] == 2.0
    s.c = 5.0
    assert s._classify(s.forward, {}, {}, {})[1] == 5.0

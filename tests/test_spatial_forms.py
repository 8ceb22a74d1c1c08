"""Spatially varying forms on the CPU: the oracle's quadrature-level Map
(``element_matrices_quadrature``, ``load_vector_quadrature``) pinned by closed
forms, and the per-element moments that the probing of ``forward`` produces
for the moments kernels, turned into element matrices / load vectors by the
kernels' formula and compared with the oracle."""
import math

import numpy as np
import pytest
import torch

from oracle import assemble as oa
from tensormesh_amd.assemble import forms
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays


def _meshes():
    pts, conn, _ = structured_rectangle_arrays(5, jitter=0.25, seed=3)
    yield "triangle", pts, conn
    pts, conn, _ = kuhn_cube_arrays(3)
    rng = np.random.default_rng(4)
    pts = pts + 0.05 * rng.uniform(-1, 1, pts.shape) * (np.abs(pts - 0.5) < 0.49)
    yield "tetra", pts, conn


def _bary_integral(alpha, d, vol):
    """int_T prod lambda_i^alpha_i = d! |T| prod alpha_i! / (sum alpha + d)!"""
    return math.factorial(d) * vol * np.prod([math.factorial(a) for a in alpha]) / math.factorial(sum(alpha) + d)


@pytest.mark.parametrize("et,pts,conn", list(_meshes()))
def test_oracle_linear_coefficient_laplace_closed_form(et, pts, conn):
    """kappa(x) linear: int kappa grad phi_a . grad phi_b = kappa(centroid) |T| grad phi_a . grad phi_b."""
    d = pts.shape[1]
    cvec = np.array([2.0, -1.5, 0.5][:d])
    K = oa.element_matrices_quadrature(pts, conn, et, lambda u, gu, v, gv, at:
                                       (1.0 + at["x"] @ cvec) * (gu * gv).sum(-1), order=2)
    _, grad, jxw = oa.geometry(pts, conn, et, 2)
    vol = jxw.sum(1)
    cen = pts[conn].mean(1)
    ref = ((1.0 + cen @ cvec) * vol)[:, None, None] * np.einsum("ead,ebd->eab", grad[:, 0], grad[:, 0])
    assert np.abs(K - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("et,pts,conn", list(_meshes()))
def test_oracle_coordinate_weighted_mass_closed_form(et, pts, conn):
    """int x_0 phi_a phi_b = sum_c x_0(c) int lambda_a lambda_b lambda_c, exact with the order-3 rule."""
    d = pts.shape[1]
    nb = d + 1
    K = oa.element_matrices_quadrature(pts, conn, et, lambda u, gu, v, gv, at: at["x"][..., 0] * u * v, order=3)
    _, _, jxw = oa.geometry(pts, conn, et, 3)
    vol = jxw.sum(1)
    ref = np.zeros_like(K)
    for a in range(nb):
        for b in range(nb):
            for c in range(nb):
                alpha = np.zeros(nb, int)
                np.add.at(alpha, [a, b, c], 1)
                ref[:, a, b] += pts[conn[:, c], 0] * np.array([_bary_integral(alpha, d, v) for v in vol])
    assert np.abs(K - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("et,pts,conn", list(_meshes()))
def test_oracle_coordinate_load_closed_form(et, pts, conn):
    """int x_0 phi_a = sum_c x_0(c) d! |T| (1 + delta_ac) / (d + 2)!."""
    d = pts.shape[1]
    nb = d + 1
    b = oa.load_vector_quadrature(pts, conn, et, len(pts), lambda v, gv, at: at["x"][..., 0] * v, order=2)
    _, _, jxw = oa.geometry(pts, conn, et, 2)
    vol = jxw.sum(1)
    be = np.zeros((len(conn), nb))
    for a in range(nb):
        for c in range(nb):
            be[:, a] += pts[conn[:, c], 0] * math.factorial(d) * vol * (2 if a == c else 1) / math.factorial(d + 2)
    ref = np.bincount(conn.ravel(), weights=be.ravel(), minlength=len(pts))
    assert np.abs(b - ref).max() <= 1e-13 * np.abs(ref).max()


def _kernel_formula(pts, conn, mom, nb, d):
    """The moments kernel's element matrix in the element's vertex order."""
    X = pts[conn]
    E = len(conn)
    out = np.zeros((E, nb, nb))
    MM = mom[:, :nb * nb].reshape(E, nb, nb)
    P = mom[:, nb * nb:nb * nb + nb * d].reshape(E, nb, d)
    Qm = mom[:, nb * nb + nb * d:nb * nb + 2 * nb * d].reshape(E, nb, d)
    S = mom[:, nb * nb + 2 * nb * d:].reshape(E, d, d)
    for e in range(E):
        J = (X[e, 1:] - X[e, 0]).T                   # columns x_b - x_0
        det = np.linalg.det(J)
        g = np.concatenate([-np.ones((1, d)), np.eye(d)]) @ np.linalg.inv(J)   # grad phi [nb, d]
        c = g * det
        out[e] = (abs(det) * MM[e] + np.sign(det) * (P[e] @ c.T + c @ Qm[e].T)
                  + c @ S[e] @ c.T / abs(det))
    return out


@pytest.mark.parametrize("et,pts,conn", list(_meshes()))
def test_probed_moments_reproduce_the_oracle(et, pts, conn):
    """Non-symmetric variable-coefficient form (anisotropic diffusion,
    convection, reaction, a nodal field and its gradient, element data):
    moments from probing forward, the kernel's formula, the oracle Map."""
    d = pts.shape[1]
    nb = d + 1
    rng = np.random.default_rng(7)
    k = rng.uniform(0.5, 2.0, len(pts))
    Ee = rng.uniform(1.0, 3.0, len(conn))

    def fwd(u, gradu, v, gradv, x, k, gradk, E):
        A = torch.eye(d, dtype=x.dtype) * (1.0 + x[0] ** 2) + 0.3 * torch.outer(x, x)
        beta = torch.stack([torch.sin(x[i] + i) for i in range(d)])
        return E * (gradu @ A @ gradv) + (beta @ gradu) * v + k * u * v + (gradk @ gradv) * u

    pt = torch.tensor(pts)
    q = forms.probe_bilinear_q(fwd, pt, torch.tensor(conn), et, 2, {"x": pt, "k": torch.tensor(k)},
                               {"E": torch.tensor(Ee)}, {})
    assert q is not None
    mom, sym = q
    assert not sym and mom.shape == (len(conn), (nb + d) ** 2)
    got = _kernel_formula(pts, conn, mom.numpy(), nb, d)

    def integrand(u, gu, v, gv, at):
        x = at["x"]
        r2 = (x ** 2)
        Agv = (1.0 + r2[..., :1]) * gv + 0.3 * x * (x * gv).sum(-1, keepdims=True)
        beta = np.stack([np.sin(x[..., i] + i) for i in range(d)], -1)
        return (at["E"] * (gu * Agv).sum(-1) + (beta * gu).sum(-1) * v + at["k"] * u * v
                + (at["gradk"] * gv).sum(-1) * u)
    ref = oa.element_matrices_quadrature(pts, conn, et, integrand, order=2, fields={"k": k},
                                         element_data={"E": Ee})
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("et,pts,conn", list(_meshes()))
def test_probed_load_moments_reproduce_the_oracle(et, pts, conn):
    d = pts.shape[1]
    nb = d + 1

    def fwd(x, v, gradv):
        return torch.sin(math.pi * x[0]) * torch.cos(x[1]) * v + (x @ gradv) * x[1]

    pt = torch.tensor(pts)
    mom = forms.probe_linear_q(fwd, pt, torch.tensor(conn), et, 2, {"x": pt}, {})
    assert mom is not None and mom.shape == (len(conn), nb + d)
    mom = mom.numpy()
    X = pts[conn]
    be = np.zeros((len(conn), nb))
    for e in range(len(conn)):
        J = (X[e, 1:] - X[e, 0]).T
        det = np.linalg.det(J)
        c = np.concatenate([-np.ones((1, d)), np.eye(d)]) @ np.linalg.inv(J) * det
        be[e] = abs(det) * mom[e, :nb] + np.sign(det) * c @ mom[e, nb:]
    got = np.bincount(conn.ravel(), weights=be.ravel(), minlength=len(pts))
    ref = oa.load_vector_quadrature(
        pts, conn, et, len(pts),
        lambda v, gv, at: np.sin(math.pi * at["x"][..., 0]) * np.cos(at["x"][..., 1]) * v
        + (at["x"] * gv).sum(-1) * at["x"][..., 1], order=2)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_probe_rejects_non_bilinear_forms():
    pts, conn, _ = structured_rectangle_arrays(3)
    pt, ct = torch.tensor(pts), torch.tensor(conn)
    pd = {"x": pt}
    for fn in (lambda u, v, x: u * u * v,                       # quadratic in u
               lambda gradu, gradv, x: (gradu @ gradv) + x[0],   # affine, not linear
               lambda gradu, gradv, x: (gradu @ gradv) * gradu[0]):
        assert forms.probe_bilinear_q(fn, pt, ct, "triangle", 2, pd, {}, {}) is None
    assert forms.probe_linear_q(lambda v, x: v * v * x[0], pt, ct, "triangle", 2, pd, {}) is None
    assert forms.probe_linear_q(lambda v, x: v + x[0], pt, ct, "triangle", 2, pd, {}) is None

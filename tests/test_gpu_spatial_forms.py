"""GPU parity of the spatially varying forms (``tm_asm_p1_moments``,
``tm_asm_p1_load_moments``) against the float64 oracle's quadrature-level Map:
forms reading ``x``, nodal fields and their gradients and element data
through a user ``forward`` (``element_assembler.py:332-346, 402-466``), FP32
and FP64 values, 2D and 3D, and ``func_node_assembler``
(``builtin.py:929-969``) plus a load with ``gradv``."""
import math

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import assemble as oa
from oracle import pattern as op
from tensormesh_amd.assemble import func_node_assembler
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays

pytestmark = pytest.mark.gpu


def _case(et, seed):
    if et == "tetra":
        pts, conn, bnd = kuhn_cube_arrays(6)
        rng = np.random.default_rng(seed)
        pts = pts + (~bnd)[:, None] * rng.uniform(-0.1, 0.1, pts.shape) / 6
    else:
        pts, conn, _ = structured_rectangle_arrays(14, jitter=0.25, seed=seed)
    return pts, conn


def _row_scaled(got, ref, eu):
    scale = np.zeros(eu.max() + 1)
    np.maximum.at(scale, eu, np.abs(ref))
    return (np.abs(got - ref) / scale[eu]).max()


@pytest.mark.parametrize("vdt,tol", [(torch.float32, 1e-6), (torch.float64, 1e-12)])
@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_variable_coefficient_form(cuda, et, vdt, tol):
    print(f"case spatial form {et} {vdt}", flush=True)
    pts, conn = _case(et, seed=8)
    d = pts.shape[1]
    rng = np.random.default_rng(9)
    k = rng.uniform(0.5, 2.0, len(pts))
    Ee = rng.uniform(1.0, 3.0, len(conn))

    class Variable(tm.ElementAssembler):
        def forward(self, u, gradu, v, gradv, x, k, gradk, E):
            A = torch.eye(d, dtype=x.dtype, device=x.device) * (1.0 + x[0] ** 2) + 0.3 * torch.outer(x, x)
            beta = torch.stack([torch.sin(x[i] + i) for i in range(d)])
            return E * (gradu @ A @ gradv) + (beta @ gradu) * v + k * u * v + (gradk @ gradv) * u

    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    asm = Variable.from_mesh(mesh).to(cuda)
    asm.value_dtype = vdt
    K = asm(point_data={"k": torch.tensor(k, device=cuda)}, element_data={"E": torch.tensor(Ee, device=cuda)})
    assert K.values.dtype == vdt and K.symmetric is False

    def integrand(u, gu, v, gv, at):
        x = at["x"]
        Agv = (1.0 + x[..., :1] ** 2) * gv + 0.3 * x * (x * gv).sum(-1, keepdims=True)
        beta = np.stack([np.sin(x[..., i] + i) for i in range(d)], -1)
        return (at["E"] * (gu * Agv).sum(-1) + (beta * gu).sum(-1) * v + at["k"] * u * v
                + (at["gradk"] * gv).sum(-1) * u)

    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices_quadrature(pts, conn, et, integrand, 2, {"k": k}, {"E": Ee}), len(eu))
    assert np.array_equal(K.row.cpu().numpy(), eu) and np.array_equal(K.col.cpu().numpy(), ev)
    err = _row_scaled(K.values.double().cpu().numpy(), ref, eu)
    assert err <= tol, err


@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_symmetric_variable_diffusion_flag_and_solve(cuda, et):
    """kappa(x) grad u . grad v: a symmetric variable form, flagged symmetric
    and solved by PCG (with a small reaction term so the system is regular)."""
    print(f"case spatial symmetric {et}", flush=True)
    pts, conn = _case(et, seed=10)

    class VarDiff(tm.ElementAssembler):
        def forward(self, u, gradu, v, gradv, x):
            return torch.exp(x.sum()) * (gradu @ gradv) + 0.1 * u * v

    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    K = VarDiff.from_mesh(mesh).to(cuda)()
    assert K.symmetric is True
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices_quadrature(
        pts, conn, et, lambda u, gu, v, gv, at: np.exp(at["x"].sum(-1)) * (gu * gv).sum(-1) + 0.1 * u * v),
        len(eu))
    assert _row_scaled(K.values.double().cpu().numpy(), ref, eu) <= 1e-6
    b = torch.ones(len(pts), dtype=torch.float64, device=cuda)
    x = K.solve(b, tol=1e-10, atol=0.0)
    assert K.last_solve_info["method"] == "jacobi-pcg" and K.last_solve_info["converged"]


@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_func_node_assembler_and_gradv_load(cuda, et):
    print(f"case spatial loads {et}", flush=True)
    pts, conn = _case(et, seed=11)
    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    f = lambda x: torch.sin(math.pi * x[..., 0]) * torch.cos(x[..., 1])
    b = func_node_assembler(f).from_mesh(mesh).to(cuda)()
    ref = oa.load_vector_quadrature(pts, conn, et, len(pts), lambda v, gv, at:
                                    np.sin(math.pi * at["x"][..., 0]) * np.cos(at["x"][..., 1]) * v)
    assert b.is_cuda and np.abs(b.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()

    class Flux(tm.NodeAssembler):
        def forward(self, x, gradv, v):
            return (x @ gradv) * x[1] + x[0] * v

    b2 = Flux.from_mesh(mesh).to(cuda)()
    ref2 = oa.load_vector_quadrature(pts, conn, et, len(pts), lambda v, gv, at:
                                     (at["x"] * gv).sum(-1) * at["x"][..., 1] + at["x"][..., 0] * v)
    assert np.abs(b2.cpu().numpy() - ref2).max() <= 1e-13 * np.abs(ref2).max()

"""Deterministic assembly (``TM_ASM_DETERMINISTIC``): repeated calls give
bit-identical values, the values stay at the float64 oracle, and Krylov
iteration counts on the assembled matrix repeat.  Covers the builtin matrix
kernels, the general-form and moments kernels and the load kernels."""
import math

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import assemble as oa
from oracle import pattern as op
from tensormesh_amd.assemble import func_node_assembler
from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays, structured_rectangle_arrays

pytestmark = pytest.mark.gpu


def _row_scaled(got, ref, eu):
    scale = np.zeros(eu.max() + 1)
    np.maximum.at(scale, eu, np.abs(ref))
    return (np.abs(got - ref) / scale[eu]).max()


@pytest.mark.parametrize("vdt,tol", [(torch.float32, 1e-6), (torch.float64, 1e-12)])
def test_kappa_laplace_bit_identical(cuda, vdt, tol):
    print(f"case deterministic kappa-Laplace {vdt}", flush=True)
    pts, conn, bnd = kuhn_cube_arrays(32)
    kappa = np.exp(2.0 * np.random.default_rng(1).standard_normal(len(pts)))
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    asm = tm.KappaLaplaceElementAssembler.from_mesh(m).to(cuda)
    asm.value_dtype = vdt
    asm.deterministic = True
    kd = torch.tensor(kappa, dtype=torch.float32, device=cuda)
    runs = [asm(point_data={"kappa": kd}).values.clone() for _ in range(4)]
    assert all(torch.equal(runs[0], r) for r in runs[1:])
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, "tetra", "kappa_laplace",
                                            kappa=kappa.astype(np.float32).astype(np.float64)), len(eu))
    assert _row_scaled(runs[0].double().cpu().numpy(), ref, eu) <= tol
    asm.deterministic = False
    plain = asm(point_data={"kappa": kd}).values
    assert _row_scaled(plain.double().cpu().numpy(), runs[0].double().cpu().numpy(), eu) <= tol


def test_mass_bit_identical(cuda):
    print("case deterministic mass", flush=True)
    pts, conn, _ = structured_rectangle_arrays(64, jitter=0.25, seed=3)
    m = Mesh(torch.from_numpy(pts), {"triangle": torch.from_numpy(conn)})
    asm = tm.MassElementAssembler.from_mesh(m).to(cuda)
    asm.deterministic = True
    runs = [asm().values.clone() for _ in range(3)]
    assert all(torch.equal(runs[0], r) for r in runs[1:])
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, "triangle", "mass"), len(eu))
    assert _row_scaled(runs[0].double().cpu().numpy(), ref, eu) <= 1e-6


def test_convection_diffusion_iterations_repeat(cuda):
    """Non-symmetric general-form matrix: bit-identical on every assembly, so
    BiCGSTAB (fixed-order reductions) repeats its iteration count and iterate."""
    print("case deterministic convection-diffusion", flush=True)
    pts, conn, bnd = structured_rectangle_arrays(128, jitter=0.2, seed=1)
    mesh = tm.Mesh(torch.from_numpy(pts), {"triangle": torch.from_numpy(conn)},
                   {"is_boundary": torch.from_numpy(bnd)})
    beta = torch.tensor([20.0, -10.0], dtype=torch.float64)

    class ConvDiff(tm.ElementAssembler):
        def forward(self, gradu, gradv, v):
            return gradu @ gradv + (beta.to(gradu.dtype) @ gradu) * v

    asm = ConvDiff.from_mesh(mesh).to(cuda)
    asm.deterministic = True
    load = tm.const_node_assembler(1.0).from_mesh(mesh).to(cuda)
    load.deterministic = True
    cond = tm.Condenser(mesh.boundary_mask.to(cuda))
    out = []
    for _ in range(3):
        K = asm(mesh.points.to(cuda))
        Ki, fi = cond(K, load())
        x = Ki.solve(fi, tol=1e-10, atol=0.0)
        out.append((K.values.clone(), Ki.last_solve_info["iterations"], x))
    for vals, its, x in out[1:]:
        assert torch.equal(vals, out[0][0]) and its == out[0][1] and torch.equal(x, out[0][2])


@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_loads_bit_identical(cuda, et):
    """Constant, nodal and spatially varying (moments) loads: bit-identical
    repeats, within 1e-13 of the oracle and of the default mode."""
    print(f"case deterministic loads {et}", flush=True)
    if et == "tetra":
        pts, conn, _ = kuhn_cube_arrays(24)
    else:
        pts, conn, _ = structured_rectangle_arrays(96, jitter=0.25, seed=5)
    m = Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    fn = np.exp(np.random.default_rng(6).standard_normal(len(pts)))
    fd = torch.tensor(fn, dtype=torch.float64, device=cuda)
    g = lambda x: torch.sin(math.pi * x[..., 0]) * torch.cos(x[..., 1]) + 2.0
    cases = [
        (tm.const_node_assembler(2.5).from_mesh(m).to(cuda), {}, oa.load_vector(pts, conn, et, len(pts), c=2.5)),
        (tm.SourceNodeAssembler.from_mesh(m).to(cuda), {"point_data": {"f": fd}},
         oa.load_vector(pts, conn, et, len(pts), f=fn)),
        (func_node_assembler(g).from_mesh(m).to(cuda), {},
         oa.load_vector_quadrature(pts, conn, et, len(pts), lambda v, gv, at:
                                   (np.sin(math.pi * at["x"][..., 0]) * np.cos(at["x"][..., 1]) + 2.0) * v)),
    ]
    for asm, kw, ref in cases:
        asm.deterministic = True
        runs = [asm(**kw).clone() for _ in range(4)]
        assert all(torch.equal(runs[0], r) for r in runs[1:])
        got = runs[0].cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
        asm.deterministic = False
        plain = asm(**kw).cpu().numpy()
        assert np.abs(plain - got).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("vdt,tol", [(torch.float32, 1e-6), (torch.float64, 1e-12)])
def test_spatially_varying_form_bit_identical(cuda, vdt, tol):
    """The moments kernel in deterministic mode: bit-identical repeats, at the
    oracle's quadrature-level Map."""
    print(f"case deterministic moments {vdt}", flush=True)
    pts, conn, _ = kuhn_cube_arrays(16)
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)})

    class VarDiff(tm.ElementAssembler):
        def forward(self, u, gradu, v, gradv, x):
            return torch.exp(x.sum()) * (gradu @ gradv) + x[0] * gradu[1] * v + 0.1 * u * v

    asm = VarDiff.from_mesh(m).to(cuda)
    asm.value_dtype = vdt
    asm.deterministic = True
    runs = [asm().values.clone() for _ in range(4)]
    assert all(torch.equal(runs[0], r) for r in runs[1:])
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices_quadrature(
        pts, conn, "tetra", lambda u, gu, v, gv, at: (np.exp(at["x"].sum(-1)) * (gu * gv).sum(-1)
                                                      + at["x"][..., 0] * gu[..., 1] * v + 0.1 * u * v)),
        len(eu))
    assert _row_scaled(runs[0].double().cpu().numpy(), ref, eu) <= tol

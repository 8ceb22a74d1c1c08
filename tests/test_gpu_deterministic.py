"""Deterministic assembly (``TM_ASM_DETERMINISTIC``): repeated calls give
bit-identical values, the values stay at the float64 oracle, and Krylov
iteration counts on the assembled matrix repeat."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import assemble as oa
from oracle import pattern as op
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
    f = tm.const_node_assembler(1.0).from_mesh(mesh).to(cuda)()
    cond = tm.Condenser(mesh.boundary_mask.to(cuda))
    out = []
    for _ in range(3):
        K = asm(mesh.points.to(cuda))
        Ki, fi = cond(K, f)
        x = Ki.solve(fi, tol=1e-10, atol=0.0)
        out.append((K.values.clone(), Ki.last_solve_info["iterations"], x))
    for vals, its, x in out[1:]:
        assert torch.equal(vals, out[0][0]) and its == out[0][1] and torch.equal(x, out[0][2])

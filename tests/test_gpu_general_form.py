"""GPU parity of the general constant-coefficient P1 kernel (``tm_asm_p1_general``)
against the float64 oracle: non-symmetric D in 2D and 3D, FP32 and FP64 values,
dispatch from a user ``forward``, and agreement with the Laplace kernel."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import assemble as oa
from oracle import pattern as op
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays

pytestmark = pytest.mark.gpu


def _case(et, seed):
    if et == "tetra":
        pts, conn, bnd = kuhn_cube_arrays(5)
        rng = np.random.default_rng(seed)
        pts = pts + (~bnd)[:, None] * rng.uniform(-0.1, 0.1, pts.shape) / 5
    else:
        pts, conn, _ = structured_rectangle_arrays(12, jitter=0.25, seed=seed)
    return pts, conn


def _custom(D):
    Dt = torch.tensor(D)

    class Custom(tm.ElementAssembler):
        def forward(self, u, gradu, v, gradv):
            U = torch.cat([u.reshape(1), gradu])
            V = torch.cat([v.reshape(1), gradv])
            return U @ Dt.to(U.dtype) @ V

    return Custom


def _row_scaled(got, ref, eu):
    scale = np.zeros(eu.max() + 1)
    np.maximum.at(scale, eu, np.abs(ref))
    return (np.abs(got - ref) / scale[eu]).max()


@pytest.mark.parametrize("vdt", [torch.float32, torch.float64])
@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_general_kernel_matches_oracle(cuda, et, vdt):
    print(f"case general form {et} {vdt}", flush=True)
    pts, conn = _case(et, seed=5)
    d = pts.shape[1]
    D = np.random.default_rng(6).standard_normal((d + 1, d + 1))
    cls = _custom(D)
    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    asm = cls.from_mesh(mesh).to(cuda)
    asm.value_dtype = vdt
    form = asm._classify(asm.forward, {}, {}, {})
    assert form is not None and form[0] == tm._lib.FORM_GENERAL
    K = asm(mesh.points.to(cuda))
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices_general(pts, conn, et, D), len(eu))
    assert np.array_equal(K.row.cpu().numpy(), eu) and np.array_equal(K.col.cpu().numpy(), ev)
    got = K.values.double().cpu().numpy()
    assert _row_scaled(got, ref, eu) <= (1e-6 if vdt == torch.float32 else 1e-12)
    # y = K x through the CSR product
    x = np.random.default_rng(7).standard_normal(len(pts))
    y = (K @ torch.tensor(x, device=cuda)).cpu().numpy()
    yr = np.bincount(eu, weights=ref * x[ev], minlength=len(pts))
    assert np.abs(y - yr).max() <= (1e-5 if vdt == torch.float32 else 1e-11) * np.abs(yr).max()


@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_general_kernel_equals_laplace_kernel(cuda, et):
    print(f"case general form = laplace {et}", flush=True)
    pts, conn = _case(et, seed=8)
    d = pts.shape[1]
    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    asm = tm.LaplaceElementAssembler.from_mesh(mesh).to(cuda)
    p64 = mesh.points.to(cuda)
    lap = asm._launch_matrix(tm._lib.FORM_LAPLACE, p64, None, 1.0, torch.float64)
    gen = asm._launch_general(p64, np.diag([0.0] + [1.0] * d), torch.float64)
    assert (gen - lap).abs().max().item() <= 1e-13 * lap.abs().max().item()


@pytest.mark.parametrize("et", ["triangle", "tetra"])
def test_anisotropic_diffusion_runs_the_general_kernel(cuda, et):
    print(f"case anisotropic diffusion {et}", flush=True)
    pts, conn = _case(et, seed=9)
    d = pts.shape[1]
    A = np.eye(d) + 0.3 * np.random.default_rng(10).standard_normal((d, d))
    At = torch.tensor(A)

    class Aniso(tm.ElementAssembler):
        def forward(self, gradu, gradv):
            return gradu @ At.to(gradu.dtype) @ gradv

    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)})
    asm = Aniso.from_mesh(mesh).to(cuda)
    asm.value_dtype = torch.float64
    assert asm._classify(asm.forward, {}, {}, {})[0] == tm._lib.FORM_GENERAL
    K = asm(mesh.points.to(cuda))
    D = np.zeros((d + 1, d + 1))
    D[1:, 1:] = A
    eu, ev, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices_general(pts, conn, et, D), len(eu))
    assert _row_scaled(K.values.cpu().numpy(), ref, eu) <= 1e-12

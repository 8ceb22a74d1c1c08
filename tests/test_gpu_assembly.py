"""GPU parity of the HIP assembly kernels against the float64 oracle."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import assemble as oa
from oracle import pattern as op
from tensormesh_amd.dataset import lognormal_kappa_cube
from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays, structured_rectangle_arrays

pytestmark = pytest.mark.gpu


def _mesh(kind, n=6, seed=0):
    if kind == "tri":
        pts, conn, bnd = structured_rectangle_arrays(n * 2, jitter=0.25, seed=seed)
        et = "triangle"
    else:
        pts, conn, bnd = kuhn_cube_arrays(n)
        rng = np.random.default_rng(seed)
        pts = pts + (~bnd)[:, None] * rng.uniform(-0.15, 0.15, pts.shape) / n   # jitter interior nodes
        et = "tetra"
    m = Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    return m, pts, conn, bnd, et


def row_scaled_err(got, ref, u, v):
    diag = np.zeros(u.max() + 1)
    d = u == v
    diag[u[d]] = np.abs(ref[d])
    return np.max(np.abs(got - ref) / diag[u])


@pytest.mark.parametrize("kind", ["tri", "tet"])
@pytest.mark.parametrize("form", ["laplace", "kappa_laplace", "mass"])
@pytest.mark.parametrize("vdt", [torch.float32, torch.float64])
def test_matrix_vs_oracle(cuda, kind, form, vdt):
    print(f"case matrix {kind} {form} {vdt}", flush=True)
    m, pts, conn, bnd, et = _mesh(kind)
    kappa = np.exp(np.random.default_rng(7).standard_normal(len(pts)))
    cls = {"laplace": tm.LaplaceElementAssembler, "mass": tm.MassElementAssembler,
           "kappa_laplace": tm.KappaLaplaceElementAssembler}[form]
    asm = cls.from_mesh(m).to(cuda)
    asm.value_dtype = vdt
    pd = {"kappa": torch.tensor(kappa, device=cuda, dtype=torch.float32)} if form == "kappa_laplace" else None
    K = asm(point_data=pd)
    assert K.values.dtype == vdt and K.values.is_cuda
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    assert np.array_equal(K.row.cpu().numpy(), u) and np.array_equal(K.col.cpu().numpy(), v)
    kk = kappa.astype(np.float32).astype(np.float64)
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, et, form, kappa=kk), len(u))
    got = K.values.double().cpu().numpy()
    tol = 1e-6 if vdt == torch.float32 else 1e-13
    assert row_scaled_err(got, ref, u, v) <= tol


def test_scaled_and_f64_coefficient(cuda):
    m, pts, conn, bnd, et = _mesh("tet", 4)
    kappa = np.exp(np.random.default_rng(1).standard_normal(len(pts)))

    class Scaled(tm.ElementAssembler):
        def forward(self, gradu, gradv, k):
            return 2.5 * k * (gradu @ gradv)

    asm = Scaled.from_mesh(m).to(cuda)
    asm.value_dtype = torch.float64
    K = asm(point_data={"k": torch.tensor(kappa, device=cuda)})
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    ref = 2.5 * oa.reduce(pe, oa.element_matrices(pts, conn, et, "kappa_laplace", kappa=kappa), len(u))
    assert np.abs(K.values.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()


def test_reaction_diffusion_form_runs_the_general_kernel(cuda):
    m, pts, conn, bnd, et = _mesh("tri", 3)

    class Reaction(tm.ElementAssembler):
        def forward(self, gradu, gradv, u, v):
            return gradu @ gradv + 3.0 * u * v

    asm = Reaction.from_mesh(m).to(cuda)
    assert asm._classify(asm.forward, {}, {}, {})[0] == tm._lib.FORM_GENERAL
    asm.value_dtype = torch.float64
    K = asm(m.points.to(cuda))
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, et, "laplace") + 3 * oa.element_matrices(pts, conn, et, "mass"),
                    len(u))
    assert np.abs(K.values.cpu().numpy() - ref).max() < 1e-12


def test_spatially_varying_form_runs_the_moments_kernel(cuda):
    """A coefficient that varies in space (here through the coordinates) is not a
    constant U^T D V: it is probed at the quadrature points and assembled by the
    moments kernel (FP32 values: 1e-6 row-scaled; FP64 values: 1e-12); a forward
    that is not bilinear runs the generic torch path in FP64."""
    m, pts, conn, bnd, et = _mesh("tri", 3)

    class Variable(tm.ElementAssembler):
        def forward(self, gradu, gradv, x):
            return (1.0 + x[0]) * (gradu @ gradv)

    asm = Variable.from_mesh(m).to(cuda)
    assert asm._classify(asm.forward, {"x": m.points}, {}, {}) is None
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, et, "kappa_laplace", kappa=1.0 + pts[:, 0]), len(u))
    K = asm(m.points.to(cuda))
    assert K.values.dtype == m.points.dtype                  # the reference's dtype rule
    asm.value_dtype = torch.float32
    K = asm(m.points.to(cuda))
    assert K.values.dtype == torch.float32
    scale = np.zeros(len(pts))
    np.maximum.at(scale, u, np.abs(ref))
    assert (np.abs(K.values.double().cpu().numpy() - ref) / scale[u]).max() <= 1e-6
    asm.value_dtype = torch.float64
    K = asm(m.points.to(cuda))
    assert np.abs(K.values.cpu().numpy() - ref).max() < 1e-12

    class NotBilinear(tm.ElementAssembler):
        def forward(self, gradu, gradv, x):
            return (1.0 + x[0]) * (gradu @ gradv) * (1.0 + 1e-3 * gradu[0] ** 2)

    ref2 = oa.reduce(pe, oa.element_matrices_quadrature(
        pts, conn, et, lambda uu, gu, vv, gv, at: (1.0 + at["x"][..., 0]) * (gu * gv).sum(-1)
        * (1.0 + 1e-3 * gu[..., 0] ** 2)), len(u))
    K = NotBilinear.from_assembler(asm)(m.points.to(cuda))
    assert K.values.dtype == torch.float64 and np.abs(K.values.cpu().numpy() - ref2).max() < 1e-12


@pytest.mark.parametrize("kind", ["tri", "tet"])
def test_load_vectors_vs_oracle(cuda, kind):
    print(f"case load {kind}", flush=True)
    m, pts, conn, bnd, et = _mesh(kind)
    f = np.sin(3 * pts[:, 0]) + pts[:, 1] ** 2
    b1 = tm.const_node_assembler(2.0).from_mesh(m).to(cuda)()
    ref1 = oa.load_vector(pts, conn, et, len(pts), c=2.0)
    assert np.abs(b1.cpu().numpy() - ref1).max() <= 1e-14 * np.abs(ref1).max()
    b2 = tm.SourceNodeAssembler.from_mesh(m).to(cuda)(point_data={"f": torch.tensor(f, device=cuda)})
    ref2 = oa.load_vector(pts, conn, et, len(pts), f=f)
    assert np.abs(b2.cpu().numpy() - ref2).max() <= 1e-13 * np.abs(ref2).max()


@pytest.mark.parametrize("kind", ["tri", "tet"])
def test_kappa_adjoint_vs_oracle(cuda, kind):
    print(f"case adjoint {kind}", flush=True)
    m, pts, conn, bnd, et = _mesh(kind, 4)
    rng = np.random.default_rng(2)
    kappa = torch.tensor(np.exp(rng.standard_normal(len(pts))), device=cuda, requires_grad=True)
    asm = tm.KappaLaplaceElementAssembler.from_mesh(m).to(cuda)
    asm.value_dtype = torch.float64
    K = asm(point_data={"kappa": kappa})
    g = torch.tensor(rng.standard_normal(K.nnz), device=cuda)
    (K.values * g).sum().backward()
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.kappa_gradient(pts, conn, et, g.cpu().numpy(), pe, len(pts))
    assert np.abs(kappa.grad.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    # central finite differences through the kernel itself (float64 values)
    for c in [0, len(pts) // 2]:
        kp = kappa.detach().clone()
        km = kappa.detach().clone()
        kp[c] += 1e-6
        km[c] -= 1e-6
        fd = ((asm(point_data={"kappa": kp}).values - asm(point_data={"kappa": km}).values) * g).sum() / 2e-6
        assert abs(fd.item() - kappa.grad[c].item()) <= 1e-6 * max(1.0, abs(kappa.grad[c].item()))


def test_assembly_is_reproducible(cuda):
    m, pts, conn, bnd, et = _mesh("tet", 8)
    asm = tm.KappaLaplaceElementAssembler.from_mesh(m).to(cuda)
    k = torch.rand(len(pts), device=cuda) + 0.5
    a = asm(point_data={"kappa": k}).values
    b = asm(point_data={"kappa": k}).values
    # FP64 accumulation in arbitrary order, rounded once to FP32: at most 1 ulp apart
    assert torch.all((a - b).abs() <= 1.2e-7 * a.abs() + 1e-30)


def test_full_size_128_properties(cuda):
    """128^3 Kuhn grid (12.58 M tets): size-independent checks of the kernel output."""
    n = 128
    print("case full-size 128^3", flush=True)
    pts, conn, bnd = kuhn_cube_arrays(n)
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    kappa = torch.tensor(lognormal_kappa_cube(n, 0.1, 1.0, 0), dtype=torch.float32, device=cuda)
    asm = tm.KappaLaplaceElementAssembler.from_mesh(m).to(cuda)
    K = asm(point_data={"kappa": kappa})
    assert K.nnz == op.kuhn_nnz(n) == 31_802_497
    vals = K.values.double()
    # K 1 = 0 row by row, relative to the diagonal
    rs = torch.zeros(K.shape[0], dtype=torch.float64, device=cuda).index_add_(0, K.row, vals)
    diag = vals[K.row == K.col]
    assert (rs.abs() / diag).max().item() < 2e-6
    assert (diag > 0).all()
    # symmetry through the transpose permutation of the canonical pattern
    key_t = K.col * K.shape[0] + K.row
    perm = torch.argsort(key_t)
    assert ((vals - vals[perm]).abs() / diag[K.row]).max().item() < 2e-6
    # checksum of a random row sample against the oracle restricted to their elements
    rows = np.random.default_rng(0).choice(len(pts), 64, replace=False)
    touch = np.isin(conn, rows).any(1)
    sub = conn[touch]
    u, v, pe = op.pattern_and_slots(sub, len(pts))
    ke = oa.element_matrices(pts, sub, "tetra", "kappa_laplace", kappa=kappa.double().cpu().numpy())
    ref = oa.reduce(pe, ke, len(u))
    sel = np.isin(u, rows)
    rp = asm.topology.row_ptr
    ci = asm.topology.col_idx
    got = vals.cpu().numpy()
    for r in rows:
        cols_ref = v[sel & (u == r)]
        vref = ref[sel & (u == r)]
        cols = ci[rp[r]:rp[r + 1]]
        assert np.array_equal(cols, cols_ref)
        d = np.abs(vref[cols_ref == r][0])
        assert np.abs(got[rp[r]:rp[r + 1]] - vref).max() <= 1e-6 * d

"""GPU parity of the edge-record matrix kernel (symmetric scalar forms): the
same values as the 16-byte-record kernel and the float64 oracle, on 8-byte
records (Kuhn tetrahedra, triangles) and 12-byte records (a Delaunay
tetrahedral mesh whose rows exceed 16 entries)."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import assemble as oa
from oracle import pattern as op
from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays, structured_rectangle_arrays
from tests.test_gpu_assembly import row_scaled_err
from tests.test_host import _delaunay_tets

pytestmark = pytest.mark.gpu

CLS = {"laplace": tm.LaplaceElementAssembler, "mass": tm.MassElementAssembler,
       "kappa_laplace": tm.KappaLaplaceElementAssembler}


def _mesh(kind):
    if kind == "tri":
        pts, conn, bnd = structured_rectangle_arrays(14, jitter=0.25, seed=3)
        et = "triangle"
    elif kind == "tet":
        pts, conn, bnd = kuhn_cube_arrays(7)
        et = "tetra"
    else:
        pts, conn = _delaunay_tets(600, 5)
        bnd = np.zeros(len(pts), dtype=bool)
        et = "tetra"
    m = Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    return m, pts, conn, et


def _without_edge(asm, pd, dev):
    """The same assembler on the 16-byte records (edge fields cleared)."""
    plan = asm.topology.device_plan(dev, edge=True)
    st = plan.struct
    saved = (st.erecs, st.eblk, st.emir, st.erows)
    st.erecs = st.eblk = st.emir = st.erows = None
    try:
        return asm(point_data=pd)
    finally:
        st.erecs, st.eblk, st.emir, st.erows = saved


@pytest.mark.parametrize("kind,words", [("tri", 2), ("tet", 2), ("delaunay", 3)])
@pytest.mark.parametrize("form", ["laplace", "kappa_laplace", "mass"])
@pytest.mark.parametrize("vdt", [torch.float32, torch.float64])
def test_edge_records_vs_oracle_and_full_records(cuda, kind, words, form, vdt):
    print(f"case edge {kind} {form} {vdt}", flush=True)
    m, pts, conn, et = _mesh(kind)
    kappa = np.exp(np.random.default_rng(11).standard_normal(len(pts)))
    asm = CLS[form].from_mesh(m).to(cuda)
    asm.value_dtype = vdt
    pd = {"kappa": torch.tensor(kappa, device=cuda, dtype=torch.float32)} if form == "kappa_laplace" else None
    K = asm(point_data=pd)
    plan = asm.topology.device_plan(cuda, edge=True)
    assert plan.has_edge and int(plan.edge_info[0]) == words
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    kk = kappa.astype(np.float32).astype(np.float64)
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, et, form, kappa=kk), len(u))
    got = K.values.double().cpu().numpy()
    # FP64: the accumulation order of the LDS atomics varies, and the edge kernel
    # forms a Laplace diagonal as minus its row sum, so the rounding scales with
    # the row's absolute mass sum_e |K_e| over its diagonal (about 2 on the
    # Kuhn and triangle meshes, up to 11 on the Delaunay one): up to ~100
    # roundings of 1.1e-16 times that ratio, bounded below by 1e-13
    absref = oa.reduce(pe, np.abs(oa.element_matrices(pts, conn, et, form, kappa=kk)), len(u))
    rowabs = np.zeros(len(pts))
    np.add.at(rowabs, u, absref)
    ratio = float(np.max(rowabs / np.abs(ref[u == v])))
    tol = 1e-6 if vdt == torch.float32 else max(1e-13, 128 * np.finfo(np.float64).eps * ratio)
    assert row_scaled_err(got, ref, u, v) <= tol
    full = _without_edge(asm, pd, cuda).values.double().cpu().numpy()
    assert row_scaled_err(full, ref, u, v) <= tol
    # the two kernels sum in different orders: last-bit differences only
    assert row_scaled_err(got, full, u, v) <= (2.5e-7 if vdt == torch.float32 else tol)


def test_deterministic_mode_keeps_full_records(cuda):
    """The deterministic mode rounds every entry to its row's grid before the
    sums, which needs both directed entries: it runs the 16-byte kernel and
    repeats bit for bit."""
    m, pts, conn, et = _mesh("tet")
    asm = tm.LaplaceElementAssembler.from_mesh(m).to(cuda)
    asm.value_dtype = torch.float32
    K0 = asm().values.clone()
    asm.deterministic = True
    K1, K2 = asm().values, asm().values
    assert torch.equal(K1, K2)
    u, v, _ = op.pattern_and_slots(conn, len(pts))
    assert row_scaled_err(K1.double().cpu().numpy(), K0.double().cpu().numpy(), u, v) <= 2.5e-7


def test_plan_beyond_edge_limits_keeps_full_records(cuda, monkeypatch):
    """One block of 729 rows (more than 8191 accumulators): the edge records
    are refused, the matrix kernel runs the 16-byte records (102 KiB of LDS)
    and matches the oracle."""
    monkeypatch.setenv("TM_ASM_MAX_ROWS", "1000")
    monkeypatch.setenv("TM_ASM_LDS_BUDGET", str(150 * 1024))
    pts, conn, bnd = kuhn_cube_arrays(8)
    m = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    asm = tm.LaplaceElementAssembler.from_mesh(m).to(cuda)
    asm.value_dtype = torch.float32
    K = asm()
    plan = asm.topology.device_plan(cuda, edge=True)
    assert plan.n_blocks == 1 and not plan.has_edge and plan.edge_refused
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, "tetra", "laplace"), len(u))
    assert row_scaled_err(K.values.double().cpu().numpy(), ref, u, v) <= 1e-6

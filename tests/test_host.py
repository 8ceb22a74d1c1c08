"""Host side of the native library (CPU only): exported symbols, bit-exact
pattern and element slots against the SciPy oracle, and a float64 emulation
of the block-plan kernel against the oracle matrices."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import assemble as oa
from oracle import pattern as op
from tensormesh_amd import _lib
from tensormesh_amd.assemble.topology import P1Topology
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays
from tests.plan_emulator import emulate_edge_matrix, emulate_matrix
from tensormesh_amd.assemble.topology import edge_plan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ERR_ARG = -1  # TM_ERR_ARG


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "tensormesh_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:const char\*|int64_t|int|void)\s+(tm_\w+)\(", hdr, re.M))
    assert declared, "no declarations parsed"
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), f"{name} missing from libtensormesh_amd.so"
    assert declared == set(_lib.declared_symbols())
    assert L.tm_version() >= 1


def _meshes():
    yield "tri", structured_rectangle_arrays(9, jitter=0.2, seed=5), "triangle"
    yield "tet", kuhn_cube_arrays(4), "tetra"


@pytest.mark.parametrize("name,arrays,et", list(_meshes()))
def test_pattern_and_slots_bit_exact(name, arrays, et):
    pts, conn, _ = arrays
    topo = P1Topology(pts, conn, len(pts), et)
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    assert topo.edges.dtype.is_floating_point is False
    assert np.array_equal(topo.edges[0].numpy(), u) and np.array_equal(topo.edges[1].numpy(), v)
    assert np.array_equal(topo.row_ptr, op.csr_from_edges(u, len(pts)))
    assert np.array_equal(topo.element_slots(), pe)


def test_pattern_with_isolated_node_and_shuffled_elements():
    pts, conn, _ = kuhn_cube_arrays(3)
    rng = np.random.default_rng(1)
    conn = conn[rng.permutation(len(conn))]
    conn = np.where(conn >= 10, conn + 1, conn)               # node 10 belongs to no element
    pts = np.insert(pts, 10, [[9.0, 9.0, 9.0]], axis=0)
    topo = P1Topology(pts, conn, len(pts), "tetra")
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    assert np.array_equal(topo.edges[0].numpy(), u) and np.array_equal(topo.edges[1].numpy(), v)
    assert topo.row_ptr[11] == topo.row_ptr[10]
    assert np.array_equal(topo.element_slots(), pe)


def test_bad_connectivity_raises():
    pts, conn, _ = kuhn_cube_arrays(1)
    conn = conn.copy()
    conn[0, 0] = 99
    with pytest.raises(ValueError, match="out of"):
        P1Topology(pts, conn, len(pts), "tetra")


def _kernel_element_matrix(topo, form, kappa, order=2):
    w, phi, G = oa.rule(topo.element_type, order)
    W = (w[:, None] * phi).sum(0)

    def f(e, x):
        J = x[1:] - x[0]
        det = np.linalg.det(J)
        Gi = np.linalg.inv(J)
        g = np.concatenate([-Gi.sum(1, keepdims=True), Gi], 1).T
        s = abs(det) * (w.sum() if form == "laplace" else (kappa[topo.conn[e]] * W).sum())
        return s * g @ g.T
    return f


@pytest.mark.parametrize("name,arrays,et", list(_meshes()))
@pytest.mark.parametrize("rows", [7, 33, 200])
def test_block_plan_emulation_matches_oracle(name, arrays, et, rows):
    pts, conn, _ = arrays
    topo = P1Topology(pts, conn, len(pts), et)
    arrs, info = topo.host_plan("scalar", max_rows=rows)
    assert info[5] <= 1024 and info[4] <= 64 * 1024
    kappa = np.exp(np.random.default_rng(3).standard_normal(len(pts)))
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    for form in ("laplace", "kappa_laplace"):
        got = emulate_matrix(topo, arrs, _kernel_element_matrix(topo, form, kappa))
        ref = oa.reduce(pe, oa.element_matrices(pts, conn, et, form, kappa=kappa), len(u))
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def _mass_element_matrix(topo, order=2):
    w, phi, _ = oa.rule(topo.element_type, order)
    M = (w[:, None, None] * phi[:, :, None] * phi[:, None, :]).sum(0)

    def f(e, x):
        return abs(np.linalg.det(x[1:] - x[0])) * M
    return f


def _delaunay_tets(n, seed):
    """Tetrahedra of a Delaunay mesh of n random points in the unit cube
    (rows longer than 16 entries: the 12-byte edge records)."""
    from scipy.spatial import Delaunay
    pts = np.random.default_rng(seed).random((n, 3))
    tri = Delaunay(pts)
    conn = tri.simplices.astype(np.int64)
    vol = np.abs(np.linalg.det(pts[conn[:, 1:]] - pts[conn[:, :1]]))
    return pts, conn[vol > 1e-9 * vol.max()]


@pytest.mark.parametrize("name,arrays,et", list(_meshes()) + [("delaunay", _delaunay_tets(300, 2) + (None,), "tetra")])
@pytest.mark.parametrize("rows", [7, 33, 200])
def test_edge_records_emulation_matches_oracle(name, arrays, et, rows):
    """Edge records: one accumulation per vertex pair in the row of the owned
    end with the smaller slot, mirrored entries, row-sum diagonals (Laplace,
    kappa-Laplace) or direct diagonals (mass), against the float64 oracle."""
    pts, conn, _ = arrays
    used = np.unique(conn)
    topo = P1Topology(pts, conn, len(pts), et)
    arrs, info = topo.host_plan("scalar", max_rows=rows)
    earrs, einfo = edge_plan(topo.nb, arrs, info)
    words = int(einfo[0])
    if name == "delaunay":
        assert einfo[3] > 15 and words == 3
    else:
        assert words == 2
    assert len(earrs["erecs"]) == words * int(info[2])
    assert (earrs["eblk"].reshape(-1, 4)[:, 1] & 0xFFFF).sum() == einfo[1] and einfo[6] == len(pts)
    kappa = np.exp(np.random.default_rng(3).standard_normal(len(pts)))
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    for form in ("laplace", "kappa_laplace", "mass"):
        f = _mass_element_matrix(topo) if form == "mass" else _kernel_element_matrix(topo, form, kappa)
        got = emulate_edge_matrix(topo, arrs, earrs, words, f, rowsum=form != "mass")
        ref = oa.reduce(pe, oa.element_matrices(pts, conn, et, form, kappa=kappa), len(u))
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), form
    assert len(used) == len(pts)


def test_edge_records_refused_beyond_limits_fall_back():
    """A plan whose blocks exceed the edge records' limits (here an LDS budget
    that lets a block's edge layout pass 64 KiB) keeps the 16-byte records:
    the builder reports TM_ERR_LIMIT and the device plan records why."""
    pts, conn, _ = kuhn_cube_arrays(12)
    topo = P1Topology(pts, conn, len(pts), "tetra")
    arrs, info = topo.host_plan("scalar", max_rows=100_000, lds_budget=1 << 20)
    with pytest.raises(_lib.TMError) as ei:
        edge_plan(topo.nb, arrs, info)
    assert ei.value.code == _lib.TM_ERR_LIMIT


def test_block_plan_respects_lds_budget():
    pts, conn, _ = kuhn_cube_arrays(6)
    topo = P1Topology(pts, conn, len(pts), "tetra")
    arrs, info = topo.host_plan("scalar", max_rows=10_000, lds_budget=24 * 1024)
    assert info[4] <= 24 * 1024
    blk = arrs["blk"].reshape(-1, 8)
    assert blk[:, 2].sum() == len(pts)                         # every node owned exactly once
    owned = np.concatenate([arrs["halo"][b[0]:b[0] + b[2]] for b in blk])
    assert np.array_equal(np.sort(owned), np.arange(len(pts)))


@pytest.mark.parametrize("n,rows", [(3, 80), (8, 64), (8, 80), (8, 256)])
def test_accumulator_free_plan_covers_the_tangent_action(n, rows):
    """The matrix-free tangent action (k_nh_apply) carves, per block, 16-byte
    aligned regions of 24·nh (coordinates), 24·nh (displacements), 24·nown
    (row sums), 4·nown (row meta) and then 24·nh (multiplier) bytes; the
    accumulator-free plan it is launched with must cover that in every block,
    and own every node exactly once."""
    pts, conn, _ = kuhn_cube_arrays(n)
    topo = P1Topology(pts, conn, len(pts), "tetra")
    budget = {64: 20 * 1024, 80: 24 * 1024}.get(rows, 64 * 1024 - 64)
    arrs, info = topo.host_plan("vector3a", max_rows=rows, lds_budget=budget)
    blk = arrs["blk"].reshape(-1, 8).astype(np.int64)
    nh, nown = blk[:, 1], blk[:, 2]
    a16 = lambda x: (x + 15) // 16 * 16
    need = a16(a16(24 * nh) + a16(24 * nh) + a16(24 * nown) + 4 * nown) + 24 * nh
    assert (need <= info[4]).all() and (blk[:, 7] >= nown).all()
    assert nown.sum() == len(pts) and nown.max() <= rows


def test_plan_limit_error_is_reported():
    pts, conn, _ = kuhn_cube_arrays(2)
    topo = P1Topology(pts, conn, len(pts), "tetra")
    with pytest.raises(_lib.TMError, match="single row"):
        topo.host_plan("scalar", max_rows=8, lds_budget=256)


def test_assembly_rejects_label_dependent_quadrature():
    """Plan records may list an element's vertices in any order, so the entry
    points refuse moments that depend on the labels (checked on the host,
    before any device work)."""
    topo = P1Topology(*kuhn_cube_arrays(2)[:2], 27, "tetra")
    L = _lib.lib()
    plan = _lib.PlanDev()
    fake = ctypes.c_void_p(64)
    good = topo.quad(2)
    bad = topo.quad(2)
    bad.W[0] *= 1.5
    for q in (bad,):
        rc = L.tm_asm_p1_matrix(_lib.FORM_KAPPA_LAPLACE, 4, 3, ctypes.byref(plan), fake, fake, _lib.F32, 1.0,
                                ctypes.byref(q), fake, _lib.F32, None)
        assert rc == ERR_ARG and b"relabelling" in L.tm_last_error()
        rc = L.tm_asm_p1_load(_lib.LOAD_CONST, 4, 3, ctypes.byref(plan), fake, None, _lib.F64, 1.0,
                              ctypes.byref(q), fake, None)
        assert rc == ERR_ARG
    # symmetric moments pass the check (an empty plan launches nothing)
    rc = L.tm_asm_p1_matrix(_lib.FORM_KAPPA_LAPLACE, 4, 3, ctypes.byref(plan), fake, fake, _lib.F32, 1.0,
                            ctypes.byref(good), fake, _lib.F32, None)
    assert rc == 0


def test_tangent_entry_points_refuse_accumulator_free_plans():
    """A 3x3-block plan states its accumulator bytes per CSR entry: the tangent
    and stiffness calls refuse one without room for 72-byte accumulators (it
    would overrun the blocks' LDS), while residual-only calls accept it; the
    checks run on the host, and an empty plan launches nothing."""
    L = _lib.lib()
    fake = ctypes.c_void_p(64)
    plan = _lib.PlanDev()
    for eb, ok in ((0, True), (72, True), (36, False), (-1, False), (0, True)):
        plan.entry_bytes = eb
        rc = L.tm_asm_nh(ctypes.byref(plan), fake, fake, None, _lib.F64, 1.0, 0.3, fake, _lib.F32, fake, None, None)
        assert (rc == 0) == ok and (ok or b"72-byte" in L.tm_last_error())
        rc = L.tm_asm_linear_elastic(ctypes.byref(plan), fake, None, _lib.F64, 1.0, 0.3, fake, _lib.F32, None)
        assert (rc == 0) == ok
        # residual without the tangent: any plan
        assert L.tm_asm_nh(ctypes.byref(plan), fake, fake, None, _lib.F64, 1.0, 0.3, None, 0, fake, None, None) == 0
    from tensormesh_amd.assemble.topology import PLAN_KINDS
    assert PLAN_KINDS["vector3"][0] == 72 and PLAN_KINDS["vector3a"][0] == 0   # stated as 72 and -1


@pytest.mark.parametrize("name,arrays,et", list(_meshes()))
def test_deterministic_mode_is_order_independent(name, arrays, et):
    """The deterministic mode's arithmetic on the CPU: with entries rounded to
    the row's power-of-two grid, every order of the additions gives the same
    bits, and the values stay within a few 2^-44 of the row's largest entry of
    the float64 oracle; without the rounding the orders disagree in the last
    bits."""
    pts, conn, _ = arrays
    topo = P1Topology(pts, conn, len(pts), et)
    arrs, info = topo.host_plan("scalar", max_rows=33)
    kappa = np.exp(2.0 * np.random.default_rng(4).standard_normal(len(pts)))
    f = _kernel_element_matrix(topo, "kappa_laplace", kappa)
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    ref = oa.reduce(pe, oa.element_matrices(pts, conn, et, "kappa_laplace", kappa=kappa), len(u))
    runs = [emulate_matrix(topo, arrs, f, det=True, shuffle=np.random.default_rng(s)) for s in range(3)]
    assert all(np.array_equal(runs[0], r) for r in runs[1:])
    scale = np.zeros(len(pts))
    np.maximum.at(scale, u, np.abs(ref))
    assert (np.abs(runs[0] - ref) / scale[u]).max() <= 64 * 2.0 ** -44
    plain = [emulate_matrix(topo, arrs, f, shuffle=np.random.default_rng(s)) for s in range(3)]
    assert not all(np.array_equal(plain[0], r) for r in plain[1:])

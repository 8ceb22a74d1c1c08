"""The oracle and the package tables against the fixtures in
tests/golden/known_answers.json (the reference's own test fixtures and their
closed-form answers; written by tests/golden/make_known_answers.py)."""
import json
import os

import numpy as np

from oracle import assemble as oa
from oracle import condense as oc
from oracle import pattern as op
from oracle import solve as osv
from tensormesh_amd import element as tel

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")))


def test_unit_triangle_fixture():
    g = GOLD["2d_pure_tri"]
    pts, conn = np.array(g["points"]), np.array(g["triangles"])
    assert np.allclose(oa.element_matrices(pts, conn, "triangle", "laplace")[0], g["laplace"], atol=1e-15)
    assert np.allclose(oa.element_matrices(pts, conn, "triangle", "mass")[0], g["mass"], atol=1e-15)


def test_kuhn_tet_fixture():
    g = GOLD["kuhn_tet"]
    K = oa.element_matrices(np.array(g["points"], float), np.array(g["tetra"]), "tetra", "laplace")[0]
    assert np.allclose(K, g["laplace"], atol=1e-15)


def test_tet_rule_fixture():
    g = GOLD["tet_order2_rule"]
    w, q = tel.quadrature("tetra", 2)
    assert np.array_equal(w, np.array(g["weights"])) and np.array_equal(q, np.array(g["points"]))


def test_kuhn_nnz_fixture():
    for n, nnz in GOLD["kuhn_nnz"].items():
        assert op.kuhn_nnz(int(n)) == nnz


def test_condenser_fixture():
    g = GOLD["condenser_simple_system"]
    row, col, val = np.array(g["row"]), np.array(g["col"]), np.array(g["val"])
    mask, f = np.array(g["mask"]), np.array(g["f"])
    kv, kr, kc, ni, fi, _ = oc.condense(val, row, col, mask, f)
    assert [ni, ni] == g["inner_shape"]
    assert list(kr) == g["inner_row"] and list(kc) == g["inner_col"] and np.allclose(kv, g["inner_val"])
    assert np.allclose(fi, g["f_inner_zero_dirichlet"])
    kv, kr, kc, ni, fi, _ = oc.condense(val, row, col, mask, f, np.array(g["dirichlet_update"]))
    assert np.allclose(fi, g["f_inner_dirichlet_update"])
    full = oc.recover(osv.direct(osv.csr(kv, kr, kc, ni), fi), mask, np.array([0.0, 0.0, 0.0, 0.0]))
    assert np.allclose(full[mask], 0.0)


def test_block_coo_fixtures():
    """SparseMatrix.from_block_coo (sparse/matrix.py:325-375) against the
    reference's block-COO fixtures: the two-block case and dense placement."""
    import torch
    import tensormesh_amd as tm
    g = GOLD["block_coo_naive"]
    m = tm.SparseMatrix.from_block_coo(torch.tensor(g["edata"]), torch.tensor(g["row"]), torch.tensor(g["col"]),
                                       tuple(g["shape"]))
    assert torch.equal(m.to_dense(), torch.tensor(g["dense"], dtype=torch.float32))
    for c in GOLD["block_coo_random"]["cases"]:
        m = tm.SparseMatrix.from_block_coo(torch.tensor(c["edata"], dtype=torch.float64), torch.tensor(c["row"]),
                                           torch.tensor(c["col"]), tuple(c["shape"]))
        assert m.shape == (c["shape"][0] * c["h"], c["shape"][1] * c["h"])
        assert torch.allclose(m.to_dense(), torch.tensor(c["dense"], dtype=torch.float64), atol=1e-12, rtol=0)

"""The product path (HIP kernels) on the fixtures of tests/golden/known_answers.json."""
import json
import os

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from tensormesh_amd.mesh import Mesh
from tensormesh_amd.sparse.matrix import SparseMatrix

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")))


def _dense(K):
    A = np.zeros(K.shape)
    A[K.row.cpu().numpy(), K.col.cpu().numpy()] = K.values.double().cpu().numpy()
    return A


@pytest.mark.parametrize("key,cell", [("2d_pure_tri", "triangle"), ("kuhn_tet", "tetra")])
def test_single_element_fixture(cuda, key, cell):
    print(f"case golden {key}", flush=True)
    g = GOLD[key]
    pts = torch.tensor(g["points"], dtype=torch.float64)
    conn = torch.tensor(g["triangles" if cell == "triangle" else "tetra"], dtype=torch.int64)
    m = Mesh(pts, {cell: conn}, {})
    K = tm.LaplaceElementAssembler.from_mesh(m).to(cuda)()
    assert np.abs(_dense(K) - np.array(g["laplace"])).max() <= 1e-6 * np.abs(g["laplace"]).max()
    if "mass" in g:
        M = tm.MassElementAssembler.from_mesh(m).to(cuda)()
        assert np.abs(_dense(M) - np.array(g["mass"])).max() <= 1e-6 * np.abs(g["mass"]).max()


def test_condenser_fixture(cuda):
    print("case golden condenser", flush=True)
    g = GOLD["condenser_simple_system"]
    K = SparseMatrix(torch.tensor(g["val"], dtype=torch.float64, device=cuda), torch.tensor(g["row"], device=cuda),
                     torch.tensor(g["col"], device=cuda), (4, 4))
    f = torch.tensor(g["f"], dtype=torch.float64, device=cuda)
    mask = torch.tensor(g["mask"], device=cuda)
    c = tm.Condenser(mask)
    Ki, fi = c(K, f)
    assert list(Ki.shape) == g["inner_shape"]
    assert Ki.row.tolist() == g["inner_row"] and Ki.col.tolist() == g["inner_col"]
    assert np.allclose(Ki.values.cpu().numpy(), g["inner_val"]) and np.allclose(fi.cpu().numpy(), g["f_inner_zero_dirichlet"])
    u = c.recover(Ki.solve(fi))
    assert u.shape == (4,) and torch.allclose(u[mask], torch.zeros(2, dtype=torch.float64, device=cuda), atol=1e-10)
    c.update_dirichlet(torch.tensor(g["dirichlet_update"], device=cuda))
    assert np.allclose(c.dirichlet_value.cpu().numpy(), g["dirichlet_update"])
    assert np.allclose(c.condense_rhs(f).cpu().numpy(), g["f_inner_dirichlet_update"], atol=1e-14)

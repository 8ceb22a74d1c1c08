"""Mesh files (``Mesh.save`` / ``Mesh.read``, ``tensormesh/mesh/mesh.py:267-325, 694-740``):
gmsh ASCII 2.2 / 4.1 and ASCII ``.vtu`` for P1 triangles and tetrahedra."""
import os

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from tensormesh_amd import mesh_io

MSH22 = """$MeshFormat
2.2 0 8
$EndMeshFormat
$PhysicalNames
1
2 7 "domain"
$EndPhysicalNames
$Nodes
5
3 0 0 0
7 1 0 0
8 1 1 0
20 0 1 0
21 0.5 0.5 0
$EndNodes
$Elements
6
1 15 2 0 1 3
2 1 2 0 1 3 7
3 2 2 7 1 3 7 21
4 2 2 7 1 7 8 21
5 2 2 7 1 8 20 21
6 2 2 7 1 20 3 21
$EndElements
$NodeData
1
"f"
1
0.0
3
0
1
5
3 1.5
7 2.5
8 3.5
20 4.5
21 5.5
$EndNodeData
"""

MSH41 = """$MeshFormat
4.1 0 8
$EndMeshFormat
$Entities
1 0 1 0
10 0 0 0 0
1 0 0 0 1 1 0 0 0
$EndEntities
$Nodes
2 4 10 13
0 10 0 1
10
0 0 0
2 1 0 3
11
12
13
1 0 0
1 1 0
0 1 0
$EndNodes
$Elements
2 3 1 3
1 1 1 1
1 10 11
2 1 2 2
2 10 11 12
3 10 12 13
$EndElements
"""


def _write(tmp_path, name, text):
    p = os.path.join(tmp_path, name)
    with open(p, "w") as f:
        f.write(text)
    return p


def test_gmsh22_known_answer(tmp_path):
    m = tm.Mesh.read(_write(tmp_path, "a.msh", MSH22))
    assert m.dim == 2 and list(m.cells.keys()) == ["triangle"]          # the line and point cells are skipped
    assert torch.equal(m.points, torch.tensor([[0., 0], [1, 0], [1, 1], [0, 1], [0.5, 0.5]], dtype=torch.float64))
    assert torch.equal(m.cells["triangle"], torch.tensor([[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]]))
    assert torch.equal(m.point_data["f"], torch.tensor([1.5, 2.5, 3.5, 4.5, 5.5], dtype=torch.float64))


def test_gmsh41_known_answer(tmp_path):
    m = tm.Mesh.read(_write(tmp_path, "b.msh", MSH41))
    assert torch.equal(m.points, torch.tensor([[0., 0], [1, 0], [1, 1], [0, 1]], dtype=torch.float64))
    assert torch.equal(m.cells["triangle"], torch.tensor([[0, 1, 2], [0, 2, 3]]))


@pytest.mark.parametrize("ext", ["vtu", "msh"])
@pytest.mark.parametrize("kind", ["tri", "tet"])
def test_round_trip_is_exact(tmp_path, ext, kind):
    m = tm.structured_rectangle(5, jitter=0.2, seed=2) if kind == "tri" else tm.kuhn_cube(3)
    rng = np.random.default_rng(0)
    m.register_point_data("kappa", torch.tensor(rng.uniform(0.5, 2.0, m.n_points)))
    p = os.path.join(tmp_path, f"m.{ext}")
    assert m.save(p) is m
    r = tm.Mesh.read(p)
    assert torch.equal(r.points, m.points)
    et = m.default_element_type
    assert torch.equal(r.cells[et], m.cells[et])
    assert r.point_data["is_boundary"].dtype == torch.bool
    assert torch.equal(r.point_data["is_boundary"], m.point_data["is_boundary"])
    assert torch.equal(r.point_data["kappa"], m.point_data["kappa"])
    if ext == "vtu":
        assert torch.equal(r.point_data["u"], torch.zeros(m.n_points, dtype=torch.float64))


def test_read_mesh_assembles_identically(tmp_path):
    m = tm.kuhn_cube(3)
    p = os.path.join(tmp_path, "c.vtu")
    m.save(p)
    r = tm.Mesh.read(p)
    K0 = tm.LaplaceElementAssembler.from_mesh(m)(m.points)
    K1 = tm.LaplaceElementAssembler.from_mesh(r)(r.points)
    assert torch.equal(K0.row, K1.row) and torch.equal(K0.col, K1.col)
    assert torch.equal(K0.values, K1.values)


def test_errors(tmp_path):
    with pytest.raises(ValueError):
        tm.Mesh.read(_write(tmp_path, "x.obj", "v 0 0 0\n"))
    with pytest.raises(NotImplementedError):
        tm.Mesh.read(_write(tmp_path, "bin.msh", "$MeshFormat\n2.2 1 8\n$EndMeshFormat\n"))
    with pytest.raises(ValueError):
        mesh_io.read(_write(tmp_path, "empty.msh", "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n"
                                                   "$EndNodes\n$Elements\n1\n1 15 0 1\n$EndElements\n"))

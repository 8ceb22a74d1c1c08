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
    with pytest.raises(ValueError):       # binary file without its endianness marker
        tm.Mesh.read(_write(tmp_path, "bin.msh", "$MeshFormat\n2.2 1 8\n$EndMeshFormat\n"))
    with pytest.raises(NotImplementedError):
        tm.Mesh.read(_write(tmp_path, "v3.msh", "$MeshFormat\n3.0 0 8\n$EndMeshFormat\n"))
    with pytest.raises(ValueError):
        mesh_io.read(_write(tmp_path, "empty.msh", "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n"
                                                   "$EndNodes\n$Elements\n1\n1 15 0 1\n$EndElements\n"))


# ---- binary files, written byte by byte here (the gmsh and VTK file-format references) ----
import base64
import struct
import zlib

_PTS = np.array([[0.0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
_TRI = np.array([[0, 1, 2], [0, 2, 3]])
_F = np.array([1.5, 2.5, 3.5, 4.5])


def _msh41_binary():
    b = bytearray(b"$MeshFormat\n4.1 1 8\n") + struct.pack("<i", 1) + b"\n$EndMeshFormat\n"
    b += b"$PhysicalNames\n1\n2 7 \"domain\"\n$EndPhysicalNames\n"
    # entities: 1 point, 0 curves, 1 surface, 0 volumes
    b += b"$Entities\n" + struct.pack("<4Q", 1, 0, 1, 0)
    b += struct.pack("<i3dQ", 10, 0.0, 0.0, 0.0, 0)
    b += struct.pack("<i6dQiQ", 1, 0, 0, 0, 1, 1, 0, 1, 7, 0)
    b += b"\n$EndEntities\n"
    # nodes: tags 10..13 in two blocks (point entity, surface entity)
    b += b"$Nodes\n" + struct.pack("<4Q", 2, 4, 10, 13)
    b += struct.pack("<3iQ", 0, 10, 0, 1) + struct.pack("<Q", 10) + struct.pack("<3d", *_PTS[0])
    b += struct.pack("<3iQ", 2, 1, 0, 3) + struct.pack("<3Q", 11, 12, 13) + _PTS[1:].astype("<f8").tobytes()
    b += b"\n$EndNodes\n"
    # elements: a point element (type 15) and two triangles (type 2)
    b += b"$Elements\n" + struct.pack("<4Q", 2, 3, 1, 3)
    b += struct.pack("<3iQ", 0, 10, 15, 1) + struct.pack("<2Q", 1, 10)
    b += struct.pack("<3iQ", 2, 1, 2, 2) + struct.pack("<8Q", 2, 10, 11, 12, 3, 10, 12, 13)
    b += b"\n$EndElements\n"
    b += b'$NodeData\n1\n"f"\n1\n0.0\n3\n0\n1\n4\n'
    b += b"".join(struct.pack("<id", 10 + k, _F[k]) for k in range(4)) + b"\n$EndNodeData\n"
    return bytes(b)


def _msh22_binary():
    b = bytearray(b"$MeshFormat\n2.2 1 8\n") + struct.pack("<i", 1) + b"\n$EndMeshFormat\n"
    b += b"$Nodes\n4\n" + b"".join(struct.pack("<i3d", 5 + k, *_PTS[k]) for k in range(4)) + b"\n$EndNodes\n"
    b += b"$Elements\n3\n"
    b += struct.pack("<3i", 1, 1, 2) + struct.pack("<5i", 1, 0, 0, 5, 6)           # one line, 2 tags
    b += struct.pack("<3i", 2, 2, 2) + struct.pack("<6i", 2, 0, 0, 5, 6, 7) + struct.pack("<6i", 3, 0, 0, 5, 7, 8)
    b += b"\n$EndElements\n"
    return bytes(b)


@pytest.mark.parametrize("make", [_msh41_binary, _msh22_binary])
def test_gmsh_binary_known_answer(tmp_path, make):
    p = os.path.join(tmp_path, "bin.msh")
    with open(p, "wb") as f:
        f.write(make())
    m = tm.Mesh.read(p)
    assert torch.equal(m.points, torch.tensor(_PTS[:, :2]))
    assert torch.equal(m.cells["triangle"], torch.tensor(_TRI))
    if make is _msh41_binary:
        assert torch.equal(m.point_data["f"], torch.tensor(_F))


def _vtu(pieces, header="UInt32", compressor=None, appended=None, order="LittleEndian"):
    attrs = f'type="UnstructuredGrid" version="1.0" byte_order="{order}" header_type="{header}"'
    if compressor:
        attrs += f' compressor="{compressor}"'
    body = ('<?xml version="1.0"?>\n<VTKFile ' + attrs + '>\n<UnstructuredGrid>\n'
            '<Piece NumberOfPoints="4" NumberOfCells="3">\n'
            f'<PointData>{pieces["f"]}</PointData>\n<Points>{pieces["pts"]}</Points>\n'
            f'<Cells>{pieces["conn"]}{pieces["offs"]}{pieces["types"]}</Cells>\n'
            '</Piece>\n</UnstructuredGrid>\n')
    out = body.encode()
    if appended is not None:
        out += appended
    return out + b"</VTKFile>\n"


def _arrays(order="<"):
    # a line cell (VTK type 3) first: it is skipped on reading
    conn = np.array([0, 1, 0, 1, 2, 0, 2, 3])
    return {"f": ("Float64", _F.astype(order + "f8"), 'Name="f"'),
            "pts": ("Float32", _PTS.astype(order + "f4"), 'NumberOfComponents="3"'),
            "conn": ("Int64", conn.astype(order + "i8"), 'Name="connectivity"'),
            "offs": ("Int32", np.array([2, 5, 8]).astype(order + "i4"), 'Name="offsets"'),
            "types": ("UInt8", np.array([3, 5, 5], np.uint8), 'Name="types"')}


def _encode(data: bytes, header: str, mode: str, order="<") -> bytes:
    hd = {"UInt32": "u4", "UInt64": "u8"}[header]
    h = lambda *v: np.array(v).astype(order + hd).tobytes()
    if mode == "joint":
        return base64.b64encode(h(len(data)) + data)
    if mode == "split":
        return base64.b64encode(h(len(data))) + base64.b64encode(data)
    if mode == "zlib":                                   # 5-byte blocks: several blocks per array
        blocks = [data[i:i + 5] for i in range(0, len(data), 5)]
        comp = [zlib.compress(bk) for bk in blocks]
        return base64.b64encode(h(len(blocks), 5, len(blocks[-1]), *[len(c) for c in comp])) + \
            base64.b64encode(b"".join(comp))
    if mode == "raw":
        return h(len(data)) + data
    if mode == "rawzlib":
        comp = [zlib.compress(data)]
        return h(1, len(data), len(data), len(comp[0])) + comp[0]
    raise ValueError(mode)


@pytest.mark.parametrize("header,mode", [("UInt32", "joint"), ("UInt64", "split"), ("UInt32", "zlib"),
                                         ("UInt64", "zlib")])
def test_vtu_inline_binary_known_answer(tmp_path, header, mode):
    pieces = {k: f'<DataArray type="{t}" {extra} format="binary">'
                 f'{_encode(a.tobytes(), header, mode).decode()}</DataArray>'
              for k, (t, a, extra) in _arrays().items()}
    p = os.path.join(tmp_path, "b.vtu")
    with open(p, "wb") as f:
        f.write(_vtu(pieces, header, "vtkZLibDataCompressor" if mode == "zlib" else None))
    _check_vtu(p)


@pytest.mark.parametrize("encoding,mode", [("raw", "raw"), ("raw", "rawzlib"), ("base64", "split"),
                                           ("base64", "zlib")])
def test_vtu_appended_known_answer(tmp_path, encoding, mode):
    blobs, pieces, off = [], {}, 0
    for k, (t, a, extra) in _arrays().items():
        enc = _encode(a.tobytes(), "UInt64", mode)
        pieces[k] = f'<DataArray type="{t}" {extra} format="appended" offset="{off}"/>'
        blobs.append(enc)
        off += len(enc)
    app = b'<AppendedData encoding="' + encoding.encode() + b'">\n_' + b"".join(blobs) + b"\n</AppendedData>\n"
    comp = "vtkZLibDataCompressor" if "zlib" in mode else None
    p = os.path.join(tmp_path, "a.vtu")
    with open(p, "wb") as f:
        f.write(_vtu(pieces, "UInt64", comp, app))
    _check_vtu(p)


def test_vtu_big_endian(tmp_path):
    pieces = {k: f'<DataArray type="{t}" {extra} format="binary">'
                 f'{_encode(a.tobytes(), "UInt32", "split", ">").decode()}</DataArray>'
              for k, (t, a, extra) in _arrays(">").items()}
    p = os.path.join(tmp_path, "be.vtu")
    with open(p, "wb") as f:
        f.write(_vtu(pieces, "UInt32", None, None, "BigEndian"))
    _check_vtu(p)


def _check_vtu(p):
    m = tm.Mesh.read(p)
    assert torch.equal(m.points, torch.tensor(_PTS[:, :2]))
    assert torch.equal(m.cells["triangle"], torch.tensor(_TRI))
    assert torch.equal(m.point_data["f"], torch.tensor(_F))

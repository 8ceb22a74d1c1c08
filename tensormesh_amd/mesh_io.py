"""Mesh files without meshio: gmsh ``.msh`` (ASCII 2.2 and 4.1) and VTK ``.vtu`` (ASCII XML).

The reference reads and writes meshes through ``meshio.read`` / ``meshio.write``
(``tensormesh/mesh/mesh.py:267-325, 694-740``); meshio is not available here,
so this module implements the two formats the path needs, for the element
types the kernels cover (P1 ``triangle`` and ``tetra``; gmsh types 2 and 4,
VTK types 5 and 10).  Lower-dimensional cells in a file (boundary lines,
points) are skipped.  P1 simplices have the same vertex order in gmsh, VTK
and TensorMesh, so no reordering is involved.

Conventions follow ``Mesh.save`` / ``Mesh.__init__`` of the reference:
points are truncated to the dimension of the highest-dimensional cells on
read (``mesh.py:139-144``) and padded to 3D for ``.vtu``; boolean point data
(``is_*`` / ``*_mask``) is written as float and read back as bool; ``.vtu``
output carries a zero ``u`` field when the mesh has none (``mesh.py:309-312``).
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET
from typing import Dict, Optional, Tuple

import numpy as np

GMSH_TYPES = {2: ("triangle", 3), 4: ("tetra", 4)}
GMSH_NODES_PER_TYPE = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 8: 3, 9: 6, 10: 9, 11: 10, 15: 1}
GMSH_CODE = {"triangle": 2, "tetra": 4}
VTK_TYPES = {5: ("triangle", 3), 10: ("tetra", 4)}
VTK_CODE = {"triangle": 5, "tetra": 10}
CELL_DIM = {"triangle": 2, "tetra": 3}

MeshArrays = Tuple[np.ndarray, Dict[str, np.ndarray], Dict[str, np.ndarray]]


def _format_of(path: str, file_format: Optional[str]) -> str:
    fmt = (file_format or os.path.splitext(path)[1].lstrip(".")).lower()
    if fmt in ("msh", "gmsh", "gmsh22"):
        return "msh"
    if fmt == "vtu":
        return "vtu"
    raise ValueError(f"unsupported mesh file format '{fmt}' (supported: msh, vtu)")


def _finish(points: np.ndarray, cells: Dict[str, list], point_data: Dict[str, np.ndarray]) -> MeshArrays:
    cells = {k: np.asarray(v, dtype=np.int64).reshape(-1, GMSH_TYPES[GMSH_CODE[k]][1]) for k, v in cells.items() if len(v)}
    if not cells:
        raise ValueError("the file holds no triangle or tetra cells")
    dim = max(CELL_DIM[k] for k in cells)
    return np.ascontiguousarray(points[:, :dim], dtype=np.float64), cells, point_data


# ---- gmsh ------------------------------------------------------------------------------------
def _sections(text: str) -> Dict[str, list]:
    out: Dict[str, list] = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("$") and not ln.startswith("$End"):
            name = ln[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != "$End" + name:
                j += 1
            if j == len(lines):
                raise ValueError(f"gmsh file: section ${name} is not closed")
            out.setdefault(name, []).append(lines[i + 1:j])
            i = j + 1
        else:
            i += 1
    return out


def read_msh(path: str) -> MeshArrays:
    with open(path, "r") as f:
        sec = _sections(f.read())
    if "MeshFormat" not in sec:
        raise ValueError(f"{path}: no $MeshFormat section")
    head = sec["MeshFormat"][0][0].split()
    version, ftype = head[0], int(head[1])
    if ftype != 0:
        raise NotImplementedError(f"{path}: binary gmsh files are not supported (ASCII only)")
    if version.startswith("2"):
        nodes = sec["Nodes"][0]
        n = int(nodes[0])
        ids = np.empty(n, np.int64)
        pts = np.empty((n, 3))
        for k in range(n):
            t = nodes[1 + k].split()
            ids[k] = int(t[0])
            pts[k] = [float(v) for v in t[1:4]]
        cells: Dict[str, list] = {}
        el = sec["Elements"][0]
        for k in range(int(el[0])):
            t = [int(v) for v in el[1 + k].split()]
            et, ntags = t[1], t[2]
            if et in GMSH_TYPES:
                name, nb = GMSH_TYPES[et]
                cells.setdefault(name, []).extend(t[3 + ntags:3 + ntags + nb])
    elif version.startswith("4"):
        nodes = sec["Nodes"][0]
        nblocks, n = (int(v) for v in nodes[0].split()[:2])
        ids = np.empty(n, np.int64)
        pts = np.empty((n, 3))
        row, filled = 1, 0
        for _ in range(nblocks):
            _, _, param, nin = (int(v) for v in nodes[row].split())
            if param:
                raise NotImplementedError(f"{path}: parametric node coordinates are not supported")
            tags = [int(nodes[row + 1 + k]) for k in range(nin)]
            for k in range(nin):
                ids[filled + k] = tags[k]
                pts[filled + k] = [float(v) for v in nodes[row + 1 + nin + k].split()[:3]]
            filled += nin
            row += 1 + 2 * nin
        cells = {}
        el = sec["Elements"][0]
        nblocks = int(el[0].split()[0])
        row = 1
        for _ in range(nblocks):
            _, _, et, nin = (int(v) for v in el[row].split())
            if et in GMSH_TYPES:
                name, nb = GMSH_TYPES[et]
                for k in range(nin):
                    t = [int(v) for v in el[row + 1 + k].split()]
                    cells.setdefault(name, []).extend(t[1:1 + nb])
            elif et not in GMSH_NODES_PER_TYPE:
                raise NotImplementedError(f"{path}: gmsh element type {et}")
            row += 1 + nin
    else:
        raise NotImplementedError(f"{path}: gmsh format version {version}")
    # node tags -> 0-based rows in file order
    lookup = np.full(int(ids.max()) + 1, -1, np.int64)
    lookup[ids] = np.arange(len(ids))
    for k in cells:
        c = lookup[np.asarray(cells[k], np.int64)]
        if (c < 0).any():
            raise ValueError(f"{path}: a {k} cell references an undefined node")
        cells[k] = c
    point_data: Dict[str, np.ndarray] = {}
    for block in sec.get("NodeData", []):
        ns = int(block[0])
        name = block[1].strip().strip('"')
        nr = int(block[1 + ns])
        ni = int(block[2 + ns + nr])
        ints = [int(v) for v in block[3 + ns + nr:3 + ns + nr + ni]]
        ncomp, count = ints[1], ints[2]
        start = 3 + ns + nr + ni
        vals = np.zeros((len(ids), ncomp))
        for k in range(count):
            t = block[start + k].split()
            vals[lookup[int(t[0])]] = [float(v) for v in t[1:1 + ncomp]]
        point_data[name] = vals[:, 0] if ncomp == 1 else vals
    return _finish(pts, cells, point_data)


def write_msh(path: str, points: np.ndarray, cells: Dict[str, np.ndarray],
              point_data: Dict[str, np.ndarray]) -> None:
    """gmsh ASCII 2.2: nodes (z = 0 for 2D), cells with two zero tags, point data as $NodeData."""
    p3 = np.zeros((points.shape[0], 3))
    p3[:, :points.shape[1]] = points
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(p3))
        for i, (x, y, z) in enumerate(p3):
            f.write(f"{i + 1} {float(x)!r} {float(y)!r} {float(z)!r}\n")
        f.write("$EndNodes\n$Elements\n")
        total = sum(len(c) for c in cells.values())
        f.write(f"{total}\n")
        eid = 1
        for name, conn in cells.items():
            code = GMSH_CODE[name]
            for row in np.asarray(conn):
                f.write(f"{eid} {code} 2 0 0 " + " ".join(str(int(v) + 1) for v in row) + "\n")
                eid += 1
        f.write("$EndElements\n")
        for name, v in point_data.items():
            v = np.asarray(v, dtype=np.float64)
            v2 = v.reshape(len(p3), -1)
            f.write(f'$NodeData\n1\n"{name}"\n1\n0.0\n3\n0\n{v2.shape[1]}\n{len(p3)}\n')
            for i in range(len(p3)):
                f.write(f"{i + 1} " + " ".join(repr(float(a)) for a in v2[i]) + "\n")
            f.write("$EndNodeData\n")


# ---- VTU -------------------------------------------------------------------------------------
def _ascii_array(node, dtype):
    if node.get("format", "ascii") != "ascii":
        raise NotImplementedError("only ASCII .vtu data arrays are supported")
    return np.array((node.text or "").split(), dtype=dtype)


def read_vtu(path: str) -> MeshArrays:
    root = ET.parse(path).getroot()
    if root.get("type") != "UnstructuredGrid":
        raise ValueError(f"{path}: not a VTK UnstructuredGrid file")
    piece = root.find("UnstructuredGrid/Piece")
    npts = int(piece.get("NumberOfPoints"))
    pts = _ascii_array(piece.find("Points/DataArray"), np.float64).reshape(npts, -1)
    arr = {a.get("Name"): a for a in piece.find("Cells").findall("DataArray")}
    conn = _ascii_array(arr["connectivity"], np.int64)
    offs = _ascii_array(arr["offsets"], np.int64)
    types = _ascii_array(arr["types"], np.int64)
    starts = np.concatenate([[0], offs[:-1]])
    cells: Dict[str, list] = {}
    for t, s, e in zip(types, starts, offs):
        if int(t) in VTK_TYPES:
            cells.setdefault(VTK_TYPES[int(t)][0], []).extend(conn[s:e].tolist())
    point_data: Dict[str, np.ndarray] = {}
    pd = piece.find("PointData")
    if pd is not None:
        for a in pd.findall("DataArray"):
            v = _ascii_array(a, np.float64)
            nc = int(a.get("NumberOfComponents", "1"))
            point_data[a.get("Name")] = v if nc == 1 else v.reshape(npts, nc)
    return _finish(pts, cells, point_data)


def write_vtu(path: str, points: np.ndarray, cells: Dict[str, np.ndarray],
              point_data: Dict[str, np.ndarray]) -> None:
    p3 = np.zeros((points.shape[0], 3))
    p3[:, :points.shape[1]] = points
    pd = {k: np.asarray(v, dtype=np.float64) for k, v in point_data.items()}
    if "u" not in pd:
        pd["u"] = np.zeros(len(p3))
    conn = np.concatenate([np.asarray(c, np.int64).reshape(-1) for c in cells.values()])
    shift = np.cumsum([0] + [np.asarray(c).size for c in cells.values()][:-1])
    offs = np.concatenate([np.arange(1, len(c) + 1) * np.asarray(c).shape[1] + s
                           for c, s in zip(cells.values(), shift)])
    types = np.concatenate([np.full(len(c), VTK_CODE[k]) for k, c in cells.items()])

    def arr(values, **attrs):
        a = " ".join(f'{k}="{v}"' for k, v in attrs.items())
        body = " ".join(repr(float(x)) if values.dtype.kind == "f" else str(int(x)) for x in values.reshape(-1))
        return f"<DataArray {a} format=\"ascii\">\n{body}\n</DataArray>\n"

    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n'
                f'<UnstructuredGrid>\n<Piece NumberOfPoints="{len(p3)}" NumberOfCells="{len(types)}">\n<PointData>\n')
        for k, v in pd.items():
            nc = 1 if v.ndim == 1 else v.shape[1]
            f.write(arr(v, type="Float64", Name=k, NumberOfComponents=nc))
        f.write("</PointData>\n<Points>\n")
        f.write(arr(p3, type="Float64", NumberOfComponents=3))
        f.write("</Points>\n<Cells>\n")
        f.write(arr(conn, type="Int64", Name="connectivity"))
        f.write(arr(offs, type="Int64", Name="offsets"))
        f.write(arr(types.astype(np.int64), type="UInt8", Name="types"))
        f.write("</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


def read(path: str, file_format: Optional[str] = None) -> MeshArrays:
    return read_msh(path) if _format_of(path, file_format) == "msh" else read_vtu(path)


def write(path: str, points: np.ndarray, cells: Dict[str, np.ndarray], point_data: Dict[str, np.ndarray],
          file_format: Optional[str] = None) -> None:
    pd = {k: (np.asarray(v).astype(np.float64)) for k, v in point_data.items()}
    if _format_of(path, file_format) == "msh":
        write_msh(path, points, cells, pd)
    else:
        write_vtu(path, points, cells, pd)

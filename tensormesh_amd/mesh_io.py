"""Mesh files without meshio: gmsh ``.msh`` (2.2 and 4.1, ASCII and binary) and VTK ``.vtu``
(XML with ASCII, base64 or appended raw / base64 data, optionally zlib-compressed).

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


class _Bytes:
    """Cursor over a binary gmsh file (little- or big-endian, fixed sizes)."""

    def __init__(self, data: bytes, path: str):
        self.d, self.i, self.path = data, 0, path
        self.end = "<"
        self.szt = "Q"

    def line(self) -> str:
        j = self.d.index(b"\n", self.i)
        out = self.d[self.i:j].decode("ascii", errors="replace").strip()
        self.i = j + 1
        return out

    def skip_ws(self):
        while self.i < len(self.d) and self.d[self.i:self.i + 1] in (b"\n", b"\r", b" "):
            self.i += 1

    def arr(self, code: str, n: int) -> np.ndarray:
        dt = np.dtype(self.end + code)
        nbytes = dt.itemsize * n
        if self.i + nbytes > len(self.d):
            raise ValueError(f"{self.path}: truncated binary data")
        out = np.frombuffer(self.d, dtype=dt, count=n, offset=self.i)
        self.i += nbytes
        return out

    def one(self, code: str) -> int:
        return int(self.arr(code, 1)[0])

    def expect_end(self, name: str):
        self.skip_ws()
        if self.line() != "$End" + name:
            raise ValueError(f"{self.path}: ${name} is not closed where expected")

    def skip_section(self, name: str):
        j = self.d.find(b"$End" + name.encode(), self.i)
        if j < 0:
            raise ValueError(f"{self.path}: section ${name} is not closed")
        self.i = j
        self.line()


def _read_msh_binary(path: str, data: bytes) -> MeshArrays:
    """gmsh binary 2.2 and 4.1 (the gmsh file-format reference, sections
    $Nodes, $Elements, $Entities, $NodeData)."""
    c = _Bytes(data, path)
    while c.line() != "$MeshFormat":
        pass
    head = c.line().split()
    version, dsize = head[0], int(head[2])
    if dsize not in (4, 8):
        raise NotImplementedError(f"{path}: data size {dsize}")
    c.szt = "Q" if dsize == 8 else "I"
    one = np.frombuffer(data, "<i4", 1, c.i)[0]
    if one != 1:
        c.end = ">"
        if np.frombuffer(data, ">i4", 1, c.i)[0] != 1:
            raise ValueError(f"{path}: bad endianness marker")
    c.i += 4
    c.expect_end("MeshFormat")
    v4 = version.startswith("4")
    if not (v4 or version.startswith("2")):
        raise NotImplementedError(f"{path}: gmsh format version {version}")
    ids = pts = None
    cells: Dict[str, list] = {}
    node_data = []
    while True:
        c.skip_ws()
        if c.i >= len(data):
            break
        name = c.line()
        if not name.startswith("$"):
            raise ValueError(f"{path}: unexpected content {name[:40]!r}")
        name = name[1:]
        if name == "Nodes" and v4:
            nblk, n, _, _ = (int(v) for v in c.arr(c.szt, 4))
            ids, pts, filled = np.empty(n, np.int64), np.empty((n, 3)), 0
            for _ in range(nblk):
                edim, _tag, param = (int(v) for v in c.arr("i", 3))
                nin = c.one(c.szt)
                if param:
                    raise NotImplementedError(f"{path}: parametric node coordinates are not supported")
                ids[filled:filled + nin] = c.arr(c.szt, nin)
                pts[filled:filled + nin] = c.arr("d", 3 * nin).reshape(nin, 3)
                filled += nin
            c.expect_end(name)
        elif name == "Nodes":
            n = int(c.line())
            rec = np.dtype([("tag", c.end + "i4"), ("x", c.end + "f8", 3)])
            r = np.frombuffer(data, rec, n, c.i)
            c.i += rec.itemsize * n
            ids, pts = r["tag"].astype(np.int64), r["x"].astype(np.float64)
            c.expect_end(name)
        elif name == "Elements" and v4:
            nblk, _, _, _ = (int(v) for v in c.arr(c.szt, 4))
            for _ in range(nblk):
                _edim, _tag, et = (int(v) for v in c.arr("i", 3))
                nin = c.one(c.szt)
                if et not in GMSH_NODES_PER_TYPE:
                    raise NotImplementedError(f"{path}: gmsh element type {et}")
                npe = GMSH_NODES_PER_TYPE[et]
                blk = c.arr(c.szt, nin * (1 + npe)).reshape(nin, 1 + npe)
                if et in GMSH_TYPES:
                    cells.setdefault(GMSH_TYPES[et][0], []).extend(blk[:, 1:].reshape(-1).tolist())
            c.expect_end(name)
        elif name == "Elements":
            total, done = int(c.line()), 0
            while done < total:
                et, nin, ntag = (int(v) for v in c.arr("i", 3))
                if et not in GMSH_NODES_PER_TYPE:
                    raise NotImplementedError(f"{path}: gmsh element type {et}")
                npe = GMSH_NODES_PER_TYPE[et]
                blk = c.arr("i", nin * (1 + ntag + npe)).reshape(nin, 1 + ntag + npe)
                if et in GMSH_TYPES:
                    cells.setdefault(GMSH_TYPES[et][0], []).extend(blk[:, 1 + ntag:].reshape(-1).tolist())
                done += nin
            c.expect_end(name)
        elif name == "Entities" and v4:
            npt, ncv, nsf, nvl = (int(v) for v in c.arr(c.szt, 4))
            for _ in range(npt):
                c.arr("i", 1), c.arr("d", 3)
                c.arr("i", c.one(c.szt))
            for _ in range(ncv + nsf + nvl):
                c.arr("i", 1), c.arr("d", 6)
                c.arr("i", c.one(c.szt))
                c.arr("i", c.one(c.szt))
            c.expect_end(name)
        elif name == "NodeData":
            hdr = []
            ns = int(c.line())
            hdr.append([c.line() for _ in range(ns)])
            nr = int(c.line())
            [c.line() for _ in range(nr)]
            ni = int(c.line())
            ints = [int(c.line()) for _ in range(ni)]
            ncomp, count = ints[1], ints[2]
            rec = np.dtype([("tag", c.end + "i4"), ("v", c.end + "f8", ncomp)])
            r = np.frombuffer(data, rec, count, c.i)
            c.i += rec.itemsize * count
            node_data.append((hdr[0][0].strip().strip('"') if hdr[0] else f"data{len(node_data)}", ncomp,
                              r["tag"].astype(np.int64), r["v"].reshape(count, ncomp).astype(np.float64)))
            c.expect_end(name)
        else:
            c.skip_section(name)
    if ids is None:
        raise ValueError(f"{path}: no $Nodes section")
    lookup = np.full(int(ids.max()) + 1, -1, np.int64)
    lookup[ids] = np.arange(len(ids))
    for k in cells:
        cc = lookup[np.asarray(cells[k], np.int64)]
        if (cc < 0).any():
            raise ValueError(f"{path}: a {k} cell references an undefined node")
        cells[k] = cc
    point_data: Dict[str, np.ndarray] = {}
    for nm, ncomp, tags, vals in node_data:
        full = np.zeros((len(ids), ncomp))
        full[lookup[tags]] = vals
        point_data[nm] = full[:, 0] if ncomp == 1 else full
    return _finish(pts, cells, point_data)


def read_msh(path: str) -> MeshArrays:
    with open(path, "rb") as f:
        raw = f.read()
    i = raw.find(b"$MeshFormat")
    if i < 0:
        raise ValueError(f"{path}: no $MeshFormat section")
    head = raw[i:i + 200].split(b"\n")[1].split()
    if len(head) >= 2 and head[1] == b"1":
        return _read_msh_binary(path, raw)
    sec = _sections(raw.decode("utf-8", errors="replace"))
    head = sec["MeshFormat"][0][0].split()
    version = head[0]
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
# VTK XML data arrays (the VTK file-format reference): "ascii", inline "binary"
# (base64 of a byte-count header and the data, encoded together or one after
# the other), and "appended" (raw bytes or base64 after the "_" of
# <AppendedData>, each array at its byte / character offset).  With
# compressor="vtkZLibDataCompressor" the header is [number of blocks, block
# size, last block size, compressed sizes...] and the data zlib blocks.
_VTK_DT = {"Int8": "i1", "UInt8": "u1", "Int16": "i2", "UInt16": "u2", "Int32": "i4", "UInt32": "u4",
           "Int64": "i8", "UInt64": "u8", "Float32": "f4", "Float64": "f8"}


def _b64_chars(nbytes: int) -> int:
    return 4 * ((nbytes + 2) // 3)


class _VtuArrays:
    def __init__(self, root, raw: Optional[bytes], text: Optional[str], path: str):
        self.path = path
        self.order = ">" if root.get("byte_order") == "BigEndian" else "<"
        self.hdt = np.dtype(self.order + _VTK_DT[root.get("header_type", "UInt32")])
        comp = root.get("compressor")
        if comp not in (None, "", "vtkZLibDataCompressor"):
            raise NotImplementedError(f"{path}: compressor {comp} is not supported (zlib only)")
        self.zlib = comp == "vtkZLibDataCompressor"
        self.raw, self.text = raw, text

    def _inflate(self, header: np.ndarray, body: bytes) -> bytes:
        import zlib
        nblk = int(header[0])
        sizes = [int(v) for v in header[3:3 + nblk]]
        out, o = [], 0
        for sz in sizes:
            out.append(zlib.decompress(body[o:o + sz]))
            o += sz
        return b"".join(out)

    def _from_raw(self, buf: bytes, off: int) -> bytes:
        hs = self.hdt.itemsize
        if not self.zlib:
            n = int(np.frombuffer(buf, self.hdt, 1, off)[0])
            return buf[off + hs:off + hs + n]
        nblk = int(np.frombuffer(buf, self.hdt, 1, off)[0])
        header = np.frombuffer(buf, self.hdt, 3 + nblk, off)
        body = off + hs * (3 + nblk)
        return self._inflate(header, buf[body:body + int(header[3:].sum())])

    def _from_b64(self, t: str) -> bytes:
        import base64
        hs = self.hdt.itemsize
        if not self.zlib:
            hc = _b64_chars(hs)
            n = int(np.frombuffer(base64.b64decode(t[:hc])[:hs], self.hdt)[0])
            if t[hc - 1] == "=":                       # header and data encoded one after the other
                return base64.b64decode(t[hc:hc + _b64_chars(n)])[:n]
            return base64.b64decode(t[:_b64_chars(hs + n)])[hs:hs + n]
        nblk = int(np.frombuffer(base64.b64decode(t[:_b64_chars(hs)])[:hs], self.hdt)[0])
        hc = _b64_chars(hs * (3 + nblk))
        header = np.frombuffer(base64.b64decode(t[:hc])[:hs * (3 + nblk)], self.hdt)
        nbody = int(header[3:].sum())
        return self._inflate(header, base64.b64decode(t[hc:hc + _b64_chars(nbody)]))

    def get(self, node, dtype) -> np.ndarray:
        dt = np.dtype(self.order + _VTK_DT[node.get("type", "Float64")])
        fmt = node.get("format", "ascii")
        if fmt == "ascii":
            return np.array((node.text or "").split(), dtype=dtype)
        if fmt == "binary":
            buf = self._from_b64("".join((node.text or "").split()))
        elif fmt == "appended":
            off = int(node.get("offset", "0"))
            if self.raw is not None:
                buf = self._from_raw(self.raw, off)
            elif self.text is not None:
                buf = self._from_b64(self.text[off:])
            else:
                raise ValueError(f"{self.path}: appended data array without <AppendedData>")
        else:
            raise NotImplementedError(f"{self.path}: data array format {fmt}")
        return np.frombuffer(buf, dt).astype(dtype)


def _parse_vtu(path: str):
    with open(path, "rb") as f:
        raw_file = f.read()
    raw = text = None
    i = raw_file.find(b"<AppendedData")
    if i >= 0:
        gt = raw_file.index(b">", i)
        enc = b'encoding="raw"' in raw_file[i:gt] or b"encoding='raw'" in raw_file[i:gt]
        us = raw_file.index(b"_", gt)
        j = raw_file.rfind(b"</AppendedData>")
        if enc:
            raw = raw_file[us + 1:j]
        else:
            text = "".join(raw_file[us + 1:j].decode("ascii").split())
        raw_file = raw_file[:gt + 1] + raw_file[j:]
    root = ET.fromstring(raw_file)
    return root, _VtuArrays(root, raw, text, path)


def read_vtu(path: str) -> MeshArrays:
    root, arrays = _parse_vtu(path)
    if root.get("type") != "UnstructuredGrid":
        raise ValueError(f"{path}: not a VTK UnstructuredGrid file")
    piece = root.find("UnstructuredGrid/Piece")
    npts = int(piece.get("NumberOfPoints"))
    pts = arrays.get(piece.find("Points/DataArray"), np.float64).reshape(npts, -1)
    arr = {a.get("Name"): a for a in piece.find("Cells").findall("DataArray")}
    conn = arrays.get(arr["connectivity"], np.int64)
    offs = arrays.get(arr["offsets"], np.int64)
    types = arrays.get(arr["types"], np.int64)
    starts = np.concatenate([[0], offs[:-1]])
    cells: Dict[str, list] = {}
    for t, s, e in zip(types, starts, offs):
        if int(t) in VTK_TYPES:
            cells.setdefault(VTK_TYPES[int(t)][0], []).extend(conn[s:e].tolist())
    point_data: Dict[str, np.ndarray] = {}
    pd = piece.find("PointData")
    if pd is not None:
        for a in pd.findall("DataArray"):
            v = arrays.get(a, np.float64)
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

"""Mesh container and the seeded structured generators used on the assembly path.

:class:`Mesh` keeps the reference's public surface (``tensormesh/mesh/mesh.py``:
``points`` buffer ``[N, d]``, ``cells`` keyed by element type with int64
connectivity ``[E, nb]`` (``:86-145``), ``point_data["is_boundary"]``,
``elements()`` (``:403-486``), ``boundary_mask`` (``:633-652``),
``default_element_type``, ``dtype``/``device``).  Mesh generation in the
reference goes through gmsh (``Mesh.gen_rectangle`` ``:742-793``,
``Mesh.gen_cube`` ``:1024-1074``), which is not available here; the
generators below produce structured / jittered simplex meshes instead:

* :func:`kuhn_cube` - unit cube, ``n^3`` cubes each split into the 6
  Kuhn (Freudenthal) tetrahedra that share the main diagonal;
* :func:`structured_rectangle` - unit square, ``m^2`` squares each split
  into 2 triangles, optional seeded jitter of the interior nodes;
* :func:`lattice_rectangle` - the ``gen_rectangle`` substitute: boundary
  nodes at spacing ``h``, interior nodes on a hexagonal lattice of spacing
  ``h``, Delaunay triangles (near-equilateral, the density of a frontal
  mesher such as gmsh's at ``chara_length = h``).

Node numbering is lexicographic with x fastest; elements are cube-major
(resp. square-major) in the same order.
"""
from __future__ import annotations

import math
from typing import Dict, Optional, Union, Iterable

import numpy as np
import torch
import torch.nn as nn

from .element import ELEMENT_DIM

# Kuhn permutations: tet = [v, v+e_a, v+e_a+e_b, v+(1,1,1)] for (a, b, c) a
# permutation of the axes.
_KUHN_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


class BufferDict(nn.Module):
    """String-keyed dict of tensors registered as module buffers (follows ``.to()``)."""

    def __init__(self, tensors: Optional[Dict[str, torch.Tensor]] = None):
        super().__init__()
        self._keys = []
        for k, v in (tensors or {}).items():
            self[k] = v

    def __setitem__(self, key, value):
        if key not in self._keys:
            self._keys.append(key)
        self.register_buffer("_b_" + key, value)

    def __getitem__(self, key):
        if key not in self._keys:
            raise KeyError(key)
        return getattr(self, "_b_" + key)

    def __contains__(self, key):
        return key in self._keys

    def __len__(self):
        return len(self._keys)

    def __iter__(self):
        return iter(self._keys)

    def keys(self):
        return list(self._keys)

    def values(self):
        return [self[k] for k in self._keys]

    def items(self):
        return [(k, self[k]) for k in self._keys]


class Mesh(nn.Module):
    """Simplex mesh: coordinates, connectivity per element type, per-point data."""

    def __init__(self, points: torch.Tensor, cells: Dict[str, torch.Tensor],
                 point_data: Optional[Dict[str, torch.Tensor]] = None):
        super().__init__()
        if not isinstance(points, torch.Tensor):
            points = torch.as_tensor(np.asarray(points))
        assert points.dim() == 2, f"points must be [n_points, dim], got {tuple(points.shape)}"
        self.register_buffer("points", points)
        self.cells = BufferDict({k: torch.as_tensor(v).long() for k, v in cells.items()})
        pd = {}
        for k, v in (point_data or {}).items():
            v = torch.as_tensor(v)
            if k.startswith("is_") or k.endswith("_mask"):
                v = v.bool()
            pd[k] = v
        self.point_data = BufferDict(pd)
        dims = {ELEMENT_DIM[k] for k in cells}
        top = max(dims)
        types = [k for k in cells if ELEMENT_DIM[k] == top]
        self.default_eletyp = types[0] if len(types) == 1 else types

    # ---- reference surface ------------------------------------------------
    def elements(self, element_type: Optional[Union[str, int, Iterable[str]]] = None):
        if element_type == "all":
            return {k: self.cells[k] for k in self.cells.keys()}
        if element_type is None:
            if isinstance(self.default_eletyp, str):
                return self.cells[self.default_eletyp]
            return {k: self.cells[k] for k in self.default_eletyp}
        if isinstance(element_type, int):
            d = self.dim if element_type == -1 else element_type
            return {k: self.cells[k] for k in self.cells.keys() if ELEMENT_DIM[k] == d}
        if isinstance(element_type, str):
            return self.cells[element_type]
        return {k: self.cells[k] for k in element_type}

    @property
    def boundary_mask(self) -> torch.Tensor:
        if "is_boundary" in self.point_data:
            return self.point_data["is_boundary"]
        if "boundary_mask" in self.point_data:
            return self.point_data["boundary_mask"]
        raise Exception("'boundary_mask' or 'is_boundary' is not found in point_data")

    @property
    def default_element_type(self):
        return self.default_eletyp

    @property
    def dtype(self) -> torch.dtype:
        return self.points.dtype

    @property
    def device(self) -> torch.device:
        return self.points.device

    @property
    def dim(self) -> int:
        return self.points.shape[1]

    @property
    def n_points(self) -> int:
        return self.points.shape[0]

    def n_elements(self) -> int:
        e = self.elements()
        if isinstance(e, torch.Tensor):
            return e.shape[0]
        return sum(v.shape[0] for v in e.values())

    def register_point_data(self, key: str, value: torch.Tensor):
        self.point_data[key] = value

    # ---- files (mesh.py:267-325, 694-740; gmsh .msh and .vtu without meshio) ----
    def save(self, file_name: str, file_format: Optional[str] = None) -> "Mesh":
        """Write ``.msh`` (gmsh ASCII 2.2) or ``.vtu`` (ASCII XML); returns ``self``."""
        from . import mesh_io
        mesh_io.write(file_name, self.points.detach().cpu().numpy(),
                      {k: v.detach().cpu().numpy() for k, v in self.cells.items()},
                      {k: v.detach().cpu().numpy() for k, v in self.point_data.items()}, file_format)
        return self

    to_file = save

    @classmethod
    def read(cls, file_name: str, file_format: Optional[str] = None, reorder: bool = False) -> "Mesh":
        """Read ``.msh`` (gmsh 2.2 / 4.1, ASCII or binary) or ``.vtu`` (XML: ASCII, base64 or
        appended raw / base64 arrays, optionally zlib-compressed) triangle / tetra meshes.
        ``reorder`` is accepted for the reference's signature; P1 simplices need none."""
        from . import mesh_io
        pts, cells, pd = mesh_io.read(file_name, file_format)
        return cls(torch.from_numpy(pts), {k: torch.from_numpy(v) for k, v in cells.items()},
                   {k: torch.from_numpy(v) for k, v in pd.items()})

    from_file = read

    # ---- generators (gmsh substitutes) --------------------------------------
    @staticmethod
    def gen_rectangle(chara_length: float = 0.1, order: int = 1, element_type: str = "tri",
                      left: float = 0.0, right: float = 1.0, bottom: float = 0.0, top: float = 1.0,
                      visualize: bool = False, cache_path: Optional[str] = None, dtype=torch.float64) -> "Mesh":
        """Triangles of edge length about ``chara_length`` on a rectangle
        (``mesh.py:743-793``; gmsh's frontal mesher there, :func:`lattice_rectangle`
        here).  P1 triangles only; ``visualize`` and ``cache_path`` are accepted
        for the reference's signature."""
        if order != 1 or element_type not in ("tri", "triangle"):
            raise NotImplementedError("gen_rectangle builds P1 triangles (order=1, element_type='tri')")
        pts, tris, bnd = lattice_rectangle_arrays(chara_length, left, right, bottom, top)
        return Mesh(torch.from_numpy(pts).to(dtype), {"triangle": torch.from_numpy(tris)},
                    {"is_boundary": torch.from_numpy(bnd)})

    @staticmethod
    def gen_cube(chara_length: float = 0.1, dtype=torch.float64) -> "Mesh":
        """Unit cube, Kuhn 6-tetrahedra split of a structured grid."""
        n = max(1, int(round(1.0 / chara_length)))
        return kuhn_cube(n, dtype=dtype)

    def __repr__(self):
        return (f"Mesh(n_points={self.n_points}, cells="
                f"{ {k: tuple(v.shape) for k, v in self.cells.items()} })")


def kuhn_cube_arrays(n: int):
    """Numpy arrays for the Kuhn grid: ``(points [N,3] f64, tets [6 n^3, 4] i64, boundary [N] bool)``."""
    n1 = n + 1
    g = np.arange(n1, dtype=np.float64) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    points = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    ii = np.arange(n1)
    kk, jj, iii = np.meshgrid(ii, ii, ii, indexing="ij")
    bnd = ((iii == 0) | (iii == n) | (jj == 0) | (jj == n) | (kk == 0) | (kk == n)).ravel()
    c = np.arange(n, dtype=np.int64)
    ck, cj, ci = np.meshgrid(c, c, c, indexing="ij")
    v0 = (ci + n1 * (cj + n1 * ck)).ravel()
    step = np.array([1, n1, n1 * n1], dtype=np.int64)
    tets = np.empty((v0.shape[0], 6, 4), dtype=np.int64)
    for p, (a, b, _) in enumerate(_KUHN_PERMS):
        tets[:, p, 0] = v0
        tets[:, p, 1] = v0 + step[a]
        tets[:, p, 2] = v0 + step[a] + step[b]
        tets[:, p, 3] = v0 + step.sum()
    return points, tets.reshape(-1, 4), bnd


def kuhn_cube(n: int, dtype=torch.float64) -> Mesh:
    pts, tets, bnd = kuhn_cube_arrays(n)
    return Mesh(torch.from_numpy(pts).to(dtype), {"tetra": torch.from_numpy(tets)},
                {"is_boundary": torch.from_numpy(bnd)})


def structured_rectangle_arrays(m: int, jitter: float = 0.0, seed: int = 0):
    """Numpy arrays for an ``m x m`` split-square triangulation of the unit square.

    Interior nodes are moved by a seeded uniform displacement inside the
    disk of radius ``jitter * h`` (``h = 1/m``), so jitter up to about 0.3
    keeps every triangle positively oriented.
    """
    m1 = m + 1
    g = np.arange(m1, dtype=np.float64) / m
    y, x = np.meshgrid(g, g, indexing="ij")
    points = np.stack([x.ravel(), y.ravel()], 1)
    jj, ii = np.meshgrid(np.arange(m1), np.arange(m1), indexing="ij")
    bnd = ((ii == 0) | (ii == m) | (jj == 0) | (jj == m)).ravel()
    if jitter > 0:
        rng = np.random.default_rng(seed)
        nint = int((~bnd).sum())
        r = jitter / m * np.sqrt(rng.uniform(0.0, 1.0, nint))
        t = rng.uniform(0.0, 2.0 * math.pi, nint)
        points[~bnd, 0] += r * np.cos(t)
        points[~bnd, 1] += r * np.sin(t)
    c = np.arange(m, dtype=np.int64)
    cj, ci = np.meshgrid(c, c, indexing="ij")
    v00 = (ci + m1 * cj).ravel()
    tris = np.empty((v00.shape[0], 2, 3), dtype=np.int64)
    tris[:, 0] = np.stack([v00, v00 + 1, v00 + m1 + 1], 1)
    tris[:, 1] = np.stack([v00, v00 + m1 + 1, v00 + m1], 1)
    return points, tris.reshape(-1, 3), bnd


def structured_rectangle(m: int, jitter: float = 0.0, seed: int = 0, dtype=torch.float64) -> Mesh:
    pts, tris, bnd = structured_rectangle_arrays(m, jitter, seed)
    return Mesh(torch.from_numpy(pts).to(dtype), {"triangle": torch.from_numpy(tris)},
                {"is_boundary": torch.from_numpy(bnd)})


def lattice_rectangle_arrays(h: float, left: float = 0.0, right: float = 1.0, bottom: float = 0.0,
                             top: float = 1.0):
    """Near-equilateral triangulation of a rectangle with edge length about ``h``.

    Boundary nodes divide each side into equal segments of length <= ``h``;
    interior nodes sit on a hexagonal lattice (rows ``h sqrt(3)/2`` apart,
    alternate rows shifted by ``h/2``) kept at least ``h/2`` from the sides;
    the triangles are the Delaunay triangulation (``scipy.spatial``),
    counter-clockwise.  At ``h = 0.05`` on the unit square: 509 nodes, 429 of
    them interior, 936 triangles."""
    import scipy.spatial
    W, H = right - left, top - bottom
    nx, ny = max(1, int(math.ceil(W / h - 1e-9))), max(1, int(math.ceil(H / h - 1e-9)))
    tx, ty = np.linspace(left, right, nx + 1), np.linspace(bottom, top, ny + 1)
    ring = [np.stack([tx, np.full_like(tx, bottom)], 1), np.stack([tx, np.full_like(tx, top)], 1),
            np.stack([np.full_like(ty[1:-1], left), ty[1:-1]], 1), np.stack([np.full_like(ty[1:-1], right), ty[1:-1]], 1)]
    bpts = np.concatenate(ring, 0)
    hx = W / nx
    dy = hx * math.sqrt(3.0) / 2.0
    rows = []
    y, j = bottom + dy, 0
    while y < top - 0.5 * hx:
        x0 = left + (0.5 if j % 2 else 1.0) * hx
        xs = np.arange(x0, right - 0.49 * hx, hx)
        rows.append(np.stack([xs, np.full_like(xs, y)], 1))
        y += dy
        j += 1
    ipts = np.concatenate(rows, 0) if rows else np.zeros((0, 2))
    pts = np.concatenate([bpts, ipts], 0)
    tri = scipy.spatial.Delaunay(pts).simplices.astype(np.int64)
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    det = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    tri = tri[np.abs(det) > 1e-12 * h * h]
    det = det[np.abs(det) > 1e-12 * h * h]
    tri[det < 0] = tri[det < 0][:, [0, 2, 1]]
    bnd = np.zeros(len(pts), dtype=bool)
    bnd[:len(bpts)] = True
    return pts, tri, bnd

"""Boundary-facet loads on the CPU: the oracle (``oracle/facet.py``) pinned
by closed forms on the unit square and cube, and the package's generic path
(CPU tensors) against it."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import facet as of
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays


def _square(m=6):
    pts, conn, bnd = structured_rectangle_arrays(m, jitter=0.25, seed=2)
    return "triangle", pts, conn, bnd


def _cube(n=3):
    pts, conn, bnd = kuhn_cube_arrays(n)
    rng = np.random.default_rng(3)
    pts = pts + (~bnd)[:, None] * rng.uniform(-0.1, 0.1, pts.shape) / n
    return "tetra", pts, conn, bnd


@pytest.mark.parametrize("case", [_square, _cube])
def test_oracle_closed_forms(case):
    """On the side x_{d-1} = 0 (unit measure): int c = c, int x_0 = 1/2,
    tractions integrate to t, a constant flux through the cell gradients to 0."""
    et, pts, conn, bnd = case()
    d = pts.shape[1]
    side = pts[:, d - 1] == 0.0
    b = of.load_vector(pts, conn, et, side, lambda v, gv, at: 2.5 * v)
    assert abs(b.sum() - 2.5) <= 1e-12
    assert np.all(b[~side] == 0.0)
    b = of.load_vector(pts, conn, et, side, lambda v, gv, at: at["x"][:, :1] * v)
    assert abs(b.sum() - 0.5) <= 1e-12
    t = np.array([1.0, -2.0, 0.5][:d])
    b = of.load_vector(pts, conn, et, side, lambda v, gv, at: v[..., None] * t, H=d).reshape(-1, d)
    assert np.abs(b.sum(0) - t).max() <= 1e-12
    beta = np.array([0.3, -0.7, 1.1][:d])
    b = of.load_vector(pts, conn, et, side, lambda v, gv, at: gv @ beta)
    assert abs(b.sum()) <= 1e-12 and np.abs(b).max() > 1e-3


def test_oracle_selection_rule():
    """The reference selects every facet whose nodes are all flagged: with the
    whole boundary flagged, an interior diagonal joining two boundary nodes
    (the corner squares) is integrated from both of its triangles."""
    et, pts, conn, bnd = _square(6)
    b = of.load_vector(pts, conn, et, bnd, lambda v, gv, at: v)
    diag = [(e, k) for e, k in of.select(conn, bnd)
            if len({round(pts[n, 0], 12) for n in np.delete(conn[e], k)}) == 2
            and len({round(pts[n, 1], 12) for n in np.delete(conn[e], k)}) == 2]
    assert len(of.select(conn, bnd)) == 24 + len(diag)
    assert abs(b.sum() - (4.0 + len(diag) * np.sqrt(2.0) / 6)) <= 1e-12


def test_oracle_square_nodal_values():
    """Constant c on the bottom side of an m x m square grid: each inner node
    of the side gets c h (two half edges), the two end nodes c h / 2."""
    et, pts, conn, bnd = _square(8)
    side = pts[:, 1] == 0.0
    b = of.load_vector(pts, conn, et, side, lambda v, gv, at: 3.0 * v)
    ends = side & ((pts[:, 0] == 0.0) | (pts[:, 0] == 1.0))
    assert np.abs(b[side & ~ends] - 3.0 / 8).max() <= 1e-13
    assert np.abs(b[ends] - 1.5 / 8).max() <= 1e-13


@pytest.mark.parametrize("case", [_square, _cube])
def test_generic_path_matches_oracle(case):
    et, pts, conn, bnd = case()
    d = pts.shape[1]
    k = np.random.default_rng(4).uniform(0.5, 1.5, len(pts))
    mesh = tm.Mesh(torch.from_numpy(pts), {et: torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})

    class Robin(tm.FacetAssembler):
        def forward(self, v, x, k, gradv):
            return torch.sin(x[0]) * k * v + 0.5 * (gradv @ x)

    b = Robin.from_mesh(mesh)(point_data={"k": torch.from_numpy(k)})
    ref = of.load_vector(pts, conn, et, bnd,
                         lambda v, gv, at: np.sin(at["x"][:, :1]) * at["k"][:, None] * v
                         + 0.5 * np.einsum("qad,qd->qa", gv, at["x"]), fields={"k": k})
    assert b.dtype == torch.float64 and np.abs(b.numpy() - ref).max() <= 1e-13 * np.abs(ref).max()

    class Traction(tm.FacetAssembler):
        def forward(self, v, x):
            t = torch.stack([x[0] ** 2, torch.ones_like(x[0])] + ([x[1]] if d == 3 else []))
            return v[:, None] * t[None, :]

    b = Traction.from_mesh(mesh)()
    ref = of.load_vector(pts, conn, et, bnd, lambda v, gv, at: v[..., None] * np.stack(
        [at["x"][:, 0] ** 2, np.ones(len(at["x"]))] + ([at["x"][:, 1]] if d == 3 else []), -1)[:, None, :], H=d)
    assert np.abs(b.numpy() - ref).max() <= 1e-13 * np.abs(ref).max()

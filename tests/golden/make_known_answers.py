"""Writes tests/golden/known_answers.json: the small fixtures the reference's
own tests hold for this path, and the closed-form answers they are checked
against (no reference code is imported or run).

* ``2d_pure_tri``: the single-triangle mesh of
  tests/assemble/test_mixed_element_type.py:126-138 (nodes 1-3 of the .inp,
  0-based here), with its P1 Laplace and mass matrices in closed form
  (K = |T| G G^T for the constant gradients G; M = |T| (1 + delta_ab) / 12).
* ``kuhn_tet``: the tetrahedron (0,0,0),(1,0,0),(1,1,0),(1,1,1) of the Kuhn
  grid and its Laplace matrix in closed form (gradients (-1,0,0), (1,-1,0),
  (0,1,-1), (0,0,1), volume 1/6).
* ``condenser_simple_system``: the 4x4 tridiagonal system of
  tests/operator/test_condenser.py:48-60 (row, col, val, f, mask), the
  expectations of its tests (:62-100: a 2x2 inner system, zero boundary
  values after recover with zero Dirichlet data, update_dirichlet to
  [1, 2]) and the condensed system by dense block elimination
  (K_ii = A[1:3, 1:3], f_i = f[1:3] - A[1:3, [0, 3]] u_o).
* ``tet_order2_rule``: the 4-point rule of tensormesh/element/quadrature.py:143,152.
* ``kuhn_nnz``: canonical-CSR sizes of the n^3 Kuhn grids, closed form
  N + 2 (3 n (n+1)^2 + 3 n^2 (n+1) + n^3).
* ``block_coo_naive``: the two 2x2 blocks of tests/sparse/test_block_coo.py:10-24
  and the 4x4 dense matrix they expand to.
* ``block_coo_random``: seeded instances of the dense-placement check of
  tests/sparse/test_block_coo.py:27-44 (block size, shape, block positions,
  block values and the dense matrix with every block added at
  ``[i*h:(i+1)*h, j*h:(j+1)*h]``).

    python tests/golden/make_known_answers.py
"""
import json
import os

import numpy as np


def main():
    out = {}
    pts = [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]
    out["2d_pure_tri"] = {
        "source": "tests/assemble/test_mixed_element_type.py:126-138",
        "points": pts, "triangles": [[0, 1, 2]],
        "laplace": [[1.0, -0.5, -0.5], [-0.5, 0.5, 0.0], [-0.5, 0.0, 0.5]],
        "mass": (np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]]) / 24.0).tolist(),
    }
    out["kuhn_tet"] = {
        "points": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]], "tetra": [[0, 1, 2, 3]],
        "laplace": (np.array([[1, -1, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -1, 1]]) / 6.0).tolist(),
    }
    row = [0, 0, 1, 1, 1, 2, 2, 2, 3, 3]
    col = [0, 1, 0, 1, 2, 1, 2, 3, 2, 3]
    val = [2., -1., -1., 2., -1., -1., 2., -1., -1., 2.]
    f = [1., 0., 0., 1.]
    mask = [True, False, False, True]
    A = np.zeros((4, 4))
    A[row, col] = val
    uo = np.array([1.0, 2.0])
    out["condenser_simple_system"] = {
        "source": "tests/operator/test_condenser.py:48-100",
        "row": row, "col": col, "val": val, "f": f, "mask": mask,
        "inner_shape": [2, 2],
        "dirichlet_update": [1.0, 2.0],
        "inner_row": [0, 0, 1, 1], "inner_col": [0, 1, 0, 1], "inner_val": A[1:3, 1:3][A[1:3, 1:3] != 0].tolist(),
        "f_inner_zero_dirichlet": np.array(f)[1:3].tolist(),
        "f_inner_dirichlet_update": (np.array(f)[1:3] - A[1:3][:, [0, 3]] @ uo).tolist(),
    }
    a, b = 0.5854101966249685, 0.1381966011250105
    out["tet_order2_rule"] = {
        "source": "tensormesh/element/quadrature.py:143,152",
        "points": [[a, b, b], [b, b, b], [b, b, a], [b, a, b]], "weights": [1.0 / 24.0] * 4,
    }
    out["kuhn_nnz"] = {str(n): int((n + 1) ** 3 + 2 * (3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3))
                       for n in (64, 96, 128)}
    out["block_coo_naive"] = {
        "source": "tests/sparse/test_block_coo.py:10-24",
        "edata": [np.ones((2, 2)).tolist(), (2 * np.ones((2, 2))).tolist()], "row": [0, 1], "col": [0, 1],
        "shape": [2, 2],
        "dense": [[1, 1, 0, 0], [1, 1, 0, 0], [0, 0, 2, 2], [0, 0, 2, 2]],
    }
    rng = np.random.default_rng(20240605)
    cases = []
    for h, r in ((1, 7), (3, 11), (5, 4)):
        nblk = int(rng.integers(1, r * r))
        row = rng.integers(0, r, nblk)
        col = rng.integers(0, r, nblk)               # repeated positions are added, as scipy's random allows
        edata = rng.uniform(-1.0, 1.0, (nblk, h, h)).round(6)
        dense = np.zeros((r * h, r * h))
        for i, j, v in zip(row, col, edata):
            dense[i * h:(i + 1) * h, j * h:(j + 1) * h] += v
        cases.append({"h": h, "shape": [r, r], "row": row.tolist(), "col": col.tolist(),
                      "edata": edata.tolist(), "dense": dense.tolist()})
    out["block_coo_random"] = {"source": "tests/sparse/test_block_coo.py:27-44", "cases": cases}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(path)


if __name__ == "__main__":
    main()

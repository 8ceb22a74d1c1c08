"""Seeded batch of uneven 2D meshes with log-normal coefficients.

Mesh k (seed ``seed + k``): ``E_k`` log-uniform in ``[e_min, e_max]`` elements,
realised as an ``m x m`` split-square triangulation with ``m = round(sqrt(E_k/2))``,
interior nodes jittered inside a disk of radius ``jitter * h``, and
``kappa = exp(g)`` with ``g`` a Gaussian random field of length scale
``U[l_min, l_max]`` and standard deviation ``sigma``, sampled on a lattice and
bilinearly interpolated to the nodes.
"""
from __future__ import annotations

import numpy as np

from ..mesh import structured_rectangle_arrays
from .grf import lognormal_kappa_points


def batched_square_meshes(n_mesh: int = 256, e_min: float = 2e3, e_max: float = 2e5, jitter: float = 0.2,
                          l_min: float = 0.05, l_max: float = 0.2, sigma: float = 1.0, seed: int = 0,
                          grid: int = 128):
    out = []
    for k in range(n_mesh):
        rng = np.random.default_rng(seed + k)
        E = float(np.exp(rng.uniform(np.log(e_min), np.log(e_max))))
        m = max(2, int(round(np.sqrt(E / 2.0))))
        pts, tris, bnd = structured_rectangle_arrays(m, jitter=jitter, seed=seed + k)
        ell = float(rng.uniform(l_min, l_max))
        kappa = lognormal_kappa_points(pts, ell, sigma, seed=seed + k, grid=grid)
        out.append((pts, tris, bnd, kappa))
    return out

"""Log-normal Poisson problems on the Kuhn cube for training and evaluating the
learned preconditioner: ``-div(kappa grad u) = 1`` on the unit cube, zero
Dirichlet data, P1 tetrahedra, ``kappa = exp(g)`` with ``g`` a seeded Gaussian
random field.  Assembly, condensation and the plain solver are the device
path of this package (κ-Laplace assembly kernel, Condenser gathers,
Jacobi-PCG)."""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np
import torch

from ..assemble import KappaLaplaceElementAssembler, const_node_assembler
from ..dataset.grf import lognormal_kappa_cube_device
from ..mesh import Mesh, kuhn_cube_arrays
from ..operator import Condenser
from .hierarchy import Hierarchy


def field_params(seed: int, length_range=(0.05, 0.2), sigma_range=(0.5, 2.3)) -> Tuple[float, float]:
    """Per-seed GRF length scale and standard deviation (uniform draws)."""
    g = np.random.default_rng(10_000_019 + int(seed))
    return float(g.uniform(*length_range)), float(g.uniform(*sigma_range))


class PoissonCube:
    def __init__(self, n: int, device: torch.device, n_levels: int = 4):
        self.n = int(n)
        self.device = torch.device(device)
        pts, conn, bnd = kuhn_cube_arrays(self.n)
        self.mesh = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)},
                         {"is_boundary": torch.from_numpy(bnd)})
        self.asm = KappaLaplaceElementAssembler.from_mesh(self.mesh).to(self.device)
        self.asm.value_dtype = torch.float32              # the systems the network was trained on
        self.load = const_node_assembler(1.0).from_mesh(self.mesh).to(self.device)()
        self.cond = Condenser(self.mesh.boundary_mask.to(self.device))
        self.inner = torch.from_numpy(~bnd).to(self.device)
        ones = torch.ones(len(pts), dtype=torch.float32, device=self.device)
        K = self.asm(point_data={"kappa": ones})
        Ki, fi = self.cond(K, self.load)
        self.csr = Ki.csr()
        self.b = fi.to(torch.float64).contiguous()
        self.n_inner = Ki.shape[0]
        rp = self.csr.row_ptr.cpu().numpy()
        ci = self.csr.col_idx.cpu().numpy()
        self.hier = Hierarchy.from_csr(pts[~bnd], rp, ci, n_levels, self.device,
                                       level0_csr=(self.csr.row_ptr, self.csr.col_idx))

    def kappa(self, seed: int) -> torch.Tensor:
        length, sigma = field_params(seed)
        return lognormal_kappa_cube_device(self.n, length, sigma, seed, self.device)

    def system(self, kappa: torch.Tensor):
        """(condensed FP32 values in canonical CSR order, FP64 right-hand side)."""
        K = self.asm(point_data={"kappa": kappa})
        Ki, fi = self.cond(K, self.load)
        return Ki.values, fi.to(torch.float64)

    def log_kappa_feature(self, kappa: torch.Tensor) -> torch.Tensor:
        """(log kappa - mean) / 2 at the interior nodes (the network's second input)."""
        lk = torch.log(kappa[self.inner].to(torch.float64))
        return ((lk - lk.mean()) / 2.0).to(torch.float32)

"""Isotropic linear elasticity: ``LinearElasticityElementAssembler``.

Drop-in for ``tensormesh.assemble.LinearElasticityElementAssembler``
(``tensormesh/assemble/builtin.py:103-179``): ``from_mesh(mesh, E=1.0, nu=0.3)``,
``__call__(points)`` returning the ``[N*d, N*d]`` stiffness as the reference's
block COO (``SparseMatrix.from_block_coo`` order), ``forward(gradu, gradv)`` =
``B_u^T C B_v`` with the Voigt strain-displacement matrix and Hooke's matrix of
``functional/elasticity.py:132-285``, and ``element_energy``.

Block ``(a, b)`` of an element (row = the ``u`` basis function ``a``, as the
reference's pair layout ``element_assembler.py:838-851`` puts the ``u`` index
on the row) is ``vol B_a^T C B_b``, in closed form::

    K_ab[i][k] = vol (mu (g_a . g_b) d_ik + lam g_a,i g_b,k + mu g_a,k g_b,i)

On GPU tensors with P1 tetrahedra and no ``func`` / extra data the matrix
comes from the HIP kernel ``tm_asm_linear_elastic`` (the neo-Hookean tangent
kernel's ``u = 0`` variant); anything else runs the generic path, which
evaluates ``forward`` exactly as the reference does.
"""
from __future__ import annotations

from typing import Optional, Union

import torch

from .. import _lib
from ..sparse.matrix import SparseMatrix
from .element_assembler import ElementAssembler
from .hyperelastic import block_matrix


# Block size of the stiffness plan: 22-row blocks in 32 KiB, five per CU (the 83-VGPR kernel
# allows five waves per SIMD), measured fastest for the u = 0 kernel at 64^3: 143.6 us against
# 150.4 us on the neo-Hookean tangent's plan (56-row leaves in 40 KiB, four blocks per CU; the
# tangent's 126 VGPRs allow only four), 153.9 / 156.2 / 164.2 us for 46 / 44 / 48-row caps in
# 32 KiB.  The builder halves a leaf that overflows the LDS budget, so a cap just above what
# fits gives blocks of half the size (on that mesh 40-row leaves in 40 KiB become 20-row
# blocks: 170.9 us).
LIN_PLAN_ROWS = 22
LIN_PLAN_LDS = 32 * 1024 - 64


def lame(E, nu):
    """``(mu, lam)`` of Young's modulus ``E`` and Poisson ratio ``nu``."""
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def voigt_shape_grad(gradu: torch.Tensor) -> torch.Tensor:
    """Strain-displacement matrix of one basis gradient ``g`` (``[d]``).

    Column ``j`` is the engineering strain of the displacement ``phi e_j``:
    2D rows ``(xx, yy, xy)`` -> ``[3, 2]``; 3D rows ``(xx, yy, zz, yz, xz, xy)``
    -> ``[6, 3]`` (``functional/elasticity.py:132-186``).
    """
    assert gradu.dim() == 1, f"gradu should be a 1D tensor of shape [dim], but got shape {gradu.shape}"
    d = gradu.shape[-1]
    assert d in (2, 3), f"dimension is only supported for 2, 3 for voigt shape grad, but got {d}"
    z = torch.zeros((), dtype=gradu.dtype, device=gradu.device)
    if d == 2:
        gx, gy = gradu[0], gradu[1]
        rows = [[gx, z], [z, gy], [gy, gx]]
    else:
        gx, gy, gz = gradu[0], gradu[1], gradu[2]
        rows = [[gx, z, z], [z, gy, z], [z, z, gz], [z, gz, gy], [gz, z, gx], [gy, gx, z]]
    return torch.stack([torch.stack(r) for r in rows])


def voigt_stiffness(E: Union[float, torch.Tensor], nu: Union[float, torch.Tensor], dim: int = 2) -> torch.Tensor:
    """Hooke's matrix in Voigt form, ``[3, 3]`` (2D) or ``[6, 6]`` (3D):
    ``lam + 2 mu`` on the normal diagonal, ``lam`` off it, ``mu`` on the shear
    diagonal (``functional/elasticity.py:188-255``)."""
    assert dim in (2, 3)
    mu, lam = lame(E, nu)
    mu = torch.as_tensor(mu, dtype=torch.float64)
    lam = torch.as_tensor(lam, dtype=torch.float64)
    nv = 3 if dim == 2 else 6
    C = torch.zeros((nv, nv), dtype=torch.float64)
    C[:dim, :dim] = lam
    i = torch.arange(nv)
    C[i, i] = C[i, i] + mu
    C[i[:dim], i[:dim]] = C[i[:dim], i[:dim]] + mu
    return C


class LinearElasticityElementAssembler(ElementAssembler):
    """``K = int sigma(u) : eps(v)`` with Hooke's law (``builtin.py:103-179``)."""

    def __post_init__(self, E=1.0, nu=0.3):
        self.E = E
        self.nu = nu

    def forward(self, gradu, gradv):
        dim = gradu.shape[0]
        Ba = voigt_shape_grad(gradu)
        Bb = voigt_shape_grad(gradv)
        C = voigt_stiffness(self.E, self.nu, dim).to(dtype=gradu.dtype, device=gradu.device)
        return Ba.T @ C @ Bb

    def element_energy(self, graddisplacement):
        """``psi = lam/2 (tr eps)^2 + mu eps : eps`` of a displacement gradient ``[d, d]``."""
        grad_u = graddisplacement
        eps = 0.5 * (grad_u + grad_u.transpose(-1, -2))
        mu, lam = lame(self.E, self.nu)
        tr = eps.diagonal(dim1=-2, dim2=-1).sum(-1)
        return 0.5 * lam * tr * tr + mu * (eps * eps).sum((-1, -2))

    def _kernel_ok(self, points, func, point_data, element_data, scalar_data) -> bool:
        return (points.is_cuda and func is None and not point_data and not element_data and not scalar_data
                and self.topology.nb == 4 and self.dimension == 3 and not points.requires_grad
                and not (torch.is_tensor(self.E) and self.E.requires_grad))

    def stiffness(self, points: Optional[torch.Tensor] = None, E=None,
                  value_dtype: Optional[torch.dtype] = None) -> SparseMatrix:
        """HIP-kernel stiffness (P1 tetrahedra on the GPU); ``E`` may be a scalar or
        a per-element tensor ``[n_elements]`` (default: the constructor's ``E``)."""
        points = self.points if points is None else points
        _lib.require_gpu(points, "points")
        assert self.topology.nb == 4 and self.dimension == 3, "the stiffness kernel covers P1 tetrahedra"
        dev = points.device
        if self.edges.device != dev:
            self.to(dev)
        E = self.E if E is None else E
        if torch.is_tensor(E) and E.dim() == 0:
            E = float(E)
        if torch.is_tensor(E):
            assert E.shape == (self.topology.n_elements,), \
                f"E must be a scalar or [{self.topology.n_elements}] per element, got {tuple(E.shape)}"
            Et, Edt, E0 = E.detach().to(dev).contiguous(), _lib.dtype_code(E), 0.0
        else:
            Et, Edt, E0 = None, _lib.F64, float(E)
        p64 = points.detach().to(dev, torch.float64).contiguous()
        kv = torch.empty(self.topology.nnz * 9, dtype=value_dtype or self._out_dtype(points), device=dev)
        plan = self.topology.device_plan(dev, "vector3", max_rows=LIN_PLAN_ROWS, lds_budget=LIN_PLAN_LDS)
        _lib.check(_lib.lib().tm_asm_linear_elastic(plan.struct, _lib.ptr(p64), _lib.ptr(Et), Edt, E0,
                                                     float(self.nu), _lib.ptr(kv), _lib.dtype_code(kv),
                                                     _lib.stream_of(dev)))
        return block_matrix(self, kv, dev)

    def __call__(self, points: Optional[torch.Tensor] = None, func=None, point_data=None, element_data=None,
                 scalar_data=None, batch_size: int = -1) -> SparseMatrix:
        pts = self.points if points is None else points
        if self._kernel_ok(pts, func, point_data, element_data, scalar_data):
            assert pts.shape[1] == self.dimension, \
                f"the dimension of points should be {self.dimension}, but got {pts.shape[1]}"
            return self.stiffness(pts)
        return super().__call__(points, func, point_data, element_data, scalar_data, batch_size)

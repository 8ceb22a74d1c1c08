"""Ready-made forms (``tensormesh/assemble/builtin.py``): Laplace (``:25-63``),
mass (``:65-101``), linear elasticity (``:103-179``, in ``elasticity.py``),
constant and function loads (``:886-969``)."""
from __future__ import annotations

from .element_assembler import ElementAssembler
from .node_assembler import NodeAssembler
from .elasticity import LinearElasticityElementAssembler  # noqa: F401  (reference module layout)


class LaplaceElementAssembler(ElementAssembler):
    """``K_ij = int grad phi_i . grad phi_j``."""

    def forward(self, gradu, gradv):
        return gradu @ gradv


class MassElementAssembler(ElementAssembler):
    """``M_ij = int phi_i phi_j``."""

    def forward(self, u, v):
        return u * v


class KappaLaplaceElementAssembler(ElementAssembler):
    """``K_ij = int kappa grad phi_i . grad phi_j`` with nodal ``point_data["kappa"]``
    (the coefficient-identification form, ``examples/inverse_design/coefficient_identification.py:50-55``)."""

    def forward(self, gradu, gradv, kappa):
        return kappa * (gradu @ gradv)


def const_node_assembler(c=1):
    """``f_i = int c phi_i``."""

    class ConstNodeAssembler(NodeAssembler):
        def __post_init__(self, c=c):
            self.c = c

        def forward(self, v):
            return self.c * v

    return ConstNodeAssembler


def func_node_assembler(f=lambda x: x):
    """``f_i = int f(x) phi_i`` (``builtin.py:929-969``; on GPU tensors the moments kernel with ``f`` evaluated
    at the quadrature points)."""

    class FuncNodeAssembler(NodeAssembler):
        def __post_init__(self, f=f):
            self.f = f

        def forward(self, x, v):
            return self.f(x) * v

    return FuncNodeAssembler


class SourceNodeAssembler(NodeAssembler):
    """``b_i = int f phi_i`` with nodal ``point_data["f"]`` (the quickstart load, ``README.md:104-106``)."""

    def forward(self, v, f):
        return f * v

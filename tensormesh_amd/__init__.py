"""MI355X-native P1 assembly, condensation and Krylov solves behind the TensorMesh API.

``Mesh -> ElementAssembler / NodeAssembler -> SparseMatrix -> Condenser -> solve``
with hand-written gfx950 HIP kernels reached through the C ABI of
``libtensormesh_amd.so`` (``include/tensormesh_amd.h``).
"""
from .mesh import Mesh, kuhn_cube, structured_rectangle
from .assemble import (ElementAssembler, NodeAssembler, FacetAssembler, LaplaceElementAssembler, MassElementAssembler,
                       KappaLaplaceElementAssembler, SourceNodeAssembler, const_node_assembler,
                       func_node_assembler, NeoHookeanModel, LinearElasticityElementAssembler)
from .sparse import SparseMatrix
from .operator import Condenser

__all__ = ["Mesh", "kuhn_cube", "structured_rectangle", "ElementAssembler", "NodeAssembler", "FacetAssembler",
           "LaplaceElementAssembler", "MassElementAssembler", "KappaLaplaceElementAssembler",
           "SourceNodeAssembler", "const_node_assembler", "func_node_assembler", "NeoHookeanModel",
           "LinearElasticityElementAssembler", "SparseMatrix", "Condenser"]

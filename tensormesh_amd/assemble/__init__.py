from .element_assembler import ElementAssembler
from .node_assembler import NodeAssembler
from .facet_assembler import FacetAssembler
from .builtin import (LaplaceElementAssembler, MassElementAssembler, KappaLaplaceElementAssembler,
                      SourceNodeAssembler, const_node_assembler, func_node_assembler)
from .topology import P1Topology, topology_for, clear_topology_cache
from .hyperelastic import NeoHookeanModel, BlockCondenser
from .elasticity import LinearElasticityElementAssembler
